/*
 * rh_hip.h — the thin C ABI between the C++ host layer and the gfx950 HIP
 * kernels (libray_hip.so).  Plain pointers and sizes only; all buffers use the
 * layouts of rh_types.h.  Every function returns 0 on success or a negative
 * status (the HIP error code, negated, or -1000-n for argument errors);
 * rh_last_error() gives the message.
 *
 * What each entry point replaces in the reference's GPU backend
 * (src/Ray/internal/RendererOCL.cpp, OpenCL):
 *   rh_create / rh_destroy     <- Ocl::Renderer::Renderer device/program setup   RendererOCL.cpp:57-349
 *   rh_resize                  <- Ocl::Renderer::Resize buffer allocation         RendererOCL.cpp:351-414
 *   rh_clear                   <- Ocl::Renderer::Clear                            RendererOCL.h:172-178
 *   rh_upload_scene            <- Ocl::Scene growable cl::Buffer uploads          SceneOCL.cpp:8-536, VectorOCL.h:71-76
 *   rh_set_halton              <- Halton upload inside RenderScene                RendererOCL.cpp:449-454
 *   rh_render_pass             <- Ocl::Renderer::RenderScene                      RendererOCL.cpp:429-640
 *   rh_read_pixels             <- framebuffer readback                            RendererOCL.cpp:634
 *   rh_get_stats/rh_reset_stats<- RendererBase::GetStats/ResetStats               RendererBase.h:93-102
 * Parity entry points, one stage on caller-supplied rays (host buffers); the
 * reference functions each one reproduces:
 *   rh_gen_primary_rays        <- Ref::GeneratePrimaryRays                        CoreRef.cpp:427-505
 *   rh_trace_rays              <- Ref::Traverse_MacroTree_WithStack_{ClosestHit,AnyHit}
 *                                                                                 CoreRef.cpp:1091-1131, 1224-1268
 *   rh_ray_hashes              <- get_ray_hash                                    CoreRef.cpp:316-333
 *   rh_sort_rays               <- Ref::SortRays_CPU                               CoreRef.cpp:615-657
 *   rh_shade_rays              <- Ref::ShadeSurface                               CoreRef.cpp:2055-2391
 */
#ifndef RH_HIP_H
#define RH_HIP_H

#include "rh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rh_context rh_context;

/* one RenderScene pass over a region (RendererRef.cpp:19-325) */
typedef struct rh_pass_desc {
    rh_camera cam;
    int32_t rect[4];      /* x, y, w, h; must lie inside the framebuffer */
    int32_t iteration;    /* region iteration after the increment (>= 1) */
    int32_t sort_rays;    /* 1 = sort secondary rays by hash before each bounce trace */
    float root_min[3];    /* macro-tree root box min (sort hash grid origin) */
    float cell_size[3];   /* (root max - root min) / 255 */
} rh_pass_desc;

const char *rh_last_error(void);

int rh_create(int device, rh_context **out);
void rh_destroy(rh_context *ctx);
int rh_resize(rh_context *ctx, int w, int h);
int rh_clear(rh_context *ctx, const float rgba[4]);
int rh_upload_scene(rh_context *ctx, const rh_scene_view *scene);
/* After instance transforms changed (SceneBase::SetMeshInstanceTransform,
 * SceneBase.h:265; the reference rebuilds its instance tree on the host and
 * re-sends the whole scene): replaces only the arrays that follow the
 * instances -- node arrays (with the rebuilt instance tree), meshes,
 * instances, transforms, tree roots, light leaves -- and keeps the
 * triangles, vertices, materials, textures and lights of the last
 * rh_upload_scene resident.  An error (nothing replaced, the context needs
 * rh_upload_scene) when any of those changed size. */
int rh_update_instances(rh_context *ctx, const rh_scene_view *scene);
int rh_set_halton(rh_context *ctx, const float *seq, size_t count);
int rh_render_pass(rh_context *ctx, const rh_pass_desc *pass);
int rh_read_pixels(rh_context *ctx, float *rgba, size_t count);
/* The accumulated spherical-harmonic data of passes with OutputSH
 * (RendererBase::get_sh_data_ref, RendererRef.h:64-66; FramebufferRef.cpp:
 * 57-94): w*h records of 12 floats (coeff_r[4], coeff_g[4], coeff_b[4]).
 * Returns 1 (and writes nothing) when no pass has produced SH data since the
 * last resize. */
int rh_read_sh(rh_context *ctx, float *out, size_t count);
int rh_synchronize(rh_context *ctx);
/* the stage timers of the passes rendered since the last reset (waits for the passes in flight) */
int rh_get_stats(rh_context *ctx, rh_stats *out);
int rh_reset_stats(rh_context *ctx);
int rh_get_ray_counts(rh_context *ctx, rh_ray_counts *out);
/* Knobs: "sort_rays" (0/1), "stage_timing" (0/1, default 1: the stage timers rh_get_stats reports, from HIP events
 *   recorded at each pass's stage boundaries and summed once the pass has finished; no pass waits for them),
 * "kernel_timing" (0/1: events around each kernel launch, see rh_get_kernel_profile),
 * "count_visits" (0/1: instrumented trace kernels count node/triangle/instance visits),
 * "lds_stack_depth" (4..64 LDS entries of the per-lane traversal stack),
 * "sort_bits" (8/16/24/32: the in-pass ray sort orders by the top bits of the key),
 * "path_tail" (-1: every bounce staged; b >= 0: bounces b.. run as one path-per-lane launch),
 * "tail_occupancy" (2 or 3 waves per SIMD for the path-tail kernel: registers vs spills),
 * "tail_refill" (0/1: the path tail as a persistent kernel whose lanes take the next path when theirs ends),
 * "tiled_primary" (1, the default: a pass queues its primary rays in 8x8 pixel tiles, one wave each; 0: in
 *   row-major order; images and ray counts are the same),
 * "refill" (0/1: closest-hit trace as a persistent kernel whose lanes take new rays as theirs finish;
 *   binary and 8-wide trees, the latter with the trace_occupancy register cap),
 * "trace_occupancy" (0, 5 or 6: the 8-wide trace and shadow kernels as compiled, or capped at 5 or 6 waves
 *   per SIMD),
 * "shadow_occupancy" (-1, the default: as trace_occupancy; 0, 5 or 6: the 8-wide shadow kernel's own setting),
 * "trace_group" (0, the default, or 2, 4 or 8: 8-wide closest-hit and any-hit traces of rh_trace, and the staged
 *   traces of bounce 2 and later, walk each ray with a group of that many lanes, each testing 8 / group of a
 *   node's slots; faster for the few rays of late bounces, slower for full queues; results are the same),
 * "max_bounces" (1..16: the clamp of max_total_depth; 10, the reference's MAX_BOUNCES, by default:
 *   larger values render deeper paths than the reference can, an extension),
 * "adjoint_staged" (1, the default: rh_adjoint_pass replays the pass as staged wavefront passes whose
 *   stages write the vertex records, then one kernel forms the gradient; 0: one lane per path),
 * "adjoint_occupancy" (2 or 3 waves per SIMD for the one-lane-per-path adjoint kernel),
 * "adjoint_tape" (0..4096: passes 1..N after a clear keep their vertex records on the device, every bounce
 *   staged, and rh_adjoint_pass of such a pass needs no replay; 0, the default: off),
 * "adjoint_tape_texels" (0/1: the tape records texel footprints too, for texel gradients),
 * "defer_nee" (0/1: shadow rays of next-event estimation as their own stage),
 * "pass_overlap" (0..3, default 3: passes in flight besides the newest; plain rh_render_pass passes
 *   (no tape, geometry camera or SH output) rotate over pass_overlap + 1 slots, each with its own stream
 *   and working set, and mix into the image after the pass before them, so the image is the serial one;
 *   any other entry point's work is waited for by the next pass). */
int rh_set_option(rh_context *ctx, const char *name, int value);
/* Any option above, and read-only values: "scene_uploads" (rh_upload_scene
 * calls), "instance_updates" (rh_update_instances calls), "bound_slot" (the
 * pass slot of the last pass), "stack_overflows" (traversal stack entries
 * dropped) and "sort_stalls" (ray-sort tile look-backs that gave up waiting;
 * 0 unless something is wrong). */
int rh_get_option(rh_context *ctx, const char *name, int *value);

/* Per-kernel profile collected while "kernel_timing" is on (HIP events around
 * every trace/shade/sort launch of rh_render_pass) and, while "count_visits"
 * is on, the traversal work of the trace kernels.  Kinds: 0 trace, 1 shade,
 * 2 sort, 3 shadow, 4 path tail (rh_kernel_profile in rh_types.h).  No counterpart in the reference
 * beyond stats_t. */
int rh_get_kernel_profile(rh_context *ctx, rh_kernel_profile *out);
int rh_reset_kernel_profile(rh_context *ctx);
/* hipDeviceSynchronize on the context's device */
int rh_device_synchronize(rh_context *ctx);

/* ---- parity entry points (host buffers in, host buffers out) ---- */
int rh_gen_primary_rays(rh_context *ctx, const rh_camera *cam, const int32_t rect[4], int w, int h, int iteration,
                        const float *halton, rh_ray_packet *out);
int rh_trace_rays(rh_context *ctx, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax, rh_hit *out);
int rh_ray_hashes(rh_context *ctx, const rh_ray_packet *rays, size_t n, const float root_min[3],
                  const float cell_size[3], uint32_t *out);
int rh_sort_rays(rh_context *ctx, rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3]);
/* Allocation-free form of rh_sort_rays on device buffers.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Sorts the n rays (d_rays, d_aux: the split layout of rh_types.h) stably by
 * their get_ray_hash key over the hash grid (root_min, cell_size) into
 * d_rays_out/d_aux_out, as rh_sort_rays does, on `stream` (null: the default
 * stream) of the current device, asynchronously.  It allocates nothing: the
 * keys, values and digit counts live in a fixed device array of the library
 * (about 42 MB per GPU), so n is at most RH_NOALLOC_MAX_RAYS (larger n is an
 * argument error) and calls that overlap on one GPU race on that array.  The
 * output must not alias the input. */
#define RH_NOALLOC_MAX_RAYS 2097152
int rh_sort_rays_noalloc(const rh_ray *d_rays, const rh_ray_aux *d_aux, size_t n, const float root_min[3],
                         const float cell_size[3], rh_ray *d_rays_out, rh_ray_aux *d_aux_out, void *stream);
/* Shades n (ray, hit) pairs, each with its own pass info; secondary rays are
 * appended in arbitrary order, out_src[k] naming the input ray of out_secondary[k]. */
int rh_shade_rays(rh_context *ctx, const rh_pass_info *pi, const rh_ray_packet *rays, const rh_hit *hits, size_t n,
                  const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_count, int *out_src);

/* ---- differentiation (no counterpart in the reference: the adjoint extension) ----
 * rh_adjoint_pass replays one pass (same camera, rect, iteration and Halton
 * table as the rh_render_pass it differentiates) and adds, for every material,
 *   scale * sum_pixels adj[pixel] . d(pass colour of pixel) / d main_color
 * to d_grad (4 floats per material: r, g, b, unused), and, when d_texel_grad is
 * not null, the same for every stored texel of the albedo textures (4 floats
 * per texel of the scene's texel array, r, g, b per channel value in [0, 1],
 * alpha unused; a block's wrap border and the mip levels are separate stored
 * texels: folding them onto the level-0 texels is the caller's).  d_adj_rgba
 * is the adjoint of the pass image (w*h RGBA floats, alpha unused).  All are
 * device pointers; the call is asynchronous on the context's stream.
 * Paths are treated with their sampling decisions fixed: Russian roulette
 * weights are not differentiated.
 * rh_set_material_colors overwrites main_color of materials [first, first+n)
 * of the uploaded scene from packed rgb triples (host or device pointer).
 * rh_read_accum copies the linear accumulated image (the Ref renderer's
 * "clean" framebuffer, FramebufferRef.cpp:80-86, before tonemapping; w*h RGBA
 * floats) to dst (host or device pointer). */
int rh_adjoint_pass(rh_context *ctx, const rh_pass_desc *pass, const float *d_adj_rgba, float *d_grad,
                    float *d_texel_grad, float scale);
int rh_set_material_colors(rh_context *ctx, uint32_t first, uint32_t n, const float *rgb);
/* limits texel gradients to the texels of the n given textures (n = 0: every
 * texture, the default); the other textures' texels keep a zero gradient */
int rh_set_gradient_textures(rh_context *ctx, const uint32_t *ids, uint32_t n);
/* overwrites stored texels [first, first+n) of the uploaded scene (RGBA8 words) */
int rh_set_texels(rh_context *ctx, uint32_t first, uint32_t n, const uint32_t *rgba);
int rh_read_accum(rh_context *ctx, float *dst, size_t count);

/* ---- batching (no counterpart in the reference: the many-scene extension) ----
 * A batch renders many scenes of different sizes, depths and sample counts in
 * one set of persistent kernels per pass; every scene's image is identical
 * to what rh_render_pass renders for it alone, pass for pass.
 * rh_batch_add_scene copies a scene (its arrays, and the camera, full-frame
 * rect and hash-grid box of desc) and returns its index (>= 0) or a negative
 * status; desc->iteration is ignored.  Pass k (1-based) renders iteration k of
 * every scene with spp >= k; rh_batch_passes gives the largest spp.  The
 * caller uploads the Halton table of iteration k with rh_batch_set_halton
 * before pass k when k == 1 or k % 256 == 0, as RenderScene refreshes it.
 * A pass's launches are captured once as a hipGraph and replayed ("use_graph",
 * default 1).  Knobs: "lds_stack_depth", "path_tail", "max_bounces" (as
 * rh_set_option; max_bounces applies to scenes added after it is set),
 * "use_graph", "blocks_per_cu" (persistent grid size), "sort_bits" (staged
 * bounces sorted by their ray hash), "queue_rays" (rays in flight at most),
 * "pass_overlap" (passes in flight besides the newest, 0..3, default 3: each
 * on its own stream, mixing after the one before).  Changing queue_rays or
 * pass_overlap lays the batch
 * out again and restarts it at the next pass.  rh_batch_passes_done gives the
 * passes rendered since the last restart (0 once such a change is pending).
 * Stats: rays counted as by rh_get_ray_counts (primary, secondary, shadow). */
typedef struct rh_batch rh_batch;
int rh_batch_create(int device, rh_batch **out);
void rh_batch_destroy(rh_batch *batch);
int rh_batch_add_scene(rh_batch *batch, const rh_scene_view *scene, const rh_pass_desc *desc, int w, int h, int spp);
int rh_batch_set_halton(rh_batch *batch, const float *seq, size_t count);
int rh_batch_passes(const rh_batch *batch);
int rh_batch_passes_done(const rh_batch *batch);
int rh_batch_render_pass(rh_batch *batch);
int rh_batch_synchronize(rh_batch *batch);
/* back to pass 0 with cleared images (the scenes stay uploaded) */
int rh_batch_restart(rh_batch *batch);
int rh_batch_read_pixels(rh_batch *batch, int scene, float *rgba, size_t count);
int rh_batch_reset_stats(rh_batch *batch);
int rh_batch_get_ray_counts(rh_batch *batch, rh_ray_counts *out);
int rh_batch_set_option(rh_batch *batch, const char *name, int value);
/* The rays entering stage `stage` (0: primary, 1 + k: bounce k) of scene
 * `scene` in ray buffer `buffer` (0 or 1) after the last pass, for tests of
 * the queue order: *count = the stage's ray count, min(count, cap) rays (the
 * traversal half, rh_ray) copied to out.  With "sort_bits" set and every
 * bounce staged, buffer 0 holds each sorted stage's rays in key order; the
 * last stage's stay there after the pass. */
int rh_batch_read_queue(rh_batch *batch, int scene, int stage, int buffer, rh_ray *out, size_t cap, int *count);

/* ---- device-resident entry point for benchmarking the trace kernel ----
 * rays/hits are device pointers in the split layout (rh_ray / rh_hit). */
int rh_trace_device(rh_context *ctx, const rh_ray *d_rays, int n, int any_hit, rh_hit *d_hits, void *stream);

/* ---- the fast BVH build on the device ----
 * Replaces the bottom half of PreprocessPrims_HLBVH (Core.cpp:532-596) with
 * EmitLBVH (Core.cpp:342-429): Morton codes of the n primitive boxes
 * (prims: 6 floats each, bbox_min xyz then bbox_max xyz), their stable sort,
 * the treelets (runs sharing the top 12 code bits) and each treelet's tree.
 * Outputs, equal to the reference's: out_order[n], the primitives in sorted
 * order; out_nodes, the treelets' nodes in the reference's order (interior
 * child links relative to out_nodes[0], leaves at index_offset + sorted
 * position), *out_node_count of them (nodes_cap >= 2 n); out_treelets, 3
 * words per treelet (first sorted position, count, root node), at most 4096,
 * *out_treelet_count of them.  The host builds the SAH top tree over the
 * treelet roots (Core.cpp:598-680).  A builder owns a stream and scratch
 * buffers that grow to the largest mesh built; calls on one builder
 * serialise. */
typedef struct rh_builder rh_builder;
int rh_builder_create(int device, rh_builder **out);
void rh_builder_destroy(rh_builder *b);
int rh_build_lbvh(rh_builder *b, const float *prims, size_t n, uint32_t index_offset, uint32_t *out_order,
                  rh_bvh_node *out_nodes, size_t nodes_cap, uint32_t *out_node_count, uint32_t *out_treelets,
                  uint32_t *out_treelet_count);

/* ---- the SAH build on the device ----
 * PreprocessPrims_SAH (Core.cpp:431-530) with SplitPrimitives_SAH's object
 * split (BVHSplit.cpp:126-218; no spatial splits) over n boxes (6 floats
 * each, as above), with bvh_settings_t's oversplit_threshold and
 * node_traversal_cost: out_nodes gets the tree, node for node the host's,
 * root first, child links relative to out_nodes[0] (*out_node_count nodes,
 * nodes_cap >= 2 n), and out_indices[n] the box indices its leaves cover
 * (leaf prim_index = RH_LEAF_NODE_BIT | offset into out_indices).  Used for
 * the instance tree (RebuildMacroBVH, SceneRef.cpp:492-516) and the fast
 * build's top tree (Core.cpp:598-680).  Where two boxes share a bbox_max
 * coordinate within a node, or a coordinate is not finite, the host's
 * std::sort order cannot be reproduced: *host_needed = 1, nothing is written
 * and the caller builds on the host. */
int rh_build_sah(rh_builder *b, const float *boxes, size_t n, float oversplit, float cost, rh_bvh_node *out_nodes,
                 size_t nodes_cap, uint32_t *out_node_count, uint32_t *out_indices, int *host_needed);

#ifdef __cplusplus
}
#endif

#endif /* RH_HIP_H */
