/*
 * rh_scene.h — C ABI of the C++ host layer (Ray::Hip::Renderer / Ray::Hip::Scene).
 *
 * This is the flat form of the Ray library's public renderer/scene API, for
 * callers that cannot use C++ classes (Python through ctypes, the bench).
 * Each entry point maps one-to-one onto a method of the reference interface:
 *
 *   rhx_create_renderer        <- Ray::CreateRenderer        RendererFactory.h:19, RendererFactory.cpp:23-75
 *   rhx_destroy_renderer       <- ~RendererBase              RendererBase.h:57
 *   rhx_renderer_type          <- RendererBase::type          RendererBase.h:60
 *   rhx_resize                 <- RendererBase::Resize        RendererBase.h:75
 *   rhx_clear                  <- RendererBase::Clear         RendererBase.h:80
 *   rhx_create_scene           <- RendererBase::CreateScene   RendererBase.h:85
 *   rhx_render_scene           <- RendererBase::RenderScene   RendererBase.h:91
 *   rhx_get_pixels             <- RendererBase::get_pixels_ref RendererBase.h:66
 *   rhx_get_stats/reset_stats  <- RendererBase::GetStats/ResetStats RendererBase.h:93-102
 *   rhx_region_*               <- RegionContext               RendererBase.h:35-51
 *   rhx_set_environment        <- SceneBase::SetEnvironment   SceneBase.h:175
 *   rhx_add_texture            <- SceneBase::AddTexture       SceneBase.h:181
 *   rhx_add_material           <- SceneBase::AddMaterial      SceneBase.h:192
 *   rhx_add_mesh               <- SceneBase::AddMesh          SceneBase.h:203
 *   rhx_remove_mesh            <- SceneBase::RemoveMesh       SceneBase.h:208
 *   rhx_add_light              <- SceneBase::AddLight         SceneBase.h:214
 *   rhx_add_mesh_instance      <- SceneBase::AddMeshInstance  SceneBase.h:226
 *   rhx_set_mesh_instance_transform <- SceneBase::SetMeshInstanceTransform SceneBase.h:232
 *   rhx_remove_mesh_instance   <- SceneBase::RemoveMeshInstance SceneBase.h:240
 *   rhx_add_camera/set_current_cam <- SceneBase::AddCamera / set_current_cam SceneBase.h:246-278
 *   rhx_triangle_count/node_count  <- SceneBase::triangle_count/node_count SceneBase.h:281-284
 *
 * rhx_scene_view is not part of the reference API: it exposes the scene's
 * preprocessed arrays (BVH nodes, triangle records, ...) so tests can compare
 * them with the reference's and hand them to the CPU oracle.
 *
 * Errors: functions returning int give 0 on success and a negative code on
 * failure (message via rhx_last_error).  Add* functions return 0xffffffff on
 * failure, as the reference does (SceneRef.cpp:80-88).
 */
#ifndef RH_SCENE_H
#define RH_SCENE_H

#include "rh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rhx_renderer rhx_renderer;
typedef struct rhx_scene rhx_scene;
typedef struct rhx_region rhx_region;

/* renderer type bits (RendererBase.h:14-21) plus this backend's bit */
enum {
    RHX_RENDERER_REF = 1 << 0,
    RHX_RENDERER_SSE2 = 1 << 1,
    RHX_RENDERER_AVX = 1 << 2,
    RHX_RENDERER_AVX2 = 1 << 3,
    RHX_RENDERER_NEON = 1 << 4,
    RHX_RENDERER_OCL = 1 << 5,
    RHX_RENDERER_HIP = 1 << 6
};

/* mat_desc_t (SceneBase.h:55-66) */
typedef struct rhx_mat_desc {
    int32_t type;
    float main_color[3];
    uint32_t main_texture;
    uint32_t normal_map;
    uint32_t mix_materials[2];
    float roughness;
    float strength;
    float fresnel;
    float int_ior;
    float ext_ior;
} rhx_mat_desc;

/* shape_desc_t (SceneBase.h:69-81) */
typedef struct rhx_shape_desc {
    uint32_t mat_index;
    uint32_t back_mat_index;
    uint64_t vtx_start;
    uint64_t vtx_count;
} rhx_shape_desc;

/* light_desc_t (SceneBase.h:113-118) */
typedef struct rhx_light_desc {
    int32_t type;
    float position[3], radius;
    float color[3];
    float direction[3], angle;
} rhx_light_desc;

/* camera_desc_t (SceneBase.h:121-145) */
typedef struct rhx_cam_desc {
    int32_t type, filter, dtype;
    float origin[3], fwd[3], up[3];
    float fov, gamma, focus_distance, focus_factor;
    uint32_t mi_index, uv_index;
    uint8_t lighting_only, skip_direct_lighting, skip_indirect_lighting, no_background;
    uint8_t clamp, output_sh, use_coherent_sampling, pad0;
    uint8_t max_diff_depth, max_glossy_depth, max_refr_depth, max_transp_depth;
    uint8_t max_total_depth, termination_start_depth, pad1[2];
} rhx_cam_desc;

/* settings_t (RendererBase.h:24-30) plus the HIP knobs */
typedef struct rhx_settings {
    int32_t w, h;
    int32_t device_index;    /* HIP device ordinal, -1 = current */
    int32_t use_wide_bvh;    /* ignored by the HIP backend: it walks binary nodes */
    int32_t sort_rays;       /* 1 = sort secondary rays by hash before tracing (default) */
    int32_t lds_stack_depth; /* 0 = default */
} rhx_settings;

const char *rhx_last_error(void);

int rhx_create_renderer(const rhx_settings *s, uint32_t flags, rhx_renderer **out);
void rhx_destroy_renderer(rhx_renderer *r);
uint32_t rhx_renderer_type(const rhx_renderer *r);
int rhx_resize(rhx_renderer *r, int w, int h);
int rhx_clear(rhx_renderer *r, const float rgba[4]);
int rhx_get_pixels(rhx_renderer *r, float *out_rgba, size_t count);
/* The accumulated SH data (RendererBase::get_sh_data_ref, RendererBase.h:69):
 * w*h records of 12 floats (shl1_data_t); returns 1 when no pass with
 * OutputSH has run since the last resize. */
int rhx_get_sh_data(rhx_renderer *r, float *out, size_t count);
int rhx_get_stats(rhx_renderer *r, rh_stats *out);
int rhx_reset_stats(rhx_renderer *r);
int rhx_get_ray_counts(rhx_renderer *r, rh_ray_counts *out);
int rhx_set_option(rhx_renderer *r, const char *name, int value);
/* the current value of a knob of rh_set_option (HIP backend only) */
int rhx_get_option(rhx_renderer *r, const char *name, int *value);
/* Device-wide synchronisation, and the per-kernel profile of the device layer
 * (HIP backend only; see rh_get_kernel_profile in rh_hip.h). */
int rhx_synchronize(rhx_renderer *r);
int rhx_get_kernel_profile(rhx_renderer *r, rh_kernel_profile *out);
int rhx_reset_kernel_profile(rhx_renderer *r);

int rhx_create_scene(rhx_renderer *r, rhx_scene **out);
/* A scene built on the host only (no renderer, no device), with binary or
 * 8-wide trees (settings use_wide_bvh); any renderer or batch of the library
 * can render it.  No counterpart in the reference, whose scenes come from
 * RendererBase::CreateScene. */
int rhx_create_scene_standalone(int use_wide_bvh, rhx_scene **out);
void rhx_destroy_scene(rhx_scene *s);

/* The fast (HLBVH) tree over n caller boxes (6 floats each: bbox_min xyz,
 * bbox_max xyz), as PreprocessPrims_HLBVH (Core.cpp:532-680) builds a mesh's
 * tree from its triangles' boxes: on the host (device < 0) or with the
 * treelets built on GPU `device`.  out_nodes (capacity nodes_cap >= 2 n)
 * receives *out_count nodes, root first, leaves indexing out_indices[n] (the
 * boxes in Morton order).  Scenes use the same build for meshes with
 * use_fast_bvh_build; this entry point has no counterpart in the reference's
 * API (it exposes that build for any boxes). */
int rhx_build_fast_bvh(const float *boxes, size_t n, int device, rh_bvh_node *out_nodes, size_t nodes_cap,
                       uint32_t *out_count, uint32_t *out_indices);
/* PreprocessPrims_SAH (Core.cpp:431-530) without spatial splits over n boxes
   (6 floats each: bbox_min, bbox_max) with bvh_settings_t's
   oversplit_threshold and node_traversal_cost, as the instance tree
   (RebuildMacroBVH, SceneRef.cpp:492-516) is built: on the host (device < 0)
   or on GPU `device` (rh_build_sah), *on_device telling which built it (the
   host builds where boxes share a bbox_max coordinate).  Same outputs as
   rhx_build_fast_bvh. */
int rhx_build_sah(const float *boxes, size_t n, int device, float oversplit, float cost, rh_bvh_node *out_nodes,
                  size_t nodes_cap, uint32_t *out_count, uint32_t *out_indices, int *on_device);

int rhx_region_create(int x, int y, int w, int h, rhx_region **out);
void rhx_region_destroy(rhx_region *reg);
int rhx_region_iteration(const rhx_region *reg);
int rhx_region_clear(rhx_region *reg);

int rhx_render_scene(rhx_renderer *r, rhx_scene *s, rhx_region *reg);

/* Differentiation (HIP backend only; no counterpart in the reference, see
 * rh_adjoint_pass in rh_hip.h).  rhx_adjoint_scene advances reg exactly as
 * rhx_render_scene would and replays that pass, adding scale * d loss /
 * d main_color (4 floats per material) to d_grad, and d loss / d stored
 * texel to d_texel_grad when not null (4 floats per texel), for the adjoint
 * image d_adj_rgba (device pointers).  rhx_set_material_colors sets
 * main_color of materials [first, first+n) from host rgb triples, and
 * rhx_set_texture_data a texture's level-0 RGBA8 texels (its bordered blocks
 * and mips rebuilt), in the scene and in its device copy.  rhx_read_accum
 * copies the linear accumulated image (w*h RGBA). */
int rhx_adjoint_scene(rhx_renderer *r, rhx_scene *s, rhx_region *reg, const float *d_adj_rgba, float *d_grad,
                      float *d_texel_grad, float scale);
int rhx_set_material_colors(rhx_renderer *r, rhx_scene *s, uint32_t first, uint32_t n, const float *rgb);
int rhx_set_texture_data(rhx_renderer *r, rhx_scene *s, uint32_t tex, const uint8_t *rgba);
/* texel gradients only for the n given textures (n = 0: all); rh_set_gradient_textures */
int rhx_set_gradient_textures(rhx_renderer *r, const uint32_t *ids, uint32_t n);
int rhx_read_accum(rhx_renderer *r, float *dst, size_t count);

/* Batching (HIP backend only; no counterpart in the reference, see
 * rh_batch_* in rh_hip.h): Ray::Hip::BatchRenderer.  rhx_batch_add copies a
 * scene with its current camera, to be rendered at w x h for spp passes, and
 * returns its index (< 0 on error); scenes are added before the first pass.
 * rhx_batch_render_pass renders one progressive pass of every scene with
 * samples left; rhx_batch_get_pixels reads scene i's tonemapped image (RGBA
 * w*h), equal to what rhx_render_scene gives for that scene after as many
 * passes. */
typedef struct rhx_batch rhx_batch;
int rhx_batch_create(int device_index, rhx_batch **out);
void rhx_batch_destroy(rhx_batch *b);
int rhx_batch_add(rhx_batch *b, rhx_scene *s, int w, int h, int spp);
int rhx_batch_passes(const rhx_batch *b);
int rhx_batch_render_pass(rhx_batch *b);
int rhx_batch_get_pixels(rhx_batch *b, int scene, float *out, size_t count);
int rhx_batch_set_option(rhx_batch *b, const char *name, int value);
int rhx_batch_get_ray_counts(rhx_batch *b, rh_ray_counts *out);
int rhx_batch_reset_stats(rhx_batch *b);
int rhx_batch_synchronize(rhx_batch *b);
int rhx_batch_restart(rhx_batch *b); /* back to pass 0 with cleared images */
/* rh_batch_read_queue of the batch's device pipeline (queue-order tests) */
int rhx_batch_read_queue(rhx_batch *b, int scene, int stage, int buffer, rh_ray *out, size_t cap, int *count);

/* The adders return 0xffffffff, and rhx_set_environment -1, with the reason in
   rhx_last_error(), when an index names no existing texture or material
   (main_texture and mix components required, normal_map and env_map optional
   as 0xffffffff), a mesh index names no vertex, or a shape range passes the
   mesh's indices; nothing is changed then.  (The reference stores such
   indices unchecked and reads through them when rendering.) */
int rhx_set_environment(rhx_scene *s, const float env_col[3], float env_clamp, uint32_t env_map);
uint32_t rhx_add_texture(rhx_scene *s, const uint8_t *rgba, int w, int h, int is_srgb, int generate_mipmaps);
uint32_t rhx_add_material(rhx_scene *s, const rhx_mat_desc *m);
/* bvh_flags: RHX_MESH_SPATIAL_SPLITS (mesh_desc_t::allow_spatial_splits),
   RHX_MESH_FAST_BVH (mesh_desc_t::use_fast_bvh_build, SceneBase.h:94: the
   HLBVH build of Core.cpp:532-680 in place of the SAH build).
   attrs_count is the number of vertices (mesh_desc_t::vtx_attrs_count,
   SceneBase.h:88), each of the layout's stride in floats. */
#define RHX_MESH_SPATIAL_SPLITS 1
#define RHX_MESH_FAST_BVH 2
uint32_t rhx_add_mesh(rhx_scene *s, int layout, const float *attrs, size_t attrs_count,
                      const uint32_t *indices, size_t indices_count, int base_vertex,
                      const rhx_shape_desc *shapes, size_t shapes_count, int bvh_flags);
void rhx_remove_mesh(rhx_scene *s, uint32_t i);
uint32_t rhx_add_light(rhx_scene *s, const rhx_light_desc *l);
uint32_t rhx_add_mesh_instance(rhx_scene *s, uint32_t mesh, const float xform[16]);
void rhx_set_mesh_instance_transform(rhx_scene *s, uint32_t mi, const float xform[16]);
void rhx_remove_mesh_instance(rhx_scene *s, uint32_t mi);
uint32_t rhx_add_camera(rhx_scene *s, const rhx_cam_desc *c);
void rhx_set_camera(rhx_scene *s, uint32_t i, const rhx_cam_desc *c);
void rhx_set_current_cam(rhx_scene *s, uint32_t i);
int rhx_get_camera_record(rhx_scene *s, uint32_t i, rh_camera *out);
uint32_t rhx_triangle_count(rhx_scene *s);
uint32_t rhx_node_count(rhx_scene *s);

int rhx_scene_view(rhx_scene *s, rh_scene_view *out);
/* The Halton table used by the next RenderScene of this region (256 x 24 floats). */
int rhx_region_halton(const rhx_region *reg, float *out, size_t count);

#ifdef __cplusplus
}
#endif

#endif /* RH_SCENE_H */
