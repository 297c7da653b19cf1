// gfx950 kernels of the HIP rendering backend and the C ABI that drives them
// (include/rh_hip.h).  One RenderScene pass is a wavefront pipeline:
//
//   gen_primary -> trace -> shade(primary) -> [ sort -> trace -> shade ] x depth -> mix_tonemap
//
// Queue lengths live on the device: each kernel reads its ray count from a
// device counter and the shade kernel appends the next bounce's rays with one
// atomic per wave, so the host never waits on the GPU inside a pass.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <vector>

#include "rh_radix.h"
#include "rh_runtime.h"

using rh_rt::arg_error;
using rh_rt::DevBuf;
using rh_rt::fail;
using rh_rt::free_ptr;
using rh_rt::release;
using rh_rt::upload_halton;

namespace {

// Queue position i -> pixel (x, y) of a rw x rh region in 8x8 tiles, one
// wave each: the tiles cover the part whose sides are multiples of 8, a row
// of tiles after another; the right and then the bottom remainder follow row
// by row.  A bijection on [0, rw rh).  A wave's rays then start from a compact
// block of pixels instead of a 64-pixel strip, and so do the bounce and
// shadow rays that keep their queue order, which keeps a wave's traversals
// on fewer nodes.
__device__ __forceinline__ void tiled_pixel(int i, int rw, int rh, int &x, int &y) {
    const int w8 = rw & ~7, h8 = rh & ~7, na = w8 * h8;
    if (i < na) {
        const int t = i >> 6, j = i & 63, tw = w8 >> 3;
        x = ((t % tw) << 3) + (j & 7);
        y = ((t / tw) << 3) + (j >> 3);
        return;
    }
    i -= na;
    const int rwr = rw - w8, nb = rwr * h8;
    if (i < nb) {
        x = w8 + i % rwr;
        y = i / rwr;
        return;
    }
    i -= nb;
    x = i % rw;
    y = h8 + i / rw;
}

// one lane per pixel of the region, in row-major order or (Tiled) in tiles
template <bool Tiled>
__global__ void __launch_bounds__(kBlock) k_gen_primary(PassConst pc, const float *__restrict__ halton,
                                                        rh_ray *__restrict__ rays, rh_ray_aux *__restrict__ aux,
                                                        int *__restrict__ count) {
    const int rw = pc.rect[2], rh_ = pc.rect[3];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = rw * rh_;
    if (i >= rw * rh_) return;
    int x = i % rw, y = i / rw;
    if (Tiled) tiled_pixel(i, rw, rh_, x, y);
    store_ray(rays, aux, (uint32_t)i, gen_primary_ray(pc, halton, pc.rect[0] + x, pc.rect[1] + y));
}

// ---- geometry camera: SampleMeshInTextureSpace (CoreRef.cpp:507-613) ----
// The lightmap pixel (x, y) of the region takes the first triangle of the
// camera's mesh instance (in triangle order) whose texture-space footprint
// covers the pixel's jittered sample; the ray starts at the surface point
// plus its normal and points back along the normal, with a made-up hit at
// t = 1 on that triangle.  Here every triangle marks the pixels it covers
// with an atomic minimum of its index, and each pixel then builds its ray
// from the winning triangle with the same arithmetic.
struct GeoArgs {
    const rh_vertex *vertices;
    const uint32_t *vtx_indices;
    const rh_mesh_instance *mesh_instances;
    const rh_mesh *meshes;
    const rh_transform *transforms;
    const float *halton;
    uint32_t *winner;
    int rect[4];
    int w, h, iteration;
    uint32_t mi_index, uv;
};

struct GeoTri {
    float t0x, t0y, t1x, t1y, t2x, t2y;
    float d01x, d01y, d12x, d12y, d20x, d20y, inv_area;
    int x0, y0, x1, y1; // clipped pixel box
};

// the triangle's texture-space setup; false when it misses the region or is
// degenerate (or back-facing) in texture space
__device__ __forceinline__ bool geo_setup(const GeoArgs &g, uint32_t tri, GeoTri &t) {
    const float W = (float)g.w, H = (float)g.h;
    const rh_vertex &v0 = g.vertices[g.vtx_indices[tri * 3 + 0]], &v1 = g.vertices[g.vtx_indices[tri * 3 + 1]],
                    &v2 = g.vertices[g.vtx_indices[tri * 3 + 2]];
    t.t0x = v0.t[g.uv][0] * W;
    t.t0y = (1.0f - v0.t[g.uv][1]) * H;
    t.t1x = v1.t[g.uv][0] * W;
    t.t1y = (1.0f - v1.t[g.uv][1]) * H;
    t.t2x = v2.t[g.uv][0] * W;
    t.t2y = (1.0f - v2.t[g.uv][1]) * H;
    const float bminx = smin(smin(t.t0x, t.t1x), t.t2x), bminy = smin(smin(t.t0y, t.t1y), t.t2y);
    const float bmaxx = smax(smax(t.t0x, t.t1x), t.t2x), bmaxy = smax(smax(t.t0y, t.t1y), t.t2y);
    int x0 = f2i_trunc(bminx), y0 = f2i_trunc(bminy);
    int x1 = f2i_trunc(__builtin_roundf(bmaxx)), y1 = f2i_trunc(__builtin_roundf(bmaxy));
    const int rx0 = g.rect[0], ry0 = g.rect[1], rx1 = g.rect[0] + g.rect[2] - 1, ry1 = g.rect[1] + g.rect[3] - 1;
    if (x1 < rx0 || y1 < ry0 || x0 > rx1 || y0 > ry1) return false;
    t.x0 = max(x0, rx0);
    t.y0 = max(y0, ry0);
    t.x1 = min(x1, rx1);
    t.y1 = min(y1, ry1);
    t.d01x = t.t0x - t.t1x;
    t.d01y = t.t0y - t.t1y;
    t.d12x = t.t1x - t.t2x;
    t.d12y = t.t1y - t.t2y;
    t.d20x = t.t2x - t.t0x;
    t.d20y = t.t2y - t.t0y;
    const float area = t.d01x * t.d20y - t.d20x * t.d01y;
    if (area < RH_FLT_EPS) return false;
    t.inv_area = 1.0f / area;
    return true;
}

// the edge functions of pixel (x, y)'s jittered sample; true when inside
__device__ __forceinline__ bool geo_sample(const GeoArgs &g, const GeoTri &t, int x, int y, float &u, float &v,
                                           float &w) {
    const int index = y * g.w + x;
    const int hi = (g.iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT;
    const int hash_val = hash(index);
    const float sx = (float)x + fract_modf(g.halton[hi + 0] + construct_float((uint32_t)hash_val));
    const float sy = (float)y + fract_modf(g.halton[hi + 1] + construct_float((uint32_t)hash(hash_val)));
    u = t.d01x * (sy - t.t0y) - t.d01y * (sx - t.t0x);
    v = t.d12x * (sy - t.t1y) - t.d12y * (sx - t.t1x);
    w = t.d20x * (sy - t.t2y) - t.d20y * (sx - t.t2x);
    return u >= -RH_FLT_EPS && v >= -RH_FLT_EPS && w >= -RH_FLT_EPS;
}

// every lane of the region: white throughput, no hit yet
__global__ void __launch_bounds__(kBlock) k_geo_init(GeoArgs g, rh_ray *__restrict__ rays, rh_ray_aux *__restrict__ aux,
                                                    rh_hit *__restrict__ hits, int *__restrict__ count) {
    const int rw = g.rect[2], n = rw * g.rect[3];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = n;
    if (i >= n) return;
    const int x = g.rect[0] + i % rw, y = g.rect[1] + i / rw;
    RayFull r;
    memset(&r, 0, sizeof(r));
    r.c = f3{1.0f, 1.0f, 1.0f};
    r.ior = 1.0f;
    r.xy = (x << 16) | y;
    store_ray(rays, aux, (uint32_t)i, r);
    rh_hit hh;
    hh.mask = 0;
    hh.obj = -1;
    hh.prim = -1;
    hh.t = RH_MAX_DIST;
    hh.u = hh.v = 0.0f;
    hh.xy = r.xy;
    hh.pad = 0;
    hits[i] = hh;
    g.winner[i] = 0xffffffffu;
}

// one wave per triangle (a grid-stride loop over the mesh's triangles), its
// lanes over the triangle's pixel box
__global__ void __launch_bounds__(kBlock) k_geo_raster(GeoArgs g) {
    const rh_mesh_instance &mi = g.mesh_instances[g.mi_index];
    const rh_mesh &m = g.meshes[mi.mesh_index];
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t waves = gridDim.x * (kBlock / 64);
    for (uint32_t k = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); k < m.tris_count; k += waves) {
        const uint32_t tri = m.tris_index + k;
        GeoTri t;
        if (!geo_setup(g, tri, t)) continue;
        const int bw = t.x1 - t.x0 + 1, bh = t.y1 - t.y0 + 1;
        const int64_t npx = (int64_t)bw * bh;
        for (int64_t q = lane; q < npx; q += 64) {
            const int x = t.x0 + (int)(q % bw), y = t.y0 + (int)(q / bw);
            float u, v, w;
            if (geo_sample(g, t, x, y, u, v, w))
                atomicMin(&g.winner[(y - g.rect[1]) * g.rect[2] + (x - g.rect[0])], tri);
        }
    }
}

// the ray and hit of each covered pixel, from its winning triangle
__global__ void __launch_bounds__(kBlock) k_geo_finish(GeoArgs g, rh_ray *__restrict__ rays,
                                                      rh_ray_aux *__restrict__ aux, rh_hit *__restrict__ hits) {
    const int rw = g.rect[2], n = rw * g.rect[3];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tri = g.winner[i];
    if (tri == 0xffffffffu) return;
    const int x = g.rect[0] + i % rw, y = g.rect[1] + i / rw;
    GeoTri t;
    float u, v, w;
    geo_setup(g, tri, t);
    geo_sample(g, t, x, y, u, v, w);
    u *= t.inv_area;
    v *= t.inv_area;
    w *= t.inv_area;
    const rh_vertex &v0 = g.vertices[g.vtx_indices[tri * 3 + 0]], &v1 = g.vertices[g.vtx_indices[tri * 3 + 1]],
                    &v2 = g.vertices[g.vtx_indices[tri * 3 + 2]];
    const rh_transform &tr = g.transforms[g.mesh_instances[g.mi_index].tr_index];
    const f3 pl = f3{v0.p[0] * v + v1.p[0] * w + v2.p[0] * u, v0.p[1] * v + v1.p[1] * w + v2.p[1] * u,
                     v0.p[2] * v + v1.p[2] * w + v2.p[2] * u};
    const f3 nl = f3{v0.n[0] * v + v1.n[0] * w + v2.n[0] * u, v0.n[1] * v + v1.n[1] * w + v2.n[1] * u,
                     v0.n[2] * v + v1.n[2] * w + v2.n[2] * u};
    const float *X = tr.xform;
    const f3 p = f3{X[0] * pl.x + X[4] * pl.y + X[8] * pl.z + X[12], X[1] * pl.x + X[5] * pl.y + X[9] * pl.z + X[13],
                    X[2] * pl.x + X[6] * pl.y + X[10] * pl.z + X[14]};
    const f3 nn = transform_normal(nl, tr.inv_xform);
    RayFull r;
    memset(&r, 0, sizeof(r));
    r.o = p + nn;
    r.d = f3{-nn.x, -nn.y, -nn.z};
    r.c = f3{1.0f, 1.0f, 1.0f};
    r.ior = 1.0f;
    r.xy = (x << 16) | y;
    store_ray(rays, aux, (uint32_t)i, r);
    rh_hit hh;
    hh.mask = -1;
    hh.obj = (int)g.mi_index;
    hh.prim = (int)tri;
    hh.t = 1.0f;
    hh.u = w;
    hh.v = u;
    hh.xy = r.xy;
    hh.pad = 0;
    hits[i] = hh;
}

// ---- spherical-harmonic output (OutputSH; FramebufferRef.h:58-90,
// FramebufferRef.cpp:48-94) ----
// Per pixel 12 floats (coeff_r[4], coeff_g[4], coeff_b[4]).  After the
// primary bounce the pixel's secondary ray stores its direction's L1 basis in
// coeff_r and its weight in coeff_g[0]; after the pass the pass image is
// projected (scaled by 2 pi / weight) and mixed into the accumulated SH.
constexpr float kShY0 = 0.282094806f, kShY1 = 0.488602519f;

__global__ void __launch_bounds__(256) k_sh_reset_weights(float *__restrict__ sh, int x0, int y0, int rw, int rh_,
                                                          int w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rw * rh_) return;
    const size_t p = (size_t)(y0 + i / rw) * w + (x0 + i % rw);
    sh[12 * p + 4] = 0.0f;
}

__global__ void __launch_bounds__(kBlock) k_sh_samples(const rh_ray *__restrict__ rays, const rh_ray_aux *__restrict__ aux,
                                                      const int *__restrict__ count, int w, float *__restrict__ sh) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const RayFull r = load_ray(rays, aux, (uint32_t)i);
    const int x = (r.xy >> 16) & 0xffff, y = r.xy & 0xffff;
    float *s = sh + 12 * ((size_t)y * w + x);
    s[0] = kShY0;
    s[1] = kShY1 * r.d.y;
    s[2] = kShY1 * r.d.z;
    s[3] = kShY1 * r.d.x;
    s[4] = r.c.x;
}

__global__ void __launch_bounds__(256) k_sh_finish(const float4 *__restrict__ temp, float *__restrict__ sh_temp,
                                                   float *__restrict__ sh_clean, int x0, int y0, int rw, int rh_, int w,
                                                   float k) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rw * rh_) return;
    const size_t p = (size_t)(y0 + i / rw) * w + (x0 + i % rw);
    float *s = sh_temp + 12 * p, *o = sh_clean + 12 * p;
    const float inv_weight = s[4] > RH_FLT_EPS ? (2.0f * RH_PI / s[4]) : 0.0f;
    float4 c = temp[p];
    c.x *= inv_weight;
    c.y *= inv_weight;
    c.z *= inv_weight;
    float coeff[4] = {s[0], s[1], s[2], s[3]};
    for (int j = 0; j < 4; j++) {
        s[4 + j] = coeff[j] * c.y;
        s[8 + j] = coeff[j] * c.z;
        s[j] = coeff[j] * c.x;
    }
    for (int j = 0; j < 12; j++) o[j] += (s[j] - o[j]) * k;
}

// Tree: the scene's tree kind (kTreeBinary / kTreeWide), one walk compiled per
// kernel.  MinBlocks 5 caps an 8-wide walk's registers at 5 waves per SIMD
// (some spills) instead of the 4 it would take (knob trace_occupancy).
template <bool AnyHit, bool Count, int Tree, int MinBlocks = 1>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_trace(TraceArgs a) {
    extern __shared__ uint32_t lds_stack[];
    trace_body<AnyHit, Count, Tree>(a, blockIdx.x * blockDim.x, lds_stack);
}

// 8-wide trees, G = 2, 4 or 8 lanes per ray (trace_group_body)
template <bool AnyHit, bool Count, int G, int MinBlocks>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_trace_group(TraceArgs a) {
    extern __shared__ uint32_t lds_stack[];
    trace_group_body<AnyHit, Count, G>(a, lds_stack);
}

// persistent closest-hit trace with lane refill (trace_refill_body)
template <bool Count> __global__ void __launch_bounds__(kBlock) k_trace_refill(TraceArgs a, int *next_ray) {
    extern __shared__ uint32_t lds_stack[];
    trace_refill_body<Count>(a, next_ray, lds_stack);
}

// the same for 8-wide trees (trace_refill_wide_body); MinBlocks as in k_trace
template <bool Count, int MinBlocks>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_trace_refill_wide(TraceArgs a, int *next_ray) {
    extern __shared__ uint32_t lds_stack[];
    trace_refill_wide_body<Count>(a, next_ray, lds_stack);
}

// Defer: direct light goes to the NEE stage (no traversal in this kernel).
template <bool Defer> __global__ void __launch_bounds__(kBlock) k_shade(ShadeArgs a) {
    extern __shared__ uint32_t lds_stack[];
    __shared__ int s_app[kMaxDeferredLights * (kBlock / 64) + 1];
    shade_body<Defer>(a, blockIdx.x * blockDim.x, lds_stack, s_app);
}

// the bounces first..last-1 of each queued ray in one lane (path_tail_body);
// with next_path a persistent grid whose lanes refill from the queue.
// MinBlocks 3 asks the compiler for 3 resident blocks (waves per SIMD) per CU,
// trading register spills for occupancy; 2 gives it 256 registers (2 waves per
// SIMD).  Which is faster is measured (tail_occupancy).
// Count: the instrumented form (traversal visits into a.visits), for the
// roofline's algorithmic bytes.
// Solid: for scenes whose triangles are all solid (compute_visibility).
template <bool Defer, int MinBlocks, int Tree, bool Count = false, bool Solid = false>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_path_tail(ShadeArgs a, int first, int last, int *next_path) {
    extern __shared__ uint32_t lds_stack[];
    path_tail_body<Defer, Tree, Count, Solid>(a, first, last, blockIdx.x * blockDim.x, lds_stack, next_path);
}

// ---------------------------------------------------------------- adjoint

struct AdjointArgs {
    Scene sc;
    PassConst pc;
    rh_pass_info pi;    // pass-level; index/rand_index derived per pixel
    const float *halton;
    const float4 *adj;  // dLoss/d(pass image) per pixel of the frame (rgb, w unused)
    float scale;        // weight of this pass in the image (1 / passes for the mean)
    float *grad;        // 4 floats per material: d loss / d main_color (rgb, pad)
    float *tex_grad;    // optional: 4 floats per stored texel, d loss / d texel value (rgb, a unused)
    const uint32_t *tex_mask; // optional: bit t set = texture t's texels get gradients (null: all)
    uint32_t n_materials;
    float4 *rec;        // per-lane vertex records, max_vertices x lanes, lane-minor
    float4 *rec_tex;    // with tex_grad: 3 records per vertex (texels, weights + raw value, raw + sRGB flag)
    uint32_t lanes;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    unsigned long long *ray_counter;
    unsigned long long *shadow_counter;
    unsigned long long *overflow;
    // light samples recorded by shading and resolved in the lane (Defer)
    float4 *nee_ray;
    int2 *nee_aux;
    float4 *nee_term;
    uint32_t nee_slots;
    // texel blocks of at most kSmallBlockTexels (the coarse mip levels, which
    // many pixels share) are summed in kTexelReplicas copies, a wave adding to
    // one of them, and folded into tex_grad after the pass (k_fold_replicas):
    // coarse_map[tex * RH_NUM_MIP_LEVELS + lod] = (block's first texel, its
    // first slot in a copy, or kNoKey when the block is not replicated)
    float *rep;
    uint32_t rep_count; // slots per copy
    const uint2 *coarse_map;
};

constexpr uint32_t kLdsGradMaterials = 1024;
constexpr uint32_t kTexelReplicas = 32;
constexpr uint32_t kSmallBlockTexels = 66 * 66;

constexpr int kAggRounds = 16;

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

constexpr uint32_t kNoKey = 0xffffffffu;

// Adds v to dst[4 key + 0..2] (lanes with key kNoKey add nothing); every lane
// of the wave calls it.  Lanes adding to the same key are summed in registers
// first and one lane adds the sum: same-address atomics serialise, and a wave
// often adds to one address (few materials, triangles a few pixels wide, the
// texels of a coarse mip level under many pixels).  After kAggRounds distinct
// keys the remaining lanes add their own values.
__device__ __forceinline__ void wave_add3(float *dst, uint32_t key, f3 v) {
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(key != kNoKey);
    for (int round = 0; pending && round < kAggRounds; round++) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t kk = __shfl(key, leader);
        const bool mine = key == kk;
        const float sx = wave_sum(mine ? v.x : 0.0f), sy = wave_sum(mine ? v.y : 0.0f),
                    sz = wave_sum(mine ? v.z : 0.0f);
        if ((int)lane == leader) {
            float *d = dst + 4 * (size_t)kk;
            if (sx != 0.0f) atomicAdd(d + 0, sx);
            if (sy != 0.0f) atomicAdd(d + 1, sy);
            if (sz != 0.0f) atomicAdd(d + 2, sz);
        }
        pending &= ~__ballot(mine);
    }
    if ((pending >> lane) & 1u) {
        float *d = dst + 4 * (size_t)key;
        if (v.x != 0.0f) atomicAdd(d + 0, v.x);
        if (v.y != 0.0f) atomicAdd(d + 1, v.y);
        if (v.z != 0.0f) atomicAdd(d + 2, v.z);
    }
}

// wave_add3 with a per-lane target: lanes with the same key have the same
// target (the key names the address within the lane's copy)
__device__ __forceinline__ void wave_add3_at(float *target, uint32_t key, f3 v) {
    const uint32_t lane = __lane_id();
    uint64_t pending = __ballot(key != kNoKey);
    for (int round = 0; pending && round < kAggRounds; round++) {
        const int leader = __builtin_ctzll(pending);
        const uint32_t kk = __shfl(key, leader);
        const bool mine = key == kk;
        const float sx = wave_sum(mine ? v.x : 0.0f), sy = wave_sum(mine ? v.y : 0.0f),
                    sz = wave_sum(mine ? v.z : 0.0f);
        if ((int)lane == leader) {
            if (sx != 0.0f) atomicAdd(target + 0, sx);
            if (sy != 0.0f) atomicAdd(target + 1, sy);
            if (sz != 0.0f) atomicAdd(target + 2, sz);
        }
        pending &= ~__ballot(mine);
    }
    if ((pending >> lane) & 1u) {
        if (v.x != 0.0f) atomicAdd(target + 0, v.x);
        if (v.y != 0.0f) atomicAdd(target + 1, v.y);
        if (v.z != 0.0f) atomicAdd(target + 2, v.z);
    }
}

// The backward half of the replay for lane i (pixel adjoint dl) over its nv
// vertex records: S_k = sum_{j >= k} C_j, added as adj . S_k / main_color to the
// material of vertex k and, with texel gradients, to its texel footprint.
// Every lane of the wave calls it (lanes with nothing to add pass nv = 0): the
// vertex loop runs over the wave's longest path so that the sums can be formed
// across the wave.
__device__ __forceinline__ void adjoint_accumulate(const AdjointArgs &a, uint32_t i, int nv, float4 dl, bool lds_grad,
                                                   float *s_grad) {
    int kmax = nv;
    for (int off = 32; off > 0; off >>= 1) kmax = max(kmax, __shfl_xor(kmax, off));
    float *mdst = lds_grad ? s_grad : a.grad;
    f3 S = f3{0.0f, 0.0f, 0.0f};
    for (int k = kmax - 1; k >= 0; k--) {
        int m = -1;
        f3 v = f3{0.0f, 0.0f, 0.0f};
        f3 g = f3{0.0f, 0.0f, 0.0f};
        if (k < nv) {
            const float4 r = a.rec[(size_t)k * a.lanes + i];
            S = S + f3{r.x, r.y, r.z};
            m = __float_as_int(r.w);
            if (m >= 0) {
                const float *mc = a.sc.materials[m].main_color;
                g = f3{dl.x * S.x, dl.y * S.y, dl.z * S.z};
                if (mc[0] != 0.0f && g.x != 0.0f) v.x = a.scale * (g.x / mc[0]);
                if (mc[1] != 0.0f && g.y != 0.0f) v.y = a.scale * (g.y / mc[1]);
                if (mc[2] != 0.0f && g.z != 0.0f) v.z = a.scale * (g.z / mc[2]);
            }
        }
        wave_add3(mdst, m >= 0 ? (uint32_t)m : kNoKey, v);
        if (!a.tex_grad) continue;
        // albedo = lin(raw) * main_color with raw the bilinear sum; S is linear
        // in the albedo, so d pixel / d raw = S / lin(raw) * lin'(raw), and
        // each footprint texel takes its bilinear weight's share
        uint32_t idx[4] = {kNoKey, kNoKey, kNoKey, kNoKey};
        float w[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        float gr[3] = {0.0f, 0.0f, 0.0f};
        uint2 cm = make_uint2(0u, kNoKey); // small texel blocks go to this wave's replica
        if (m >= 0) {
            const float4 *rt = a.rec_tex + (size_t)k * 3 * a.lanes + i;
            const float4 t2 = rt[2 * (size_t)a.lanes];
            const uint32_t tex = __float_as_uint(t2.z);
            if (!a.tex_mask || ((a.tex_mask[tex >> 5] >> (tex & 31)) & 1u)) {
                if (a.rep) cm = a.coarse_map[tex * RH_NUM_MIP_LEVELS + __float_as_int(t2.w)];
                const float4 t0 = rt[0], t1 = rt[a.lanes];
                const float raw[3] = {t1.z, t1.w, t2.x};
                const float gch[3] = {g.x, g.y, g.z};
                const bool srgb = t2.y != 0.0f;
                const float kx = t1.x, ky = t1.y;
                w[0] = (1.0f - kx) * (1.0f - ky);
                w[1] = kx * (1.0f - ky);
                w[2] = (1.0f - kx) * ky;
                w[3] = kx * ky;
                idx[0] = __float_as_uint(t0.x);
                idx[1] = __float_as_uint(t0.y);
                idx[2] = __float_as_uint(t0.z);
                idx[3] = __float_as_uint(t0.w);
#pragma unroll
                for (int ch = 0; ch < 3; ch++) {
                    float lin, dlin;
                    if (!srgb) {
                        lin = raw[ch];
                        dlin = 1.0f;
                    } else if (raw[ch] > 0.04045f) {
                        const float b = (raw[ch] + 0.055f) / 1.055f;
                        lin = rpow(b, 2.4f);
                        dlin = 2.4f / 1.055f * rpow(b, 1.4f);
                    } else {
                        lin = raw[ch] / 12.92f;
                        dlin = 1.0f / 12.92f;
                    }
                    if (lin != 0.0f && gch[ch] != 0.0f) gr[ch] = a.scale * (gch[ch] / lin) * dlin;
                }
            }
        }
        const uint32_t copy = (blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) % kTexelReplicas;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const uint32_t key = w[q] != 0.0f ? idx[q] : kNoKey;
            float *target = cm.y != kNoKey && key != kNoKey
                                ? a.rep + 4 * ((size_t)copy * a.rep_count + cm.y + (idx[q] - cm.x))
                                : a.tex_grad + 4 * (size_t)(key == kNoKey ? 0u : key);
            wave_add3_at(target, key, f3{gr[0] * w[q], gr[1] * w[q], gr[2] * w[q]});
        }
    }
}

// Path replay for one pass: each lane regenerates its pixel's primary ray with
// the pass's iteration and Halton table, follows the same path as the forward
// pass (trace, ShadeSurface with visibility traced inline, next ray) and
// records each vertex's colour C_j and the material whose main_color scales it.
// A vertex's main_color m multiplies C_j for every j >= its own index exactly
// once (the throughput carries it; diffuse NEE and emission carry it at the
// vertex itself), so with S_m = sum_{j >= m} C_j:
//     d(pixel) / d main_color[c] = S_m[c] / main_color[c].
// The suffix sums are formed backwards over the records; the gradient adds
// adj[pixel] * scale * S_m / main_color with float atomics (their order is the
// only run-to-run difference).  A scene of at most kLdsGradMaterials materials
// sums the block's material gradients in LDS first: with few materials every
// lane would otherwise add to the same few addresses.
template <bool Defer, int MinBlocks, int Tree>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_adjoint(AdjointArgs a) {
    extern __shared__ uint32_t lds_stack[];
    __shared__ float s_grad[4 * kLdsGradMaterials];
    const bool lds_grad = a.n_materials <= kLdsGradMaterials;
    if (lds_grad) {
        for (uint32_t j = threadIdx.x; j < 4 * a.n_materials; j += blockDim.x) s_grad[j] = 0.0f;
        __syncthreads();
    }
    const int rw = a.pc.rect[2], rh_ = a.pc.rect[3];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = (int)i < rw * rh_;
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    uint32_t traced = 0;
    Stack st;
    st.init(lds_stack, a.spill + i, a.spill_stride, a.lds_depth, kSpillDepth);
    int nv = 0;
    float4 dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (active) {
        const int x = a.pc.rect[0] + (int)i % rw, y = a.pc.rect[1] + (int)i / rw;
        const int p = y * a.pc.w + x;
        dl = a.adj[p];
        if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) {
            RayFull ray = gen_primary_ray(a.pc, a.halton, x, y);
            rh_pass_info pi = a.pi;
            pi.index = p;
            pi.rand_index = p;
            if (pi.settings.flags & RH_USE_COHERENT_SAMPLING) pi.rand_index = c_sampling_pattern[(y % 8) * 8 + (x % 8)];
            const int last = pi.settings.max_total_depth;
            const bool skip_indirect = (pi.settings.flags & RH_SKIP_INDIRECT_LIGHT) != 0;
            for (int b = -1;;) {
                Hit hit;
                hit.mask = 0;
                hit.obj = -1;
                hit.prim = -1;
                hit.t = RH_MAX_DIST;
                hit.u = hit.v = 0.0f;
                st.sp = 0;
                if (a.sc.macro_root != RH_INVALID) traverse<false, Tree>(a.sc, ray.o, ray.d, hit, st, nullptr);
                traced++;
                pi.bounce = b < 0 ? 2 : b + 3;
                ShadeOut so;
                so.emit = false;
                NeeSink sink;
                if (Defer) sink = NeeSink{a.nee_ray, a.nee_aux, a.nee_term, i * a.nee_slots, 0};
                shade_surface<Tree>(a.sc, pi, hit, ray, a.halton, st, cnt, so, Defer ? &sink : nullptr);
                f3 cj = f3{so.col.x, so.col.y, so.col.z};
                if (Defer && so.nee_pending && sink.count > 0)
                    cj = cj + resolve_light_samples<Tree>(a.sc, a.halton, sink, so, ray.c, st, cnt);
                write_vertex_record(AdjRecords{a.rec, a.tex_grad ? a.rec_tex : nullptr, nullptr, a.lanes, 0, {0, 0, 0}},
                                    i, nv, cj, so);
                nv++;
                b++;
                if (!so.emit || skip_indirect || b >= last) break;
                ray = so.next;
            }
        }
    }
    adjoint_accumulate(a, i, nv, dl, lds_grad, s_grad);
    if (lds_grad) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < 4 * a.n_materials; j += blockDim.x)
            if (s_grad[j] != 0.0f) atomicAdd(&a.grad[j], s_grad[j]);
    }
    if (a.ray_counter) wave_add_u64(a.ray_counter, traced);
    if (a.shadow_counter) wave_add_u64(a.shadow_counter, cnt.shadow_rays);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
}

// The backward half alone, after a staged replay (the forward stage kernels
// wrote the vertex records and per-lane vertex counts).
__global__ void __launch_bounds__(kBlock) k_adjoint_backward(AdjointArgs a, const int *nverts) {
    __shared__ float s_grad[4 * kLdsGradMaterials];
    const bool lds_grad = a.n_materials <= kLdsGradMaterials;
    if (lds_grad) {
        for (uint32_t j = threadIdx.x; j < 4 * a.n_materials; j += blockDim.x) s_grad[j] = 0.0f;
        __syncthreads();
    }
    const int rw = a.pc.rect[2], rh_ = a.pc.rect[3];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    int nv = 0;
    float4 dl = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if ((int)i < rw * rh_) {
        const int x = a.pc.rect[0] + (int)i % rw, y = a.pc.rect[1] + (int)i / rw;
        dl = a.adj[y * a.pc.w + x];
        if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) nv = nverts[i];
    }
    adjoint_accumulate(a, i, nv, dl, lds_grad, s_grad);
    if (lds_grad) {
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < 4 * a.n_materials; j += blockDim.x)
            if (s_grad[j] != 0.0f) atomicAdd(&a.grad[j], s_grad[j]);
    }
}

// Sums the replicas of the small texel blocks into the texel gradient and
// clears them for the next pass (one lane per replicated texel).
__global__ void __launch_bounds__(256) k_fold_replicas(float *__restrict__ rep, uint32_t count,
                                                       const uint32_t *__restrict__ texel, float *__restrict__ tex_grad) {
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= count) return;
    float s[3] = {0.0f, 0.0f, 0.0f};
    for (uint32_t r = 0; r < kTexelReplicas; r++) {
        float *p = rep + 4 * ((size_t)r * count + j);
        for (int c = 0; c < 3; c++) {
            s[c] += p[c];
            p[c] = 0.0f;
        }
    }
    float *g = tex_grad + 4 * (size_t)texel[j];
    for (int c = 0; c < 3; c++) g[c] += s[c];
}

// main_color of materials [first, first + n) from packed rgb triples
__global__ void k_set_material_colors(rh_material *mats, const float *__restrict__ rgb, uint32_t first, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    rh_material &m = mats[first + k];
    m.main_color[0] = rgb[3 * k + 0];
    m.main_color[1] = rgb[3 * k + 1];
    m.main_color[2] = rgb[3 * k + 2];
}

// Solid: for scenes whose triangles are all solid (compute_visibility)
template <int Tree, int MinBlocks = 1, bool Count = false, bool Solid = false>
__global__ void __launch_bounds__(kBlock, MinBlocks) k_shadow(ShadowArgs a) {
    extern __shared__ uint32_t lds_stack[];
    shadow_body<Tree, Count, Solid>(a, blockIdx.x * blockDim.x, lds_stack);
}

__global__ void __launch_bounds__(kBlock) k_nee_combine(CombineArgs a) { combine_body(a, blockIdx.x * blockDim.x); }

__global__ void k_mix_tonemap(const float4 *__restrict__ temp, float4 *__restrict__ clean, float4 *__restrict__ final_,
                              int x0, int y0, int rw, int rh_, int w, float k, int dtype, float gamma, uint32_t flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rw * rh_) return;
    const int x = x0 + i % rw, y = y0 + i / rw;
    mix_tonemap_pixel(temp, clean, final_, (size_t)y * w + x, k, dtype, gamma, flags);
}

__global__ void k_fill(float4 *buf, size_t n, float4 v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) buf[i] = v;
}

__global__ void k_set_int(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}


// ------------------------------------------------------------------ sorting
// Keys for the queue (get_ray_hash) and the identity permutation.
__global__ void __launch_bounds__(kBlock) k_ray_keys(const rh_ray *__restrict__ rays, const int *__restrict__ count,
                                                     int n_static, PassConst pc, uint32_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals) {
    const int n = count ? *count : n_static;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 *p = reinterpret_cast<const float4 *>(rays + i);
    const float4 a = p[0], b = p[1];
    keys[i] = ray_hash(f3{a.x, a.y, a.z}, f3{b.x, b.y, b.z}, pc.root_min, pc.cell_size);
    vals[i] = (uint32_t)i;
}

} // namespace

// ====================================================================== host

// The working set of one pass in flight: its stream, the pass's framebuffer,
// the ray queues, hits, sort scratch, traversal spill slices, the deferred
// NEE buffers and the queue counters.  Up to kMaxPassSlots passes may be in
// flight (option pass_overlap: passes in flight besides the newest): the
// context's fields of these names are the bound slot's, and store[] holds
// the other slots' (bind_slot moves them).
constexpr int kMaxPassSlots = 4;
struct PassSlot {
    hipStream_t stream = nullptr;
    float4 *temp = nullptr;
    size_t temp_cap = 0;
    size_t max_rays = 0;
    rh_ray *rays[2] = {nullptr, nullptr};
    rh_ray_aux *aux[2] = {nullptr, nullptr};
    rh_hit *hits = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;
    uint32_t *spill = nullptr;
    size_t spill_cap = 0, spill_words = 0;
    size_t nee_cap = 0;
    float4 *nee_ray = nullptr, *nee_term = nullptr;
    int2 *nee_aux = nullptr;
    uint8_t *nee_vis = nullptr;
    float4 *pend_c = nullptr, *pend_alb = nullptr;
    int *pend_count = nullptr;
    uint32_t *shadow_queue = nullptr;
    int *counters = nullptr;
};
#define RH_SLOT_FIELDS(X)                                                                                               \
    X(stream) X(temp) X(temp_cap) X(max_rays) X(rays) X(aux) X(hits) X(keys) X(vals) X(hist) X(spill) X(spill_cap)     \
        X(spill_words) X(nee_cap) X(nee_ray) X(nee_term) X(nee_aux) X(nee_vis) X(pend_c) X(pend_alb) X(pend_count)     \
            X(shadow_queue) X(counters)

struct rh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int w = 0, h = 0;
    size_t max_rays = 0;
    // passes in flight: the other slots (store[bound] is not used: the bound
    // slot lives in the fields above and below), which slot is bound, and the
    // order between them (see begin_overlapped_pass)
    PassSlot store[kMaxPassSlots];
    int bound = 0;
    int pass_overlap = kMaxPassSlots - 1; // option: plain passes rotate over pass_overlap + 1 slots
    bool dirty = true;    // device work other than an overlapped pass was queued since the last pass
    hipEvent_t ev_mix = nullptr;  // the last pass's mix (passes mix into clean in pass order)
    bool mix_recorded = false;
    hipEvent_t ev_join = nullptr; // orders an overlapped pass after everything queued before it
    size_t temp_cap = 0;          // pixels the bound slot's temp holds

    // scene
    std::vector<DevBuf> scene_bufs;    // the arrays instance transforms leave alone
    std::vector<DevBuf> instance_bufs; // the arrays rh_update_instances replaces
    size_t static_counts[9] = {};      // their sizes at upload (rh_update_instances checks them)
    int n_uploads = 0, n_instance_updates = 0; // options "scene_uploads", "instance_updates" (read-only)
    Scene sc = {};
    bool have_scene = false;

    // framebuffers
    float4 *temp = nullptr, *clean = nullptr, *final_ = nullptr;
    // ray queues (ping-pong), hits, sort scratch
    rh_ray *rays[2] = {nullptr, nullptr};
    rh_ray_aux *aux[2] = {nullptr, nullptr};
    rh_hit *hits = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;
    uint32_t *spill = nullptr;
    size_t spill_cap = 0; // lanes with a spill slice (stride of the spill array)
    size_t spill_words = 0; // words per lane the spill array holds
    // deferred next-event estimation: slots per shaded ray = light count (0: inline)
    uint32_t nee_slots = 0;
    size_t nee_cap = 0;
    float4 *nee_ray = nullptr, *nee_term = nullptr;
    int2 *nee_aux = nullptr;
    uint8_t *nee_vis = nullptr;
    float4 *pend_c = nullptr, *pend_alb = nullptr;
    int *pend_count = nullptr;
    uint32_t *shadow_queue = nullptr; // compacted shadow work: slot indices
    int defer_nee = 1;
    // adjoint replay: per-lane vertex records (max vertices x lanes)
    float4 *adj_rec = nullptr;
    size_t adj_rec_cap = 0; // lanes
    float4 *adj_rec_tex = nullptr; // texture footprints of the vertices (texel gradients)
    size_t adj_rec_tex_cap = 0;
    float *color_stage = nullptr;
    size_t color_stage_cap = 0; // floats
    uint32_t n_texels = 0;
    uint32_t n_materials = 0;
    uint32_t n_mesh_instances = 0;
    int *counters = nullptr; // [0], [1]: ray queue counts, [2]: shadow queue count, [3]: refill fetch
    unsigned long long *stats_dev = nullptr; // [0] primary [1] secondary [2] shadow [3] sorted [4] overflow
    float *halton = nullptr;

    // profiling: event pairs around trace / shade / sort launches, read back on demand
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<size_t, size_t>>> ev_marks; // kind + 8 (stage + 1), (start, stop)
    int mark_stage = -1; // trace stage KernelMark tags a trace launch with (-1: none)
    // Count-mode traversal work: [0..3] trace (nodes, tris, instances, rays),
    // [4..8] shadow (nodes, tris, instances, entries, segments), [9..13] path
    // tail (nodes, tris, instances, segments, shadow segments)
    unsigned long long *visits = nullptr;

    // options
    int kernel_timing = 0;
    int count_visits = 0;
    int sort_rays = 1;
    int sort_bits = 32; // key bits the in-pass ray sort orders by (top bits)
    int path_tail = 2;  // first bounce run by the path-tail kernel (-1: none)
    int tail_occupancy = 2;    // waves per SIMD the path-tail kernel is compiled for (2 or 3)
    int refill = 0;            // closest-hit trace: persistent kernel with lane refill
    int tail_refill = 0;       // path tail: persistent kernel whose lanes take a new path when theirs ends
    int tiled_primary = 1;     // primary rays queued in 8x8 pixel tiles (0: row-major)
    std::vector<std::pair<std::pair<const void *, size_t>, int>> resident; // (kernel, LDS bytes) -> blocks per CU
    int trace_occupancy = 0;   // 8-wide trace/shadow kernels: 0 = as compiled (4 waves/SIMD), 5 or 6 = capped registers
    int shadow_occupancy = -1; // the shadow kernel's own setting (-1: as trace_occupancy)
    int trace_group = 0;       // 8-wide trees: lanes per ray (k_trace_group) of rh_trace and late staged bounces; 0 = one
    int max_bounces = RH_MAX_BOUNCES; // max_total_depth clamp (the reference's 10; up to RH_MAX_BOUNCES_EXT)
    int adjoint_staged = 1;           // adjoint replay as staged wavefront passes (0: one lane per path)
    uint32_t *geo_winner = nullptr;   // geometry camera: winning triangle per lane
    size_t geo_winner_cap = 0;
    float *sh_temp = nullptr, *sh_clean = nullptr; // OutputSH: 12 floats per pixel, allocated on demand
    std::vector<rh_tex_mip> host_tex_mips;  // the scene's mip table (small texel blocks are replicated)
    uint2 *coarse_map = nullptr;            // per (texture, lod): block offset, replica slot or kNoKey
    uint32_t *coarse_texel = nullptr;       // per replica slot: its texel
    float *rep = nullptr;                   // kTexelReplicas copies of the slots' gradients
    uint32_t rep_count = 0;
    bool coarse_valid = false;
    // the adjoint tape: passes 1..tape_passes after a clear keep their vertex
    // records (slot = iteration - 1) so that the adjoint needs no replay
    int tape_passes = 0, tape_texels = 0;
    size_t tape_lanes = 0, tape_vmax = 0;
    int tape_texel_slots = 0;
    float4 *tape_rec = nullptr, *tape_rec_tex = nullptr;
    int *tape_nverts = nullptr;
    std::vector<std::array<int, 4>> tape_rect; // recorded rect per slot (w = 0: empty)
    std::vector<uint64_t> tape_key;            // pass_key of the pass recorded in each slot
    uint32_t n_textures = 0;
    uint32_t *grad_tex_mask = nullptr; // textures whose texels get gradients (null: all)
    size_t grad_tex_words = 0;
    int *adj_nverts = nullptr;        // staged replay: vertices recorded per lane
    size_t adj_nverts_cap = 0;
    int cus = 0;               // compute units of the device
    int adjoint_occupancy = 2; // the same for the adjoint replay kernel
    int lds_depth = 24;
    // stage timers (RendererBase::stats_t, in the order of rh_stats): every
    // pass records events at its stage boundaries, each interval tagged with
    // the timer it adds to (-1: none); a pass's intervals are summed once its
    // last event has completed (at a later pass without waiting, or in
    // rh_get_stats), so no pass waits for the device
    int stage_timing = 1;
    std::deque<std::vector<std::pair<hipEvent_t, int>>> stage_pending;
    std::vector<hipEvent_t> stage_free;
    double stage_us[6] = {0, 0, 0, 0, 0, 0};
};

namespace {

// Makes slot j the bound one: its working set into the context's fields,
// the bound one's back to the store.
void bind_slot(rh_context *c, int j) {
    if (j == c->bound) return;
#define RH_BIND_FIELD(f)                                                                                                \
    std::swap(c->f, c->store[j].f);                                                                                    \
    std::swap(c->store[j].f, c->store[c->bound].f);
    RH_SLOT_FIELDS(RH_BIND_FIELD)
#undef RH_BIND_FIELD
    c->bound = j;
}

// Waits for every slot's stream.
int sync_slots(rh_context *c) {
    RH_CHECK(hipStreamSynchronize(c->stream));
    for (int j = 0; j < kMaxPassSlots; j++)
        if (j != c->bound && c->store[j].stream) RH_CHECK(hipStreamSynchronize(c->store[j].stream));
    return 0;
}

// Frees the bound slot's frame-sized buffers (its queue counters and stream
// stay); the next pass on the slot allocates them again.
void free_slot_buffers(rh_context *c) {
    release(c->temp);
    for (int k = 0; k < 2; k++) {
        release(c->rays[k]);
        release(c->aux[k]);
        release(c->keys[k]);
        release(c->vals[k]);
    }
    release(c->hits);
    release(c->hist);
    release(c->spill);
    release(c->nee_ray);
    release(c->nee_term);
    release(c->nee_aux);
    release(c->nee_vis);
    release(c->pend_c);
    release(c->pend_alb);
    release(c->pend_count);
    release(c->shadow_queue);
    c->temp_cap = c->max_rays = c->spill_cap = c->spill_words = c->nee_cap = 0;
}

int ensure_rays(rh_context *c, size_t n);

int alloc_frame(rh_context *c, int w, int h) {
    const size_t np = (size_t)w * h;
    int rc = sync_slots(c);
    if (rc) return rc;
    // spherical-harmonic data starts again (zeroed) at the next OutputSH pass
    release(c->sh_temp);
    release(c->sh_clean);
    release(c->clean);
    release(c->final_);
    // every slot's pass buffers are sized for the old frame: the bound slot's
    // are allocated again below, the others' at their next pass
    const int keep = c->bound;
    for (int j = 0; j < kMaxPassSlots; j++) {
        bind_slot(c, j);
        free_slot_buffers(c);
    }
    bind_slot(c, keep);
    c->w = w;
    c->h = h;
    RH_CHECK(hipMalloc(&c->clean, np * sizeof(float4)));
    RH_CHECK(hipMalloc(&c->final_, np * sizeof(float4)));
    RH_CHECK(hipMemsetAsync(c->clean, 0, np * sizeof(float4), c->stream));
    RH_CHECK(hipMemsetAsync(c->final_, 0, np * sizeof(float4), c->stream));
    return ensure_rays(c, std::max<size_t>(np, 1));
}

int ensure_nee(rh_context *c);

// The ray queues for n rays.  Capacities are recorded only once every buffer
// is allocated: after a failed allocation the context holds null pointers and
// no capacity, and the next call allocates again.
int ensure_rays(rh_context *c, size_t n) {
    const size_t np = (size_t)c->w * c->h;
    if (np && (np > c->temp_cap || !c->temp)) {
        // the bound slot's pass framebuffer (its last pass may still read the old one)
        RH_CHECK(hipStreamSynchronize(c->stream));
        release(c->temp);
        c->temp_cap = 0;
        RH_CHECK(hipMalloc(&c->temp, np * sizeof(float4)));
        RH_CHECK(hipMemsetAsync(c->temp, 0, np * sizeof(float4), c->stream));
        c->temp_cap = np;
    }
    if (n <= c->max_rays && c->rays[0]) return ensure_nee(c);
    const size_t m = std::max(n, c->max_rays);
    RH_CHECK(hipStreamSynchronize(c->stream)); // the slot's last pass may still use its queues
    c->max_rays = 0;
    for (int k = 0; k < 2; k++) {
        release(c->rays[k]);
        release(c->aux[k]);
        release(c->keys[k]);
        release(c->vals[k]);
    }
    release(c->hits);
    release(c->hist);
    for (int k = 0; k < 2; k++) {
        RH_CHECK(hipMalloc(&c->rays[k], m * sizeof(rh_ray)));
        RH_CHECK(hipMalloc(&c->aux[k], m * sizeof(rh_ray_aux)));
        RH_CHECK(hipMalloc(&c->keys[k], m * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&c->vals[k], m * sizeof(uint32_t)));
    }
    RH_CHECK(hipMalloc(&c->hits, m * sizeof(rh_hit)));
    // the ray sort's digit counts, tile tickets and per-tile status words
    RH_CHECK(hipMalloc(&c->hist, sort_scratch_words(m) * sizeof(uint32_t)));
    c->max_rays = m;
    return ensure_nee(c);
}

// Spill slices are indexed by lane (ray or shadow slot): the array holds
// kSpillDepth entries for each of spill_cap lanes, stride spill_cap, and for
// a scene with 8-wide trees as many entry distances after them.
size_t spill_words(const rh_context *c) { return (size_t)kSpillDepth * (c->sc.wnodes ? 2 : 1); }

int ensure_spill(rh_context *c, size_t lanes) {
    lanes = (lanes + kBlock - 1) / kBlock * kBlock; // whole blocks: persistent lanes index slices too
    if (lanes <= c->spill_cap && c->spill && c->spill_words >= spill_words(c)) return 0;
    RH_CHECK(hipStreamSynchronize(c->stream)); // the slot's last pass may still use its slices
    release(c->spill);
    const size_t cap = std::max(lanes, c->spill_cap), words = spill_words(c);
    c->spill_cap = 0;
    c->spill_words = 0;
    RH_CHECK(hipMalloc(&c->spill, words * cap * sizeof(uint32_t)));
    c->spill_cap = cap;
    c->spill_words = words;
    return 0;
}

// Buffers of the deferred NEE stage for the current scene and frame size.
int ensure_nee(rh_context *c) {
    const size_t m = c->max_rays;
    if (!c->nee_slots) return ensure_spill(c, m);
    const size_t cap = m * c->nee_slots;
    if (cap > c->nee_cap || !c->nee_ray || !c->shadow_queue) {
        RH_CHECK(hipStreamSynchronize(c->stream)); // the slot's last pass may still use them
        release(c->nee_ray);
        release(c->nee_term);
        release(c->nee_aux);
        release(c->nee_vis);
        release(c->pend_c);
        release(c->pend_alb);
        release(c->pend_count);
        release(c->shadow_queue);
        c->nee_cap = 0;
        RH_CHECK(hipMalloc(&c->nee_ray, cap * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->nee_term, cap * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->nee_aux, cap * sizeof(int2)));
        RH_CHECK(hipMalloc(&c->nee_vis, cap * sizeof(uint8_t)));
        RH_CHECK(hipMalloc(&c->pend_c, m * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->pend_alb, m * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->pend_count, m * sizeof(int)));
        RH_CHECK(hipMalloc(&c->shadow_queue, cap * sizeof(uint32_t)));
        c->nee_cap = cap;
    }
    return ensure_spill(c, std::max(m, cap));
}


// the LDS part of the traversal stacks: entries, and their distances for 8-wide trees
size_t lds_bytes(const rh_context *c) {
    return (size_t)c->lds_depth * kBlock * sizeof(uint32_t) * (c->sc.wnodes ? 2 : 1);
}

TraceArgs trace_args(rh_context *c, const rh_ray *rays, const int *count, int n, rh_hit *hits, const float *tmax,
                     unsigned long long *ray_counter) {
    TraceArgs a;
    a.sc = c->sc;
    a.rays = rays;
    a.count = count;
    a.n = n;
    a.tmax = tmax;
    a.hits = hits;
    a.spill = c->spill;
    a.spill_stride = (uint32_t)c->spill_cap;
    a.lds_depth = (uint32_t)c->lds_depth;
    a.ray_counter = ray_counter;
    a.overflow = c->stats_dev + 4;
    a.visits = c->visits;
    return a;
}

size_t ev_next(rh_context *c) {
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return (size_t)-1;
        c->ev_pool.push_back(e);
    }
    return c->ev_used++;
}

// kind: 0 trace, 1 shade, 2 sort
struct KernelMark {
    rh_context *c;
    int kind;
    size_t a = (size_t)-1;
    KernelMark(rh_context *ctx, int k) : c(ctx), kind(k) {
        if (c->kernel_timing) {
            a = ev_next(c);
            if (a != (size_t)-1) (void)hipEventRecord(c->ev_pool[a], c->stream);
        }
    }
    ~KernelMark() {
        if (c->kernel_timing && a != (size_t)-1) {
            const size_t b = ev_next(c);
            if (b == (size_t)-1) return;
            (void)hipEventRecord(c->ev_pool[b], c->stream);
            c->ev_marks.push_back({kind + (kind == 0 ? 8 * (c->mark_stage + 1) : 0), {a, b}});
        }
    }
};

// Blocks of `kern` (kBlock threads, `lds` bytes of dynamic LDS) resident on
// one CU at once, from the occupancy calculator, remembered per (kernel, LDS).
int resident_blocks(rh_context *c, const void *kern, size_t lds, int *out) {
    for (const auto &r : c->resident)
        if (r.first.first == kern && r.first.second == lds) {
            *out = r.second;
            return 0;
        }
    int per_cu = 0;
    RH_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, lds));
    per_cu = std::max(1, per_cu);
    c->resident.push_back({{kern, lds}, per_cu});
    *out = per_cu;
    return 0;
}

// Staged bounces from this trace stage on (bounce 2) may be walked by lane
// groups (option trace_group): their queues hold few rays, which one lane per
// ray leaves too few to keep the device busy.
constexpr int kGroupFromStage = 3;

// stage: the pass's trace stage (0 primary rays, b + 1 bounce b), or -1 for a
// trace of the caller's rays (rh_trace)
int launch_trace(rh_context *c, const TraceArgs &a, int grid_rays, bool any_hit, int stage = -1) {
    const int blocks = std::max(1, (grid_rays + kBlock - 1) / kBlock);
    const bool count = c->count_visits && a.visits;
    if (!any_hit && c->refill) {
        // persistent grid: as many blocks as fit on the device at once (and no
        // more than the rays need); spill slices exist for every lane of them
        const bool wide = c->sc.wnodes != nullptr;
        const int occ = wide ? c->trace_occupancy : 0;
        const void *kern = !wide      ? (const void *)k_trace_refill<false>
                           : occ == 5 ? (const void *)k_trace_refill_wide<false, 5>
                           : occ == 6 ? (const void *)k_trace_refill_wide<false, 6>
                                      : (const void *)k_trace_refill_wide<false, 1>;
        int per_cu = 0;
        if (const int rc = resident_blocks(c, kern, lds_bytes(c), &per_cu)) return rc;
        const int grid = std::min(blocks, per_cu * c->cus);
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->counters + 3, 0);
        KernelMark km(c, 0);
        const dim3 g(grid), blk(kBlock);
        const size_t lds = lds_bytes(c);
        int *next = c->counters + 3;
        if (!wide) {
            if (count) hipLaunchKernelGGL((k_trace_refill<true>), g, blk, lds, c->stream, a, next);
            else hipLaunchKernelGGL((k_trace_refill<false>), g, blk, lds, c->stream, a, next);
        } else if (count) {
            // the instrumented walk visits what the plain one does; its grid
            // is sized by its own occupancy
            int pc = 0;
            if (const int rc = resident_blocks(c, (const void *)k_trace_refill_wide<true, 1>, lds, &pc)) return rc;
            hipLaunchKernelGGL((k_trace_refill_wide<true, 1>), dim3(std::min(blocks, pc * c->cus)), blk, lds,
                               c->stream, a, next);
        } else if (occ == 5) {
            hipLaunchKernelGGL((k_trace_refill_wide<false, 5>), g, blk, lds, c->stream, a, next);
        } else if (occ == 6) {
            hipLaunchKernelGGL((k_trace_refill_wide<false, 6>), g, blk, lds, c->stream, a, next);
        } else {
            hipLaunchKernelGGL((k_trace_refill_wide<false, 1>), g, blk, lds, c->stream, a, next);
        }
        RH_CHECK(hipGetLastError());
        return 0;
    }
    KernelMark km(c, 0);
    const dim3 grid(blocks), blk(kBlock);
    const size_t lds = lds_bytes(c);
    if (c->sc.wnodes && c->trace_group && (stage < 0 || stage >= kGroupFromStage)) {
        // G lanes per ray: G times the threads, one LDS stack per ray
        const int g = c->trace_group;
        const dim3 gg((unsigned)std::max<size_t>(1, ((size_t)grid_rays * g + kBlock - 1) / kBlock));
        const size_t glds = (size_t)c->lds_depth * (kBlock / g) * sizeof(uint32_t) * 2;
#define RH_TRACE_GROUP(G)                                                                                              \
    if (any_hit) {                                                                                                     \
        if (count) hipLaunchKernelGGL((k_trace_group<true, true, G, 1>), gg, blk, glds, c->stream, a);                \
        else hipLaunchKernelGGL((k_trace_group<true, false, G, 1>), gg, blk, glds, c->stream, a);                     \
    } else {                                                                                                           \
        if (count) hipLaunchKernelGGL((k_trace_group<false, true, G, 1>), gg, blk, glds, c->stream, a);               \
        else hipLaunchKernelGGL((k_trace_group<false, false, G, 1>), gg, blk, glds, c->stream, a);                    \
    }
        if (g == 2) {
            RH_TRACE_GROUP(2)
        } else if (g == 4) {
            RH_TRACE_GROUP(4)
        } else {
            RH_TRACE_GROUP(8)
        }
#undef RH_TRACE_GROUP
    } else if (c->sc.wnodes && c->trace_occupancy == 5) {
        if (any_hit) {
            if (count) hipLaunchKernelGGL((k_trace<true, true, kTreeWide, 5>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<true, false, kTreeWide, 5>), grid, blk, lds, c->stream, a);
        } else {
            if (count) hipLaunchKernelGGL((k_trace<false, true, kTreeWide, 5>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<false, false, kTreeWide, 5>), grid, blk, lds, c->stream, a);
        }
    } else if (c->sc.wnodes && c->trace_occupancy == 6) {
        if (any_hit) {
            if (count) hipLaunchKernelGGL((k_trace<true, true, kTreeWide, 6>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<true, false, kTreeWide, 6>), grid, blk, lds, c->stream, a);
        } else {
            if (count) hipLaunchKernelGGL((k_trace<false, true, kTreeWide, 6>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<false, false, kTreeWide, 6>), grid, blk, lds, c->stream, a);
        }
    } else if (c->sc.wnodes) {
        if (any_hit) {
            if (count) hipLaunchKernelGGL((k_trace<true, true, kTreeWide>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<true, false, kTreeWide>), grid, blk, lds, c->stream, a);
        } else {
            if (count) hipLaunchKernelGGL((k_trace<false, true, kTreeWide>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<false, false, kTreeWide>), grid, blk, lds, c->stream, a);
        }
    } else {
        if (any_hit) {
            if (count) hipLaunchKernelGGL((k_trace<true, true, kTreeBinary>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<true, false, kTreeBinary>), grid, blk, lds, c->stream, a);
        } else {
            if (count) hipLaunchKernelGGL((k_trace<false, true, kTreeBinary>), grid, blk, lds, c->stream, a);
            else hipLaunchKernelGGL((k_trace<false, false, kTreeBinary>), grid, blk, lds, c->stream, a);
        }
    }
    RH_CHECK(hipGetLastError());
    return 0;
}

ShadeArgs shade_args(rh_context *c) {
    ShadeArgs a;
    memset(&a, 0, sizeof(a));
    a.sc = c->sc;
    a.spill = c->spill;
    a.spill_stride = (uint32_t)c->spill_cap;
    a.lds_depth = (uint32_t)c->lds_depth;
    a.shadow_counter = c->stats_dev + 2;
    a.overflow = c->stats_dev + 4;
    if (c->defer_nee && c->nee_slots) {
        a.nee_ray = c->nee_ray;
        a.nee_aux = c->nee_aux;
        a.nee_term = c->nee_term;
        a.nee_slots = c->nee_slots;
        a.pend_c = c->pend_c;
        a.pend_alb = c->pend_alb;
        a.pend_count = c->pend_count;
        a.shadow_queue = c->shadow_queue;
        a.shadow_count = c->counters + 2;
    }
    return a;
}

// shade, then (deferred NEE) visibility of the recorded light samples and
// their combination into the pixels / per-ray colours
int launch_shade(rh_context *c, const ShadeArgs &a, int grid_rays) {
    const int blocks = std::max(1, (grid_rays + kBlock - 1) / kBlock);
    {
        KernelMark km(c, 1);
        if (a.nee_ray) {
            hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, a.shadow_count, 0);
            hipLaunchKernelGGL(k_shade<true>, dim3(blocks), dim3(kBlock), 0, c->stream, a);
        } else {
            hipLaunchKernelGGL(k_shade<false>, dim3(blocks), dim3(kBlock), lds_bytes(c), c->stream, a);
        }
        RH_CHECK(hipGetLastError());
    }
    if (!a.nee_ray) return 0;
    ShadowArgs s;
    s.sc = c->sc;
    s.queue = a.shadow_queue;
    s.count = a.shadow_count;
    s.nee_ray = a.nee_ray;
    s.nee_aux = a.nee_aux;
    s.nee_vis = c->nee_vis;
    s.halton = a.halton;
    s.spill = c->spill;
    s.spill_stride = (uint32_t)c->spill_cap;
    s.lds_depth = (uint32_t)c->lds_depth;
    s.shadow_counter = a.shadow_counter;
    s.overflow = c->stats_dev + 4;
    const bool count = c->count_visits && c->visits;
    s.visits = count ? c->visits + 4 : nullptr;
    const size_t lanes = (size_t)grid_rays * a.nee_slots;
    {
        KernelMark km(c, 3);
        const dim3 grid((unsigned)std::max<size_t>(1, (lanes + kBlock - 1) / kBlock)), blk(kBlock);
        const int occ = c->shadow_occupancy >= 0 ? c->shadow_occupancy : c->trace_occupancy;
        if (count && c->sc.wnodes)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 1, true>), grid, blk, lds_bytes(c), c->stream, s);
        else if (count)
            hipLaunchKernelGGL((k_shadow<kTreeBinary, 1, true>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes && c->sc.all_solid && occ == 5)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 5, false, true>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes && c->sc.all_solid && occ == 6)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 6, false, true>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes && c->sc.all_solid)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 1, false, true>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes && occ == 5)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 5>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes && occ == 6)
            hipLaunchKernelGGL((k_shadow<kTreeWide, 6>), grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.wnodes) hipLaunchKernelGGL(k_shadow<kTreeWide>, grid, blk, lds_bytes(c), c->stream, s);
        else if (c->sc.all_solid)
            hipLaunchKernelGGL((k_shadow<kTreeBinary, 1, false, true>), grid, blk, lds_bytes(c), c->stream, s);
        else hipLaunchKernelGGL(k_shadow<kTreeBinary>, grid, blk, lds_bytes(c), c->stream, s);
        RH_CHECK(hipGetLastError());
    }
    CombineArgs cb;
    cb.count = a.count;
    cb.n = a.n;
    cb.slots = a.nee_slots;
    cb.pend_count = a.pend_count;
    cb.pend_c = a.pend_c;
    cb.pend_alb = a.pend_alb;
    cb.nee_term = a.nee_term;
    cb.nee_vis = c->nee_vis;
    cb.temp = a.temp;
    cb.out_rgba = a.out_rgba;
    cb.rec = a.rec;
    cb.frame_w = a.w;
    hipLaunchKernelGGL(k_nee_combine, dim3(blocks), dim3(kBlock), 0, c->stream, cb);
    RH_CHECK(hipGetLastError());
    return 0;
}

// Sorts the queue rays_in/aux_in (`count` rays on the device, at most max_n)
// by the top `bits` bits (8, 16, 24 or 32) of their hash into
// rays_out/aux_out, stable, on `stream`: k_sort_clear, k_sort_keys, then
// one k_sort_scatter per 8-bit digit, the last moving the rays and writing
// the count to count_out (when given).  keys/vals
// are two ping-pong buffers of max_n words each; scr holds
// sort_scratch_words(max_n) words.
int sort_queue(hipStream_t stream, const rh_ray *rays_in, const rh_ray_aux *aux_in, rh_ray *rays_out,
               rh_ray_aux *aux_out, const int *count, int *count_out, int max_n, int bits, const PassConst &pc,
               uint32_t *const keys[2], uint32_t *const vals[2], uint32_t *scr, unsigned long long *stalls) {
    if (bits != 8 && bits != 16 && bits != 24 && bits != 32) return arg_error("sort bits must be 8, 16, 24 or 32");
    if ((long long)max_n > kSortMaxRays) return arg_error("ray queue too long for the sort (2^30 rays)");
    const int tiles = (int)std::max<size_t>(1, sort_tiles((size_t)max_n));
    const int npass = bits / 8, shift0 = 32 - bits;
    const size_t clear_words = kSortStatus + (size_t)npass * tiles * 256;
    const unsigned clear_blocks = (unsigned)std::min<size_t>(1024, (clear_words + 255) / 256);
    hipLaunchKernelGGL(k_sort_clear, dim3(clear_blocks), dim3(256), 0, stream, scr, count, max_n, npass, tiles);
    hipLaunchKernelGGL(k_sort_keys, dim3(tiles), dim3(kBlock), 0, stream, rays_in, count, max_n, pc, keys[0], vals[0],
                       scr, shift0, npass);
    int src = 0;
    for (int p = 0; p < npass; p++) {
        const int shift = shift0 + 8 * p;
        if (p + 1 < npass)
            hipLaunchKernelGGL(k_sort_scatter<false>, dim3(tiles), dim3(kBlock), 0, stream, keys[src], vals[src],
                               keys[src ^ 1], vals[src ^ 1], count, max_n, shift, p, scr, tiles, nullptr, nullptr,
                               nullptr, nullptr, nullptr, stalls);
        else
            hipLaunchKernelGGL(k_sort_scatter<true>, dim3(tiles), dim3(kBlock), 0, stream, keys[src], vals[src],
                               nullptr, nullptr, count, max_n, shift, p, scr, tiles, rays_in, aux_in, rays_out,
                               aux_out, count_out, stalls);
        src ^= 1;
    }
    RH_CHECK(hipGetLastError());
    return 0;
}

PassConst make_pass_const(rh_context *c, const rh_pass_desc *p) { return make_pass_const(c->w, c->h, p); }

// the Count-mode counters of rh_context::visits: kVisitWords words, the pass
// trace launches' per stage from kStageVisits (4 per stage)
constexpr int kStageVisits = 16;
constexpr int kVisitWords = kStageVisits + 4 * RH_TRACE_STAGES;

// stats_t timers a pass's intervals add to
enum StageTimerIndex : int {
    kNoTimer = -1,
    kRayGen = 0,
    kPrimaryTrace = 1,
    kPrimaryShade = 2,
    kSecondarySort = 3,
    kSecondaryTrace = 4,
    kSecondaryShade = 5
};

// Sums the intervals of finished passes into the stage timers; with wait, of
// every recorded pass.
int stage_harvest(rh_context *c, bool wait) {
    while (!c->stage_pending.empty()) {
        auto &marks = c->stage_pending.front();
        const hipEvent_t last = marks.back().first;
        if (wait) {
            RH_CHECK(hipEventSynchronize(last));
        } else {
            const hipError_t q = hipEventQuery(last);
            if (q == hipErrorNotReady) break;
            RH_CHECK(q);
        }
        for (size_t k = 1; k < marks.size(); k++) {
            float ms = 0.0f;
            RH_CHECK(hipEventElapsedTime(&ms, marks[k - 1].first, marks[k].first));
            if (marks[k].second >= 0) c->stage_us[marks[k].second] += 1000.0 * ms;
        }
        for (auto &m : marks) c->stage_free.push_back(m.first);
        c->stage_pending.pop_front();
    }
    return 0;
}

// The stage boundaries of one pass: mark(t) records an event ending an
// interval that adds to timer t.  commit() hands the pass to the harvest;
// a pass that fails before it gives its events back.
struct PassStages {
    rh_context *c;
    bool on;
    std::vector<std::pair<hipEvent_t, int>> marks;
    PassStages(rh_context *ctx, bool enabled) : c(ctx), on(enabled && ctx->stage_timing) {}
    void mark(int timer) {
        if (!on) return;
        hipEvent_t e = nullptr;
        if (!c->stage_free.empty()) {
            e = c->stage_free.back();
            c->stage_free.pop_back();
        } else if (hipEventCreate(&e) != hipSuccess) {
            on = false; // no timing for this pass
            return;
        }
        if (hipEventRecord(e, c->stream) != hipSuccess) {
            c->stage_free.push_back(e);
            on = false;
            return;
        }
        marks.emplace_back(e, timer);
    }
    void commit() {
        if (on && marks.size() > 1) c->stage_pending.push_back(std::move(marks));
        marks.clear();
    }
    ~PassStages() {
        for (auto &m : marks) c->stage_free.push_back(m.first);
    }
};

} // namespace

// ---- scratch of the allocation-free ray sort (rh_sort_rays_noalloc) ----
// Keys and values twice (the ping-pong buffers of the radix passes), then the
// sort's digit counts, tile tickets and per-tile status words, for at most
// RH_NOALLOC_MAX_RAYS rays: 42 MB on every GPU the library runs on, held until
// the process ends.  Host code never uses the array's name as an address: the
// device address of each GPU's copy comes from hipGetSymbolAddress (scratch_ptr).
constexpr size_t kNoAllocWords = 4 * (size_t)RH_NOALLOC_MAX_RAYS + sort_scratch_words(RH_NOALLOC_MAX_RAYS);
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t *g_scratch[kScratchDevices] = {};

// The current device's copy of `scratch`, looked up once per device.
int scratch_ptr(uint32_t **out) {
    int dev = 0;
    RH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices) return arg_error("device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        RH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<uint32_t *>(p);
    }
    *out = g_scratch[dev];
    return 0;
}
} // namespace

extern "C" {

const char *rh_last_error(void) { return rh_rt::err().c_str(); }

int rh_create(int device, rh_context **out) {
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        rh_rt::err() = "no HIP device available";
        return -1002;
    }
    if (device < 0) {
        RH_CHECK(hipGetDevice(&device));
    }
    if (device >= ndev) return arg_error("device index out of range");
    RH_CHECK(hipSetDevice(device));
    auto *c = new rh_context;
    c->device = device;
    // one stream per pass slot, from the device's shared pool
    static_assert(kMaxPassSlots <= rh_rt::kPassStreams, "a stream per pass slot");
    hipStream_t streams[rh_rt::kPassStreams];
    const int src = rh_rt::pass_streams(device, streams);
    if (src) {
        delete c;
        return src;
    }
    c->stream = streams[0];
    for (int j = 1; j < kMaxPassSlots; j++) c->store[j].stream = streams[j];
    RH_CHECK(hipEventCreateWithFlags(&c->ev_mix, hipEventDisableTiming));
    RH_CHECK(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    hipDeviceProp_t prop;
    RH_CHECK(hipGetDeviceProperties(&prop, device));
    c->cus = prop.multiProcessorCount;
    RH_CHECK(hipMalloc(&c->counters, 16 * sizeof(int)));
    for (int j = 1; j < kMaxPassSlots; j++) {
        RH_CHECK(hipMalloc(&c->store[j].counters, 16 * sizeof(int)));
        RH_CHECK(hipMemsetAsync(c->store[j].counters, 0, 16 * sizeof(int), c->stream));
    }
    RH_CHECK(hipMalloc(&c->stats_dev, 8 * sizeof(unsigned long long)));
    RH_CHECK(hipMemsetAsync(c->counters, 0, 16 * sizeof(int), c->stream));
    RH_CHECK(hipMemsetAsync(c->stats_dev, 0, 8 * sizeof(unsigned long long), c->stream));
    RH_CHECK(hipMalloc(&c->visits, kVisitWords * sizeof(unsigned long long)));
    RH_CHECK(hipMemsetAsync(c->visits, 0, kVisitWords * sizeof(unsigned long long), c->stream));
    RH_CHECK(hipMalloc(&c->halton, RH_HALTON_TABLE_FLOATS * sizeof(float)));
    RH_CHECK(hipMemsetAsync(c->halton, 0, RH_HALTON_TABLE_FLOATS * sizeof(float), c->stream));
    *out = c;
    return 0;
}

void rh_destroy(rh_context *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)sync_slots(c);
    for (auto &b : c->scene_bufs) free_ptr(b.p);
    for (auto &b : c->instance_bufs) free_ptr(b.p);
    free_ptr(c->geo_winner);
    free_ptr(c->sh_temp);
    free_ptr(c->sh_clean);
    free_ptr(c->coarse_map);
    free_ptr(c->coarse_texel);
    free_ptr(c->rep);
    free_ptr(c->tape_rec);
    free_ptr(c->tape_rec_tex);
    free_ptr(c->tape_nverts);
    free_ptr(c->clean);
    free_ptr(c->final_);
    for (int j = 0; j < kMaxPassSlots; j++) {
        bind_slot(c, j);
        free_slot_buffers(c);
        free_ptr(c->counters);
        c->counters = nullptr;
    }
    free_ptr(c->stats_dev);
    free_ptr(c->halton);
    free_ptr(c->visits);
    free_ptr(c->adj_rec);
    free_ptr(c->adj_rec_tex);
    free_ptr(c->adj_nverts);
    free_ptr(c->grad_tex_mask);
    free_ptr(c->color_stage);
    for (auto &e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto &p : c->stage_pending)
        for (auto &m : p) (void)hipEventDestroy(m.first);
    for (auto &e : c->stage_free) (void)hipEventDestroy(e);
    if (c->ev_mix) (void)hipEventDestroy(c->ev_mix);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    delete c; // the streams stay in the device's pool
}

int rh_resize(rh_context *c, int w, int h) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return arg_error("framebuffer size must be 1..65535 per side");
    RH_CHECK(hipSetDevice(c->device));
    int rc = alloc_frame(c, w, h);
    if (rc) return rc;
    return ensure_rays(c, (size_t)w * h);
}

static void tape_forget(rh_context *c);

int rh_clear(rh_context *c, const float rgba[4]) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const size_t n = (size_t)c->w * c->h;
    if (!n) return 0;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->clean, n,
                       make_float4(rgba[0], rgba[1], rgba[2], rgba[3]));
    RH_CHECK(hipGetLastError());
    tape_forget(c);
    // the accumulated SH data too (FramebufferRef.cpp Clear)
    if (c->sh_clean) RH_CHECK(hipMemsetAsync(c->sh_clean, 0, n * 12 * sizeof(float), c->stream));
    return 0;
}

int rh_read_sh(rh_context *c, float *out, size_t count) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const size_t n = (size_t)c->w * c->h * 12;
    if (!c->sh_clean) return 1;
    if (count < n) return arg_error("output buffer too small");
    RH_CHECK(hipMemcpyAsync(out, c->sh_clean, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

static void static_counts(const rh_scene_view *v, size_t out[9]) {
    const size_t k[9] = {v->tris_count,      v->tri_indices_count, v->vertices_count,
                         v->vtx_indices_count, v->materials_count,   v->textures_count,
                         v->texels_count,     v->lights_count,      v->li_indices_count};
    memcpy(out, k, sizeof(k));
}

int rh_upload_scene(rh_context *c, const rh_scene_view *v) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    RH_CHECK(hipSetDevice(c->device));
    RH_CHECK(hipStreamSynchronize(c->stream));
    c->have_scene = false;
    for (auto &b : c->scene_bufs) free_ptr(b.p);
    for (auto &b : c->instance_bufs) free_ptr(b.p);
    c->scene_bufs.clear();
    c->instance_bufs.clear();
    Scene s = {};
    int rc = build_device_static(v, c->stream, c->scene_bufs, s);
    if (rc) return rc;
    rc = build_device_instances(v, c->stream, c->instance_bufs, s);
    if (rc) return rc;
    static_counts(v, c->static_counts);
    c->n_uploads++;
    // deferred NEE keeps one slot per light for every shaded ray; scenes with
    // many lights trace visibility inline instead
    c->nee_slots = (v->lights_count >= 1 && v->lights_count <= kMaxDeferredLights) ? v->lights_count : 0;
    c->sc = s;
    c->n_materials = (uint32_t)v->materials_count;
    c->n_mesh_instances = (uint32_t)v->mesh_instances_count;
    c->host_tex_mips.assign(v->tex_mips, v->tex_mips + (size_t)v->textures_count * RH_NUM_MIP_LEVELS);
    c->coarse_valid = false;
    tape_forget(c);
    c->n_texels = (uint32_t)v->texels_count;
    c->n_textures = (uint32_t)v->textures_count;
    c->have_scene = true;
    if (c->max_rays) {
        // spill slices sized for this scene's trees
        const int rc2 = ensure_nee(c);
        if (rc2) return rc2;
    }
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_update_instances(rh_context *c, const rh_scene_view *v) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("rh_update_instances: no scene uploaded");
    size_t k[9];
    static_counts(v, k);
    if (memcmp(k, c->static_counts, sizeof(k)) != 0)
        return arg_error("rh_update_instances: the scene's triangles, materials, textures or lights changed");
    RH_CHECK(hipSetDevice(c->device));
    RH_CHECK(hipStreamSynchronize(c->stream));
    c->have_scene = false; // until the new arrays are in place
    for (auto &b : c->instance_bufs) free_ptr(b.p);
    c->instance_bufs.clear();
    Scene s = c->sc;
    const int rc = build_device_instances(v, c->stream, c->instance_bufs, s);
    if (rc) return rc;
    c->sc = s;
    c->n_mesh_instances = (uint32_t)v->mesh_instances_count;
    tape_forget(c);
    c->have_scene = true;
    c->n_instance_updates++;
    if (c->max_rays) {
        const int rc2 = ensure_nee(c);
        if (rc2) return rc2;
    }
    return 0;
}

int rh_set_halton(rh_context *c, const float *seq, size_t count) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (count < (size_t)RH_HALTON_COUNT * RH_HALTON_SEQ_LEN) return arg_error("Halton table needs 256 x 24 floats");
    RH_CHECK(upload_halton(c->halton, seq, c->stream));
    // the host table may change right after this call returns
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// One pass in wavefront form.  rec == nullptr: RenderScene (pixels, path tail,
// mix and tonemap).  Otherwise the staged adjoint replay of the same pass: the
// stages write vertex records instead of pixels, every bounce is staged, and
// rays are counted as adjoint rays.
// rec: the pass writes vertex records (every bounce staged); keep_image: and
// renders its image as usual (the adjoint tape), else the image is left alone
// and the rays are counted as the adjoint's (the replay)
int run_pass(rh_context *c, const rh_pass_desc *p, const AdjRecords *rec, bool keep_image = false) {
    if (!c->have_scene) return arg_error("no scene uploaded");
    const int rx = p->rect[0], ry = p->rect[1], rw = p->rect[2], rhh = p->rect[3];
    if (rw <= 0 || rhh <= 0 || rx < 0 || ry < 0 || rx + rw > c->w || ry + rhh > c->h)
        return arg_error("region outside the framebuffer");
    if (p->iteration < 1) return arg_error("iteration must be >= 1");
    const int n = rw * rhh;
    int rc = ensure_rays(c, (size_t)n);
    if (rc) return rc;

    const PassConst pc = make_pass_const(c, p);
    rh_pass_info pi;
    memset(&pi, 0, sizeof(pi));
    pi.iteration = p->iteration;
    pi.bounce = 2;
    pi.settings = p->cam.pass_settings;
    pi.settings.max_total_depth = (uint8_t)std::min<int>(pi.settings.max_total_depth, c->max_bounces);

    const bool geo = p->cam.type == RH_CAM_GEO;
    const bool out_sh = !rec && (p->cam.pass_settings.flags & RH_OUTPUT_SH);
    if (geo) {
        if (p->cam.mi_index >= c->n_mesh_instances) return arg_error("geometry camera: no such mesh instance");
        if (p->cam.uv_index > 1) return arg_error("geometry camera: uv layer must be 0 or 1");
        if ((size_t)n > c->geo_winner_cap || !c->geo_winner) {
            release(c->geo_winner);
            c->geo_winner_cap = 0;
            RH_CHECK(hipMalloc(&c->geo_winner, (size_t)n * sizeof(uint32_t)));
            c->geo_winner_cap = (size_t)n;
        }
    }
    if (out_sh && (!c->sh_temp || !c->sh_clean)) {
        const size_t bytes = (size_t)c->w * c->h * 12 * sizeof(float);
        release(c->sh_temp);
        release(c->sh_clean);
        RH_CHECK(hipMalloc(&c->sh_temp, bytes));
        RH_CHECK(hipMalloc(&c->sh_clean, bytes));
        RH_CHECK(hipMemsetAsync(c->sh_temp, 0, bytes, c->stream));
        RH_CHECK(hipMemsetAsync(c->sh_clean, 0, bytes, c->stream));
    }

    const bool replay = rec && !keep_image;
    // finished passes' stage times (no wait), then this pass's boundaries
    rc = stage_harvest(c, false);
    if (rc) return rc;
    PassStages tm(c, !replay);
    tm.mark(kNoTimer);
    int *cnt[2] = {c->counters + 0, c->counters + 1};
    unsigned long long *count_primary = replay ? c->stats_dev + 5 : c->stats_dev + 0;
    unsigned long long *count_secondary = replay ? c->stats_dev + 5 : c->stats_dev + 1;
    if (geo) {
        // rays and hits from the mesh's texture-space footprint, no trace
        GeoArgs g;
        g.vertices = c->sc.vertices;
        g.vtx_indices = c->sc.vtx_indices;
        g.mesh_instances = c->sc.mesh_instances;
        g.meshes = c->sc.meshes;
        g.transforms = c->sc.transforms;
        g.halton = c->halton;
        g.winner = c->geo_winner;
        memcpy(g.rect, p->rect, sizeof(g.rect));
        g.w = c->w;
        g.h = c->h;
        g.iteration = p->iteration;
        g.mi_index = p->cam.mi_index;
        g.uv = p->cam.uv_index;
        const dim3 grid((n + kBlock - 1) / kBlock), blk(kBlock);
        hipLaunchKernelGGL(k_geo_init, grid, blk, 0, c->stream, g, c->rays[0], c->aux[0], c->hits, cnt[0]);
        hipLaunchKernelGGL(k_geo_raster, dim3(2 * c->cus * 4), blk, 0, c->stream, g);
        hipLaunchKernelGGL(k_geo_finish, grid, blk, 0, c->stream, g, c->rays[0], c->aux[0], c->hits);
        RH_CHECK(hipGetLastError());
        tm.mark(kRayGen); // the geometry camera makes its hits too: no primary trace
    } else {
        if (c->tiled_primary)
            hipLaunchKernelGGL(k_gen_primary<true>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, pc,
                               c->halton, c->rays[0], c->aux[0], cnt[0]);
        else
            hipLaunchKernelGGL(k_gen_primary<false>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, pc,
                               c->halton, c->rays[0], c->aux[0], cnt[0]);
        RH_CHECK(hipGetLastError());
        tm.mark(kRayGen);
        TraceArgs ta = trace_args(c, c->rays[0], cnt[0], n, c->hits, nullptr, count_primary);
        ta.visits = c->visits + kStageVisits; // stage 0
        c->mark_stage = 0;
        rc = launch_trace(c, ta, n, false, 0);
        c->mark_stage = -1;
        if (rc) return rc;
        tm.mark(kPrimaryTrace);
    }

    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, cnt[1], 0);
    ShadeArgs sa = shade_args(c);
    sa.rays = c->rays[0];
    sa.aux = c->aux[0];
    sa.hits = c->hits;
    sa.count = cnt[0];
    sa.n = n;
    sa.pi = pi;
    sa.w = c->w;
    sa.halton = c->halton;
    sa.temp = replay ? nullptr : c->temp;
    sa.primary = 1;
    if (rec) {
        sa.rec = *rec;
        sa.rec.vertex = 0;
        if (replay) sa.shadow_counter = c->stats_dev + 6;
    }
    sa.out_rays = c->rays[1];
    sa.out_aux = c->aux[1];
    sa.out_count = cnt[1];
    rc = launch_shade(c, sa, n);
    if (rc) return rc;
    if (out_sh) {
        // each pixel's secondary ray from the primary bounce (RendererRef.cpp:185-197)
        hipLaunchKernelGGL(k_sh_reset_weights, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->sh_temp, rx, ry, rw,
                           rhh, c->w);
        hipLaunchKernelGGL(k_sh_samples, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->rays[1],
                           c->aux[1], cnt[1], c->w, c->sh_temp);
        RH_CHECK(hipGetLastError());
    }
    tm.mark(kPrimaryShade);

    int q = 1; // queue holding the rays to trace next
    const bool skip_indirect = (pi.settings.flags & RH_SKIP_INDIRECT_LIGHT) != 0;
    const int sort_on = c->sort_rays && p->sort_rays;
    for (int bounce = 0; bounce < pi.settings.max_total_depth && !skip_indirect; bounce++) {
        if (!rec && c->path_tail >= 0 && bounce >= c->path_tail) {
            // the remaining bounces, one lane per path
            ShadeArgs tb = sa;
            tb.rays = c->rays[q];
            tb.aux = c->aux[q];
            tb.count = cnt[q];
            tb.primary = 0;
            tb.ray_counter = c->stats_dev + 1;
            KernelMark km(c, 4);
            // with deferred NEE on, the lane resolves its light samples itself
            // after shading (the slots of a path are free during the tail)
            const int last = pi.settings.max_total_depth;
            const size_t lds = lds_bytes(c);
            const bool occ3 = c->tail_occupancy == 3;
            const void *kern;
            const bool solid = c->sc.all_solid && tb.nee_ray; // the lane's light samples resolved by first hit
            if (c->sc.wnodes && solid)
                kern = occ3 ? (const void *)k_path_tail<true, 3, kTreeWide, false, true>
                            : (const void *)k_path_tail<true, 2, kTreeWide, false, true>;
            else if (c->sc.wnodes)
                kern = tb.nee_ray ? (occ3 ? (const void *)k_path_tail<true, 3, kTreeWide> : (const void *)k_path_tail<true, 2, kTreeWide>)
                                  : (occ3 ? (const void *)k_path_tail<false, 3, kTreeWide> : (const void *)k_path_tail<false, 2, kTreeWide>);
            else if (solid)
                kern = occ3 ? (const void *)k_path_tail<true, 3, kTreeBinary, false, true>
                            : (const void *)k_path_tail<true, 2, kTreeBinary, false, true>;
            else
                kern = tb.nee_ray ? (occ3 ? (const void *)k_path_tail<true, 3, kTreeBinary> : (const void *)k_path_tail<true, 2, kTreeBinary>)
                                  : (occ3 ? (const void *)k_path_tail<false, 3, kTreeBinary> : (const void *)k_path_tail<false, 2, kTreeBinary>);
            if (c->count_visits && c->visits) {
                // the instrumented tail (2 waves per SIMD: counts do not depend on the knob)
                tb.visits = c->visits + 9;
                if (c->sc.wnodes)
                    kern = tb.nee_ray ? (const void *)k_path_tail<true, 2, kTreeWide, true>
                                      : (const void *)k_path_tail<false, 2, kTreeWide, true>;
                else
                    kern = tb.nee_ray ? (const void *)k_path_tail<true, 2, kTreeBinary, true>
                                      : (const void *)k_path_tail<false, 2, kTreeBinary, true>;
            }
            int blocks = (n + kBlock - 1) / kBlock;
            int *next_path = nullptr;
            if (c->tail_refill) {
                // persistent grid: the blocks that are resident at once, at most the queue's
                int per_cu = 0;
                rc = resident_blocks(c, kern, lds, &per_cu);
                if (rc) return rc;
                blocks = std::max(1, std::min(blocks, per_cu * c->cus));
                next_path = c->counters + 3;
                hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, next_path, 0);
            }
            void *args[] = {&tb, &bounce, (void *)&last, &next_path};
            RH_CHECK(hipLaunchKernel(kern, dim3(blocks), dim3(kBlock), args, lds, c->stream));
            // the path tail traces and shades in one kernel: its time counts as secondary tracing
            tm.mark(kSecondaryTrace);
            break;
        }
        if (sort_on) {
            {
                KernelMark km(c, 2);
                // the sorted rays go to the other buffer, their count with them
                rc = sort_queue(c->stream, c->rays[q], c->aux[q], c->rays[q ^ 1], c->aux[q ^ 1], cnt[q], cnt[q ^ 1],
                                n, c->sort_bits, pc, c->keys, c->vals, c->hist, c->stats_dev + 7);
            }
            if (rc) return rc;
            q ^= 1;
            tm.mark(kSecondarySort);
        }
        TraceArgs ta = trace_args(c, c->rays[q], cnt[q], n, c->hits, nullptr, count_secondary);
        ta.visits = c->visits + kStageVisits + 4 * (bounce + 1); // stage bounce + 1
        c->mark_stage = bounce + 1;
        rc = launch_trace(c, ta, n, false, bounce + 1);
        c->mark_stage = -1;
        if (rc) return rc;
        tm.mark(kSecondaryTrace);
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, cnt[q ^ 1], 0);
        ShadeArgs sb = sa;
        sb.rays = c->rays[q];
        sb.aux = c->aux[q];
        sb.count = cnt[q];
        sb.pi.bounce = bounce + 3;
        sb.primary = 0;
        sb.rec.vertex = bounce + 1;
        sb.out_rays = c->rays[q ^ 1];
        sb.out_aux = c->aux[q ^ 1];
        sb.out_count = cnt[q ^ 1];
        rc = launch_shade(c, sb, n);
        if (rc) return rc;
        tm.mark(kSecondaryShade);
        q ^= 1;
    }

    if (replay) return 0; // the replay leaves the image alone
    // passes mix into the accumulated image in pass order: this one after the
    // last pass's mix, which may sit on the other slot's stream
    if (c->mix_recorded) RH_CHECK(hipStreamWaitEvent(c->stream, c->ev_mix, 0));
    if (out_sh) {
        hipLaunchKernelGGL(k_sh_finish, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->temp, c->sh_temp,
                           c->sh_clean, rx, ry, rw, rhh, c->w, pc.mix_k);
        RH_CHECK(hipGetLastError());
    }
    tm.commit(); // mixing and tonemapping are outside the stage timers, as in RendererRef.cpp:280-290
    hipLaunchKernelGGL(k_mix_tonemap, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->temp, c->clean, c->final_, rx,
                       ry, rw, rhh, c->w, pc.mix_k, p->cam.dtype, p->cam.gamma, p->cam.pass_settings.flags);
    RH_CHECK(hipGetLastError());
    RH_CHECK(hipEventRecord(c->ev_mix, c->stream));
    c->mix_recorded = true;
    return 0;
}

// A plain pass goes to the slot the last pass did not use, so that its first
// stages run while the last pass's path tail drains (a tail keeps few waves
// resident once most of its paths have ended).  Passes share only the
// accumulated image: each mixes after the last one's mix (run_pass), so the
// image is the one passes run one after another give.  The bound stream is
// then ordered after every earlier pass, and other entry points queue their
// work there; when any did since the last pass (dirty), the new pass waits
// for all of it.
int begin_overlapped_pass(rh_context *c) {
    if (c->dirty) RH_CHECK(hipEventRecord(c->ev_join, c->stream));
    bind_slot(c, (c->bound + 1) % (c->pass_overlap + 1));
    if (c->dirty) {
        RH_CHECK(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        c->dirty = false;
    }
    return 0;
}

static void tape_forget(rh_context *c) {
    for (auto &r : c->tape_rect) r = {0, 0, 0, 0};
    for (auto &k : c->tape_key) k = 0;
}

// FNV-1a over everything of a pass that decides its paths: the camera and its
// pass settings (field by field: the settings record has padding), the rect,
// the iteration and the depth clamp.  A taped pass is reused by the adjoint
// only for a pass with the same key.
static uint64_t pass_key(const rh_context *c, const rh_pass_desc *p) {
    uint64_t h = 1469598103934665603ull;
    auto mix = [&](const void *data, size_t n) {
        const unsigned char *b = static_cast<const unsigned char *>(data);
        for (size_t i = 0; i < n; i++) h = (h ^ b[i]) * 1099511628211ull;
    };
    const rh_camera &cam = p->cam;
    mix(&cam.type, offsetof(rh_camera, pass_settings) - offsetof(rh_camera, type));
    const rh_pass_settings &ps = cam.pass_settings;
    const uint8_t depths[6] = {ps.max_diff_depth,   ps.max_glossy_depth, ps.max_refr_depth,
                               ps.max_transp_depth, ps.max_total_depth,  ps.termination_start_depth};
    mix(depths, sizeof(depths));
    mix(&ps.flags, sizeof(ps.flags));
    mix(p->rect, sizeof(p->rect));
    mix(&p->iteration, sizeof(p->iteration));
    mix(&c->max_bounces, sizeof(c->max_bounces));
    return h | 1ull; // never 0, the key of an empty slot
}

// Buffers of the tape for passes of `lanes` lanes; false when it is off or the
// pass cannot be taped (geometry cameras, SH output).
static int tape_slot(rh_context *c, const rh_pass_desc *p, size_t lanes) {
    if (c->tape_passes <= 0 || p->iteration < 1 || p->iteration > c->tape_passes) return -1;
    if (p->cam.type == RH_CAM_GEO || (p->cam.pass_settings.flags & RH_OUTPUT_SH)) return -1;
    return p->iteration - 1;
}

static int tape_alloc(rh_context *c, size_t lanes) {
    const size_t vmax = (size_t)std::min(c->max_bounces, RH_MAX_BOUNCES_EXT) + 1;
    const int slots = c->tape_passes;
    if (c->tape_rec && c->tape_lanes == lanes && c->tape_vmax == vmax && (int)c->tape_rect.size() == slots &&
        c->tape_texel_slots == (c->tape_texels ? slots : 0))
        return 0;
    RH_CHECK(hipStreamSynchronize(c->stream));
    release(c->tape_rec);
    release(c->tape_rec_tex);
    release(c->tape_nverts);
    c->tape_rect.assign(slots, {0, 0, 0, 0});
    c->tape_key.assign(slots, 0);
    c->tape_lanes = c->tape_vmax = 0;
    c->tape_texel_slots = 0;
    RH_CHECK(hipMalloc(&c->tape_rec, (size_t)slots * vmax * lanes * sizeof(float4)));
    if (c->tape_texels) RH_CHECK(hipMalloc(&c->tape_rec_tex, (size_t)slots * vmax * 3 * lanes * sizeof(float4)));
    RH_CHECK(hipMalloc(&c->tape_nverts, (size_t)slots * lanes * sizeof(int)));
    c->tape_lanes = lanes;
    c->tape_vmax = vmax;
    c->tape_texel_slots = c->tape_texels ? slots : 0;
    return 0;
}

static AdjRecords tape_records(rh_context *c, int slot, const rh_pass_desc *p, bool texels) {
    const size_t lanes = c->tape_lanes, vmax = c->tape_vmax;
    AdjRecords r{c->tape_rec + (size_t)slot * vmax * lanes,
                 texels ? c->tape_rec_tex + (size_t)slot * vmax * 3 * lanes : nullptr,
                 c->tape_nverts + (size_t)slot * lanes, (uint32_t)lanes, 0, {p->rect[0], p->rect[1], p->rect[2]}};
    return r;
}

int rh_render_pass(rh_context *c, const rh_pass_desc *p) {
    const size_t lanes = (size_t)std::max(0, p->rect[2]) * std::max(0, p->rect[3]);
    const int slot = tape_slot(c, p, lanes);
    const bool plain = slot < 0 && p->cam.type != RH_CAM_GEO && !(p->cam.pass_settings.flags & RH_OUTPUT_SH);
    if (plain && c->pass_overlap) {
        int rc = begin_overlapped_pass(c);
        if (rc == 0) rc = run_pass(c, p, nullptr);
        if (rc) {
            // a pass that failed part-way has no mix the next work can follow
            (void)sync_slots(c);
            c->dirty = true;
        }
        return rc;
    }
    c->dirty = true;
    if (slot < 0) return run_pass(c, p, nullptr);
    int rc = tape_alloc(c, lanes);
    if (rc) return rc;
    c->tape_rect[slot] = {0, 0, 0, 0};
    AdjRecords rec = tape_records(c, slot, p, c->tape_texels != 0);
    RH_CHECK(hipMemsetAsync(rec.nverts, 0, lanes * sizeof(int), c->stream));
    rc = run_pass(c, p, &rec, true);
    if (rc) return rc;
    c->tape_rect[slot] = {p->rect[0], p->rect[1], p->rect[2], p->rect[3]};
    c->tape_key[slot] = pass_key(c, p);
    return 0;
}

int rh_read_pixels(rh_context *c, float *rgba, size_t count) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const size_t n = (size_t)c->w * c->h * 4;
    if (count < n) return arg_error("output buffer too small");
    RH_CHECK(hipMemcpyAsync(rgba, c->final_, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_synchronize(rh_context *c) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    return sync_slots(c);
}

int rh_get_stats(rh_context *c, rh_stats *out) {
    const int rc = stage_harvest(c, true);
    if (rc) return rc;
    unsigned long long *f = &out->time_primary_ray_gen_us;
    for (int k = 0; k < 6; k++) f[k] = (unsigned long long)c->stage_us[k];
    return 0;
}

int rh_reset_stats(rh_context *c) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const int rc = stage_harvest(c, true);
    if (rc) return rc;
    for (double &t : c->stage_us) t = 0.0;
    RH_CHECK(hipMemsetAsync(c->stats_dev, 0, 8 * sizeof(unsigned long long), c->stream));
    return 0;
}

int rh_get_ray_counts(rh_context *c, rh_ray_counts *out) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    unsigned long long v[8];
    RH_CHECK(hipMemcpyAsync(v, c->stats_dev, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    out->primary = v[0];
    out->secondary = v[1];
    out->shadow = v[2];
    out->sorted = v[3];
    out->adjoint = v[5];
    out->adjoint_shadow = v[6];
    out->overflow = v[4];
    return 0;
}

int rh_set_option(rh_context *c, const char *name, int value) {
    const std::string n = name;
    if (n == "sort_rays") c->sort_rays = value != 0;
    else if (n == "stage_timing") c->stage_timing = value != 0;
    else if (n == "kernel_timing") c->kernel_timing = value != 0;
    else if (n == "count_visits") c->count_visits = value != 0;
    else if (n == "defer_nee") c->defer_nee = value != 0;
    else if (n == "lds_stack_depth") {
        if (value < 4 || value > 64) return arg_error("lds_stack_depth must be 4..64");
        c->lds_depth = value;
    } else if (n == "path_tail") {
        if (value < -1 || value > 255) return arg_error("path_tail must be -1..255");
        c->path_tail = value;
    } else if (n == "tail_occupancy") {
        if (value != 2 && value != 3) return arg_error("tail_occupancy must be 2 or 3");
        c->tail_occupancy = value;
    } else if (n == "refill") {
        c->refill = value != 0;
    } else if (n == "tail_refill") {
        c->tail_refill = value != 0;
    } else if (n == "tiled_primary") {
        c->tiled_primary = value != 0;
    } else if (n == "pass_overlap") {
        if (value < 0 || value >= kMaxPassSlots) return arg_error("pass_overlap must be 0..3");
        c->pass_overlap = value;
    } else if (n == "shadow_occupancy") {
        if (value != -1 && value != 0 && value != 5 && value != 6)
            return arg_error("shadow_occupancy must be -1, 0, 5 or 6");
        c->shadow_occupancy = value;
    } else if (n == "adjoint_tape") {
        if (value < 0 || value > 4096) return arg_error("adjoint_tape must be 0..4096 passes");
        c->tape_passes = value;
        tape_forget(c);
    } else if (n == "adjoint_tape_texels") {
        c->tape_texels = value != 0;
        tape_forget(c);
    } else if (n == "adjoint_staged") {
        c->adjoint_staged = value != 0;
    } else if (n == "max_bounces") {
        if (value < 1 || value > RH_MAX_BOUNCES_EXT) return arg_error("max_bounces must be 1..16");
        c->max_bounces = value;
    } else if (n == "trace_occupancy") {
        if (value != 0 && value != 5 && value != 6) return arg_error("trace_occupancy must be 0, 5 or 6");
        c->trace_occupancy = value;
    } else if (n == "trace_group") {
        if (value != 0 && value != 2 && value != 4 && value != 8) return arg_error("trace_group must be 0, 2, 4 or 8");
        c->trace_group = value;
    } else if (n == "adjoint_occupancy") {
        if (value != 2 && value != 3) return arg_error("adjoint_occupancy must be 2 or 3");
        c->adjoint_occupancy = value;
    } else if (n == "sort_bits") {
        if (value != 8 && value != 16 && value != 24 && value != 32) return arg_error("sort_bits must be 8, 16, 24 or 32");
        c->sort_bits = value;
    } else return arg_error("unknown option");
    return 0;
}

int rh_get_option(rh_context *c, const char *name, int *value) {
    const std::string n = name;
    if (n == "sort_rays") *value = c->sort_rays;
    else if (n == "stage_timing") *value = c->stage_timing;
    else if (n == "lds_stack_depth") *value = c->lds_depth;
    else if (n == "sort_bits") *value = c->sort_bits;
    else if (n == "path_tail") *value = c->path_tail;
    else if (n == "tail_occupancy") *value = c->tail_occupancy;
    else if (n == "refill") *value = c->refill;
    else if (n == "tail_refill") *value = c->tail_refill;
    else if (n == "tiled_primary") *value = c->tiled_primary;
    else if (n == "scene_uploads") *value = c->n_uploads;
    else if (n == "instance_updates") *value = c->n_instance_updates;
    else if (n == "trace_occupancy") *value = c->trace_occupancy;
    else if (n == "trace_group") *value = c->trace_group;
    else if (n == "max_bounces") *value = c->max_bounces;
    else if (n == "adjoint_staged") *value = c->adjoint_staged;
    else if (n == "adjoint_tape") *value = c->tape_passes;
    else if (n == "shadow_occupancy") *value = c->shadow_occupancy;
    else if (n == "adjoint_tape_texels") *value = c->tape_texels;
    else if (n == "adjoint_occupancy") *value = c->adjoint_occupancy;
    else if (n == "defer_nee") *value = c->defer_nee;
    else if (n == "pass_overlap") *value = c->pass_overlap;
    else if (n == "bound_slot") *value = c->bound; // read-only: the slot of the last pass
    else if (n == "stack_overflows" || n == "sort_stalls") {
        // read-only: dropped traversal stack entries; sort look-backs that gave up
        c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
        unsigned long long v[8];
        RH_CHECK(hipMemcpyAsync(v, c->stats_dev, sizeof(v), hipMemcpyDeviceToHost, c->stream));
        RH_CHECK(hipStreamSynchronize(c->stream));
        *value = (int)v[n == "sort_stalls" ? 7 : 4];
    } else return arg_error("unknown option");
    return 0;
}

int rh_device_synchronize(rh_context *c) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    RH_CHECK(hipSetDevice(c->device));
    RH_CHECK(hipDeviceSynchronize());
    return 0;
}

int rh_get_kernel_profile(rh_context *c, rh_kernel_profile *out) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    memset(out, 0, sizeof(*out));
    int rc = sync_slots(c);
    if (rc) return rc;
    for (const auto &m : c->ev_marks) {
        float ms = 0;
        RH_CHECK(hipEventElapsedTime(&ms, c->ev_pool[m.second.first], c->ev_pool[m.second.second]));
        const int kind = m.first & 7, stage = (m.first >> 3) - 1;
        out->ms[kind] += ms;
        out->launches[kind]++;
        if (stage >= 0 && stage < RH_TRACE_STAGES) {
            out->trace_stage_ms[stage] += ms;
            out->trace_stage_launches[stage]++;
        }
    }
    unsigned long long v[kVisitWords];
    RH_CHECK(hipMemcpy(v, c->visits, sizeof(v), hipMemcpyDeviceToHost));
    // the trace totals: the parity entry points' launches and every stage's
    unsigned long long tot[4] = {v[0], v[1], v[2], v[3]};
    for (int k = 0; k < RH_TRACE_STAGES; k++)
        for (int j = 0; j < 4; j++) {
            out->trace_stage_visits[k][j] = v[kStageVisits + 4 * k + j];
            tot[j] += v[kStageVisits + 4 * k + j];
        }
    out->node_visits = tot[0];
    out->counted_rays = tot[3];
    out->tri_tests = tot[1];
    out->instance_tests = tot[2];
    for (int k = 0; k < 5; k++) {
        out->shadow_visits[k] = v[4 + k];
        out->tail_visits[k] = v[9 + k];
    }
    return 0;
}

int rh_reset_kernel_profile(rh_context *c) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    int rc = sync_slots(c);
    if (rc) return rc;
    c->ev_marks.clear();
    c->ev_used = 0;
    RH_CHECK(hipMemsetAsync(c->visits, 0, kVisitWords * sizeof(unsigned long long), c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

// ---------------------------------------------------------------- parity entry points

static void packet_to_split(const rh_ray_packet &p, rh_ray &r, rh_ray_aux &a) {
    memcpy(r.o, p.o, sizeof(r.o));
    memcpy(r.d, p.d, sizeof(r.d));
    r.xy = p.xy;
    r.depth = p.ray_depth;
    memcpy(a.c, p.c, sizeof(a.c));
    a.ior = p.ior;
    memcpy(a.do_dx, p.do_dx, sizeof(a.do_dx));
    memcpy(a.dd_dx, p.dd_dx, sizeof(a.dd_dx));
    memcpy(a.do_dy, p.do_dy, sizeof(a.do_dy));
    memcpy(a.dd_dy, p.dd_dy, sizeof(a.dd_dy));
}
static void split_to_packet(const rh_ray &r, const rh_ray_aux &a, rh_ray_packet &p) {
    memcpy(p.o, r.o, sizeof(r.o));
    memcpy(p.d, r.d, sizeof(r.d));
    p.xy = r.xy;
    p.ray_depth = r.depth;
    memcpy(p.c, a.c, sizeof(a.c));
    p.ior = a.ior;
    memcpy(p.do_dx, a.do_dx, sizeof(a.do_dx));
    memcpy(p.dd_dx, a.dd_dx, sizeof(a.dd_dx));
    memcpy(p.do_dy, a.do_dy, sizeof(a.do_dy));
    memcpy(p.dd_dy, a.dd_dy, sizeof(a.dd_dy));
}

static int upload_packets(rh_context *c, const rh_ray_packet *rays, size_t n, int q) {
    std::vector<rh_ray> r(n);
    std::vector<rh_ray_aux> a(n);
    for (size_t i = 0; i < n; i++) packet_to_split(rays[i], r[i], a[i]);
    RH_CHECK(hipMemcpyAsync(c->rays[q], r.data(), n * sizeof(rh_ray), hipMemcpyHostToDevice, c->stream));
    RH_CHECK(hipMemcpyAsync(c->aux[q], a.data(), n * sizeof(rh_ray_aux), hipMemcpyHostToDevice, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

static int download_packets(rh_context *c, rh_ray_packet *out, size_t n, int q) {
    std::vector<rh_ray> r(n);
    std::vector<rh_ray_aux> a(n);
    RH_CHECK(hipMemcpyAsync(r.data(), c->rays[q], n * sizeof(rh_ray), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipMemcpyAsync(a.data(), c->aux[q], n * sizeof(rh_ray_aux), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < n; i++) split_to_packet(r[i], a[i], out[i]);
    return 0;
}

int rh_gen_primary_rays(rh_context *c, const rh_camera *cam, const int32_t rect[4], int w, int h, int iteration,
                        const float *halton, rh_ray_packet *out) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const int n = rect[2] * rect[3];
    if (n <= 0) return 0;
    int rc = ensure_rays(c, (size_t)n);
    if (rc) return rc;
    const int save_w = c->w, save_h = c->h;
    rh_pass_desc p;
    memset(&p, 0, sizeof(p));
    p.cam = *cam;
    memcpy(p.rect, rect, sizeof(p.rect));
    p.iteration = iteration;
    c->w = w;
    c->h = h;
    PassConst pc = make_pass_const(c, &p);
    c->w = save_w;
    c->h = save_h;
    RH_CHECK(upload_halton(c->halton, halton, c->stream));
    hipLaunchKernelGGL(k_gen_primary<false>, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, pc, c->halton,
                       c->rays[0], c->aux[0], c->counters + 0);
    RH_CHECK(hipGetLastError());
    return download_packets(c, out, (size_t)n, 0);
}

int rh_trace_rays(rh_context *c, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax, rh_hit *out) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    if (!n) return 0;
    int rc = ensure_rays(c, n);
    if (rc) return rc;
    rc = upload_packets(c, rays, n, 0);
    if (rc) return rc;
    float *d_tmax = nullptr;
    if (tmax) {
        RH_CHECK(hipMalloc(&d_tmax, n * sizeof(float)));
        RH_CHECK(hipMemcpyAsync(d_tmax, tmax, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    }
    rc = launch_trace(c, trace_args(c, c->rays[0], nullptr, (int)n, c->hits, d_tmax, nullptr), (int)n, any_hit != 0);
    if (rc) return rc;
    RH_CHECK(hipMemcpyAsync(out, c->hits, n * sizeof(rh_hit), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    free_ptr(d_tmax);
    return 0;
}

int rh_trace_device(rh_context *c, const rh_ray *d_rays, int n, int any_hit, rh_hit *d_hits, void *stream) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    if (n <= 0) return 0;
    if ((size_t)n > c->max_rays) return arg_error("n exceeds the context's ray capacity (resize first)");
    TraceArgs a = trace_args(c, d_rays, nullptr, n, d_hits, nullptr, c->stats_dev + 1);
    const int blocks = (n + kBlock - 1) / kBlock;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    const bool wide = c->sc.wnodes != nullptr;
    if (any_hit) {
        if (wide) hipLaunchKernelGGL((k_trace<true, false, kTreeWide>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
        else hipLaunchKernelGGL((k_trace<true, false, kTreeBinary>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
    } else {
        if (wide) hipLaunchKernelGGL((k_trace<false, false, kTreeWide>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
        else hipLaunchKernelGGL((k_trace<false, false, kTreeBinary>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
    }
    RH_CHECK(hipGetLastError());
    return 0;
}

int rh_ray_hashes(rh_context *c, const rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3],
                  uint32_t *out) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!n) return 0;
    int rc = ensure_rays(c, n);
    if (rc) return rc;
    rc = upload_packets(c, rays, n, 0);
    if (rc) return rc;
    PassConst pc;
    memset(&pc, 0, sizeof(pc));
    memcpy(pc.root_min, root_min, sizeof(pc.root_min));
    memcpy(pc.cell_size, cell_size, sizeof(pc.cell_size));
    hipLaunchKernelGGL(k_ray_keys, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream, c->rays[0],
                       nullptr, (int)n, pc, c->keys[0], c->vals[0]);
    RH_CHECK(hipGetLastError());
    RH_CHECK(hipMemcpyAsync(out, c->keys[0], n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_sort_rays(rh_context *c, rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3]) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!n) return 0;
    int rc = ensure_rays(c, n);
    if (rc) return rc;
    rc = upload_packets(c, rays, n, 0);
    if (rc) return rc;
    PassConst pc;
    memset(&pc, 0, sizeof(pc));
    memcpy(pc.root_min, root_min, sizeof(pc.root_min));
    memcpy(pc.cell_size, cell_size, sizeof(pc.cell_size));
    // the full key, as SortRays_CPU
    rc = sort_queue(c->stream, c->rays[0], c->aux[0], c->rays[1], c->aux[1], nullptr, nullptr, (int)n, 32, pc, c->keys,
                    c->vals, c->hist, c->stats_dev + 7);
    if (rc) return rc;
    return download_packets(c, rays, n, 1);
}

int rh_sort_rays_noalloc(const rh_ray *d_rays, const rh_ray_aux *d_aux, size_t n, const float root_min[3],
                         const float cell_size[3], rh_ray *d_rays_out, rh_ray_aux *d_aux_out, void *stream) {
    if (!n) return 0;
    if (n > (size_t)RH_NOALLOC_MAX_RAYS) return arg_error("rh_sort_rays_noalloc: more rays than RH_NOALLOC_MAX_RAYS");
    if (!d_rays || !d_aux || !d_rays_out || !d_aux_out || !root_min || !cell_size)
        return arg_error("rh_sort_rays_noalloc: null buffer");
    if ((const void *)d_rays == (const void *)d_rays_out || (const void *)d_aux == (const void *)d_aux_out)
        return arg_error("rh_sort_rays_noalloc: the output must not be the input");
    uint32_t *base = nullptr;
    const int rc0 = scratch_ptr(&base);
    if (rc0) return rc0;
    const size_t m = RH_NOALLOC_MAX_RAYS;
    uint32_t *const keys[2] = {base, base + m}, *const vals[2] = {base + 2 * m, base + 3 * m};
    uint32_t *hist = base + 4 * m;
    PassConst pc;
    memset(&pc, 0, sizeof(pc));
    memcpy(pc.root_min, root_min, sizeof(pc.root_min));
    memcpy(pc.cell_size, cell_size, sizeof(pc.cell_size));
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the full key, as SortRays_CPU; a look-back that gives up is not reported here (no device counter)
    return sort_queue(s, d_rays, d_aux, d_rays_out, d_aux_out, nullptr, nullptr, (int)n, 32, pc, keys, vals, hist,
                      nullptr);
}

int rh_shade_rays(rh_context *c, const rh_pass_info *pi, const rh_ray_packet *rays, const rh_hit *hits, size_t n,
                  const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_count, int *out_src) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    *out_count = 0;
    if (!n) return 0;
    int rc = ensure_rays(c, n);
    if (rc) return rc;
    rc = upload_packets(c, rays, n, 0);
    if (rc) return rc;
    rh_pass_info *d_pi = nullptr;
    float4 *d_rgba = nullptr;
    int *d_src = nullptr;
    RH_CHECK(hipMalloc(&d_pi, n * sizeof(rh_pass_info)));
    RH_CHECK(hipMalloc(&d_rgba, n * sizeof(float4)));
    RH_CHECK(hipMalloc(&d_src, n * sizeof(int)));
    RH_CHECK(hipMemcpyAsync(d_pi, pi, n * sizeof(rh_pass_info), hipMemcpyHostToDevice, c->stream));
    RH_CHECK(hipMemcpyAsync(c->hits, hits, n * sizeof(rh_hit), hipMemcpyHostToDevice, c->stream));
    RH_CHECK(upload_halton(c->halton, halton, c->stream));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, c->counters + 1, 0);
    ShadeArgs a = shade_args(c);
    a.rays = c->rays[0];
    a.aux = c->aux[0];
    a.hits = c->hits;
    a.count = nullptr;
    a.n = (int)n;
    a.per_ray_pi = d_pi;
    a.halton = c->halton;
    a.temp = nullptr;
    a.out_rgba = d_rgba;
    a.out_rays = c->rays[1];
    a.out_aux = c->aux[1];
    a.out_count = c->counters + 1;
    a.out_src = d_src;
    rc = launch_shade(c, a, (int)n);
    if (rc) return rc;
    int cnt = 0;
    RH_CHECK(hipMemcpyAsync(&cnt, c->counters + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipMemcpyAsync(out_rgba, d_rgba, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    if (cnt > 0) {
        rc = download_packets(c, out_secondary, (size_t)cnt, 1);
        if (rc) return rc;
        if (out_src) RH_CHECK(hipMemcpy(out_src, d_src, cnt * sizeof(int), hipMemcpyDeviceToHost));
    }
    *out_count = cnt;
    free_ptr(d_pi);
    free_ptr(d_rgba);
    free_ptr(d_src);
    return 0;
}


// ---- differentiation: path-replay adjoint of one pass, material colours, accumulated image

int rh_adjoint_pass(rh_context *c, const rh_pass_desc *p, const float *d_adj_rgba, float *d_grad, float *d_texel_grad,
                    float scale) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    const int rx = p->rect[0], ry = p->rect[1], rw = p->rect[2], rhh = p->rect[3];
    if (rw <= 0 || rhh <= 0 || rx < 0 || ry < 0 || rx + rw > c->w || ry + rhh > c->h)
        return arg_error("region outside the framebuffer");
    if (p->iteration < 1) return arg_error("iteration must be >= 1");
    if (!d_adj_rgba || !d_grad) return arg_error("adjoint image and gradient buffers are required");
    if (p->cam.type == RH_CAM_GEO) return arg_error("the adjoint replay takes perspective cameras only");
    const int depth = std::min<int>(p->cam.pass_settings.max_total_depth, c->max_bounces);
    const size_t lanes = (size_t)rw * rhh;
    // vertex records: at most depth + 1 <= RH_MAX_BOUNCES_EXT + 1 per lane
    if (lanes > c->adj_rec_cap || !c->adj_rec) {
        release(c->adj_rec);
        c->adj_rec_cap = 0;
        RH_CHECK(hipMalloc(&c->adj_rec, (RH_MAX_BOUNCES_EXT + 1) * lanes * sizeof(float4)));
        c->adj_rec_cap = lanes;
    }
    if (d_texel_grad && (lanes > c->adj_rec_tex_cap || !c->adj_rec_tex)) {
        release(c->adj_rec_tex);
        c->adj_rec_tex_cap = 0;
        RH_CHECK(hipMalloc(&c->adj_rec_tex, 3 * (RH_MAX_BOUNCES_EXT + 1) * lanes * sizeof(float4)));
        c->adj_rec_tex_cap = lanes;
    }
    const bool defer = c->defer_nee && c->nee_slots;
    // NEE slots for every lane when deferred (this may grow the spill array,
    // so it comes before any buffer pointer is taken)
    int rc = defer ? ensure_rays(c, lanes) : 0;
    if (rc) return rc;
    rc = ensure_spill(c, lanes);
    if (rc) return rc;
    if (d_texel_grad && !c->coarse_valid) {
        // replica slots for the small texel blocks (each block once, however
        // many (texture, lod) entries name it)
        free_ptr(c->coarse_map);
        free_ptr(c->coarse_texel);
        free_ptr(c->rep);
        c->coarse_map = nullptr;
        c->coarse_texel = nullptr;
        c->rep = nullptr;
        std::vector<uint2> map(c->host_tex_mips.size());
        std::vector<uint32_t> texel;
        std::vector<std::pair<uint32_t, uint32_t>> seen; // block offset -> first slot
        for (size_t e = 0; e < map.size(); e++) {
            const rh_tex_mip &m = c->host_tex_mips[e];
            const uint32_t n = (uint32_t)m.bw * m.bh;
            map[e] = make_uint2(m.offset, kNoKey);
            if (!n || n > kSmallBlockTexels) continue;
            uint32_t slot = kNoKey;
            for (auto &b : seen)
                if (b.first == m.offset) slot = b.second;
            if (slot == kNoKey) {
                slot = (uint32_t)texel.size();
                seen.emplace_back(m.offset, slot);
                for (uint32_t t = 0; t < n; t++) texel.push_back(m.offset + t);
            }
            map[e].y = slot;
        }
        c->rep_count = (uint32_t)texel.size();
        if (!map.empty()) {
            RH_CHECK(hipMalloc(&c->coarse_map, map.size() * sizeof(uint2)));
            RH_CHECK(hipMemcpyAsync(c->coarse_map, map.data(), map.size() * sizeof(uint2), hipMemcpyHostToDevice,
                                    c->stream));
        }
        if (c->rep_count) {
            const size_t bytes = (size_t)kTexelReplicas * c->rep_count * 4 * sizeof(float);
            RH_CHECK(hipMalloc(&c->coarse_texel, texel.size() * sizeof(uint32_t)));
            RH_CHECK(hipMemcpyAsync(c->coarse_texel, texel.data(), texel.size() * sizeof(uint32_t),
                                    hipMemcpyHostToDevice, c->stream));
            RH_CHECK(hipMalloc(&c->rep, bytes));
            RH_CHECK(hipMemsetAsync(c->rep, 0, bytes, c->stream));
        }
        RH_CHECK(hipStreamSynchronize(c->stream));
        c->coarse_valid = true;
    }
    AdjointArgs a;
    memset(&a, 0, sizeof(a));
    a.sc = c->sc;
    a.pc = make_pass_const(c, p);
    a.pi.iteration = p->iteration;
    a.pi.bounce = 2;
    a.pi.settings = p->cam.pass_settings;
    a.pi.settings.max_total_depth = (uint8_t)depth;
    a.halton = c->halton;
    a.adj = reinterpret_cast<const float4 *>(d_adj_rgba);
    a.scale = scale;
    a.grad = d_grad;
    a.n_materials = c->n_materials;
    a.rec = c->adj_rec;
    a.lanes = (uint32_t)lanes;
    a.tex_grad = d_texel_grad;
    a.tex_mask = c->grad_tex_mask;
    a.rec_tex = d_texel_grad ? c->adj_rec_tex : nullptr;
    if (d_texel_grad && c->rep_count) {
        a.rep = c->rep;
        a.rep_count = c->rep_count;
        a.coarse_map = c->coarse_map;
    }
    a.spill = c->spill;
    a.spill_stride = (uint32_t)c->spill_cap;
    a.lds_depth = (uint32_t)c->lds_depth;
    a.ray_counter = c->stats_dev + 5;
    a.shadow_counter = c->stats_dev + 6;
    a.overflow = c->stats_dev + 4;
    if (defer) {
        a.nee_ray = c->nee_ray;
        a.nee_aux = c->nee_aux;
        a.nee_term = c->nee_term;
        a.nee_slots = c->nee_slots;
    }
    const int tslot = (p->iteration >= 1 && p->iteration <= (int)c->tape_rect.size()) ? p->iteration - 1 : -1;
    if (tslot >= 0 && c->tape_rect[tslot][0] == rx && c->tape_rect[tslot][1] == ry && c->tape_rect[tslot][2] == rw &&
        c->tape_rect[tslot][3] == rhh && c->tape_key[tslot] == pass_key(c, p) && (!d_texel_grad || c->tape_texels)) {
        // the forward pass kept this pass's vertex records: no replay
        const AdjRecords tr = tape_records(c, tslot, p, d_texel_grad != nullptr);
        a.rec = tr.rec;
        a.rec_tex = tr.rec_tex;
        KernelMark km(c, 4);
        hipLaunchKernelGGL(k_adjoint_backward, dim3((unsigned)((lanes + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, a, (const int *)tr.nverts);
        if (a.rep)
            hipLaunchKernelGGL(k_fold_replicas, dim3((a.rep_count + 255) / 256), dim3(256), 0, c->stream, a.rep,
                               a.rep_count, (const uint32_t *)c->coarse_texel, d_texel_grad);
        RH_CHECK(hipGetLastError());
        return 0;
    }
    if (c->adjoint_staged) {
        // the forward stage kernels replay the pass and write the vertex
        // records (every bounce staged, as many rays in flight as the pass has);
        // one kernel then forms the suffix sums
        if (lanes > c->adj_nverts_cap || !c->adj_nverts) {
            release(c->adj_nverts);
            c->adj_nverts_cap = 0;
            RH_CHECK(hipMalloc(&c->adj_nverts, lanes * sizeof(int)));
            c->adj_nverts_cap = lanes;
        }
        RH_CHECK(hipMemsetAsync(c->adj_nverts, 0, lanes * sizeof(int), c->stream));
        AdjRecords rec{c->adj_rec, d_texel_grad ? c->adj_rec_tex : nullptr, c->adj_nverts, (uint32_t)lanes, 0,
                       {rx, ry, rw}};
        rc = run_pass(c, p, &rec);
        if (rc) return rc;
        KernelMark km(c, 4);
        hipLaunchKernelGGL(k_adjoint_backward, dim3((unsigned)((lanes + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                           c->stream, a, (const int *)c->adj_nverts);
        if (a.rep)
            hipLaunchKernelGGL(k_fold_replicas, dim3((a.rep_count + 255) / 256), dim3(256), 0, c->stream, a.rep,
                               a.rep_count, (const uint32_t *)c->coarse_texel, d_texel_grad);
        RH_CHECK(hipGetLastError());
        return 0;
    }
    KernelMark km(c, 4);
    const dim3 grid((unsigned)((lanes + kBlock - 1) / kBlock)), blk(kBlock);
    const size_t lds = lds_bytes(c);
    const bool occ3 = c->adjoint_occupancy == 3;
    if (c->sc.wnodes) {
        if (defer && occ3) hipLaunchKernelGGL((k_adjoint<true, 3, kTreeWide>), grid, blk, lds, c->stream, a);
        else if (defer) hipLaunchKernelGGL((k_adjoint<true, 2, kTreeWide>), grid, blk, lds, c->stream, a);
        else if (occ3) hipLaunchKernelGGL((k_adjoint<false, 3, kTreeWide>), grid, blk, lds, c->stream, a);
        else hipLaunchKernelGGL((k_adjoint<false, 2, kTreeWide>), grid, blk, lds, c->stream, a);
    } else {
        if (defer && occ3) hipLaunchKernelGGL((k_adjoint<true, 3, kTreeBinary>), grid, blk, lds, c->stream, a);
        else if (defer) hipLaunchKernelGGL((k_adjoint<true, 2, kTreeBinary>), grid, blk, lds, c->stream, a);
        else if (occ3) hipLaunchKernelGGL((k_adjoint<false, 3, kTreeBinary>), grid, blk, lds, c->stream, a);
        else hipLaunchKernelGGL((k_adjoint<false, 2, kTreeBinary>), grid, blk, lds, c->stream, a);
    }
    if (a.rep)
        hipLaunchKernelGGL(k_fold_replicas, dim3((a.rep_count + 255) / 256), dim3(256), 0, c->stream, a.rep,
                           a.rep_count, (const uint32_t *)c->coarse_texel, d_texel_grad);
    RH_CHECK(hipGetLastError());
    return 0;
}

int rh_set_gradient_textures(rh_context *c, const uint32_t *ids, uint32_t n) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!n) {
        free_ptr(c->grad_tex_mask);
        c->grad_tex_mask = nullptr;
        c->grad_tex_words = 0;
        return 0;
    }
    if (!ids) return arg_error("texture ids required");
    uint32_t max_id = 0;
    for (uint32_t k = 0; k < n; k++) max_id = std::max(max_id, ids[k]);
    std::vector<uint32_t> mask(max_id / 32 + 1, 0u);
    for (uint32_t k = 0; k < n; k++) mask[ids[k] >> 5] |= 1u << (ids[k] & 31);
    // texture ids past the mask read as unselected: size it for every texture of the scene
    mask.resize(std::max<size_t>(mask.size(), (c->n_textures + 31) / 32 + 1), 0u);
    if (mask.size() > c->grad_tex_words || !c->grad_tex_mask) {
        release(c->grad_tex_mask);
        c->grad_tex_words = 0;
        RH_CHECK(hipMalloc(&c->grad_tex_mask, mask.size() * sizeof(uint32_t)));
        c->grad_tex_words = mask.size();
    }
    RH_CHECK(hipMemcpyAsync(c->grad_tex_mask, mask.data(), mask.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                            c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_set_material_colors(rh_context *c, uint32_t first, uint32_t n, const float *rgb) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    if ((size_t)first + n > c->n_materials) return arg_error("material range outside the scene");
    if (!n) return 0;
    // taped records hold colours of the old materials: the adjoint replays
    tape_forget(c);
    const size_t floats = (size_t)3 * n;
    if (floats > c->color_stage_cap || !c->color_stage) {
        release(c->color_stage);
        c->color_stage_cap = 0;
        RH_CHECK(hipMalloc(&c->color_stage, floats * sizeof(float)));
        c->color_stage_cap = floats;
    }
    RH_CHECK(hipMemcpyAsync(c->color_stage, rgb, floats * sizeof(float), hipMemcpyDefault, c->stream));
    hipLaunchKernelGGL(k_set_material_colors, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       const_cast<rh_material *>(c->sc.materials), c->color_stage, first, n);
    RH_CHECK(hipGetLastError());
    // the caller's array may change right after this call returns
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_set_texels(rh_context *c, uint32_t first, uint32_t n, const uint32_t *rgba) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    if (!c->have_scene) return arg_error("no scene uploaded");
    if ((size_t)first + n > c->n_texels) return arg_error("texel range outside the scene");
    if (!n) return 0;
    // taped records hold path colours and raw values of the old texels: the adjoint replays
    tape_forget(c);
    RH_CHECK(hipMemcpyAsync(const_cast<uint32_t *>(c->sc.texels) + first, rgba, (size_t)n * sizeof(uint32_t),
                            hipMemcpyDefault, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_read_accum(rh_context *c, float *dst, size_t count) {
    c->dirty = true; // work other than an overlapped pass (begin_overlapped_pass)
    const size_t n = (size_t)c->w * c->h * 4;
    if (count < n) return arg_error("output buffer too small");
    RH_CHECK(hipMemcpyAsync(dst, c->clean, n * sizeof(float), hipMemcpyDefault, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

} // extern "C"
