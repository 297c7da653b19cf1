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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rh_hip.h"
#include "rh_shade.h"

using namespace rh;

namespace {

thread_local std::string g_err;

constexpr int kBlock = 256;        // threads per block of the ray kernels (4 waves)
constexpr int kSpillDepth = 48;    // per-ray global stack entries beyond the LDS part
constexpr int kSortItems = 16;     // keys per thread in the radix passes
constexpr int kSortTile = kBlock * kSortItems;
constexpr uint32_t kMaxDeferredLights = 8; // lights per scene for the deferred NEE stage

// ------------------------------------------------------------------ tables
// Ray-hash tables of the reference (Core.cpp:106-150): sphere-sector lookup
// for the direction part of the key.  The 8-bit Morton spreads
// (morton_table_16/256) are computed by bit interleaving instead of stored.
__constant__ int8_t c_omega_table[33] = {15, 14, 13, 12, 12, 11, 11, 11, 10, 10, 9, 9, 9, 8, 8, 8, 8,
                                         7,  7,  7,  6,  6,  6,  5,  5,  4,  4, 4, 3, 3, 2, 1, 0};
__constant__ int8_t c_phi_table[17][17] = {
    {2, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6},
    {1, 2, 2, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 5, 5, 6, 6},
    {1, 1, 2, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 5, 6, 6, 6},
    {1, 1, 1, 2, 2, 2, 3, 3, 4, 4, 4, 5, 5, 6, 6, 6, 6},
    {1, 1, 1, 1, 2, 2, 2, 3, 4, 4, 5, 5, 6, 6, 6, 6, 6},
    {0, 1, 1, 1, 1, 2, 2, 3, 4, 4, 5, 6, 6, 6, 6, 6, 7},
    {0, 0, 0, 0, 1, 1, 2, 2, 4, 5, 6, 6, 6, 7, 7, 7, 7},
    {0, 0, 0, 0, 0, 0, 1, 2, 4, 6, 6, 7, 7, 7, 7, 7, 7},
    {15, 15, 15, 15, 15, 15, 15, 15, 8, 8, 8, 8, 8, 8, 8, 8, 8},
    {15, 15, 15, 15, 15, 15, 14, 14, 12, 10, 9, 8, 8, 8, 8, 8, 8},
    {15, 15, 15, 15, 14, 14, 14, 13, 12, 10, 10, 9, 9, 8, 8, 8, 8},
    {15, 14, 14, 14, 14, 14, 13, 12, 12, 11, 10, 10, 9, 9, 9, 9, 8},
    {14, 14, 14, 14, 14, 13, 13, 12, 12, 11, 10, 10, 10, 9, 9, 9, 9},
    {14, 14, 14, 14, 13, 13, 12, 12, 12, 11, 11, 10, 10, 10, 9, 9, 9},
    {14, 14, 14, 13, 13, 13, 12, 12, 12, 11, 11, 10, 10, 10, 10, 9, 9},
    {14, 14, 13, 13, 13, 13, 12, 12, 12, 11, 11, 10, 10, 10, 10, 10, 9},
    {14, 13, 13, 13, 13, 12, 12, 12, 12, 11, 11, 11, 10, 10, 10, 10, 10}};
// 8x8 coherent-sampling pattern (Core.cpp:152-160)
__constant__ int c_sampling_pattern[64] = {0, 1, 2, 3, 1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2, 1, 2, 3, 0, 2, 1,
                                           0, 3, 3, 0, 1, 2, 0, 3, 2, 1, 2, 3, 0, 1, 1, 0, 3, 2, 0, 1, 2, 3,
                                           0, 3, 2, 1, 3, 2, 1, 0, 3, 2, 1, 0, 1, 0, 3, 2, 2, 1, 0, 3};

__device__ __forceinline__ uint32_t spread2(uint32_t v) { // bit k -> bit 2k (morton_table_16)
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) r |= ((v >> k) & 1u) << (2 * k);
    return r;
}
__device__ __forceinline__ uint32_t spread3(uint32_t v) { // bit k -> bit 3k (morton_table_256)
    uint32_t r = 0;
    for (int k = 0; k < 8; k++) r |= ((v >> k) & 1u) << (3 * k);
    return r;
}
// float -> int as the x86 cvttss2si instruction: out-of-range and NaN give INT_MIN
__device__ __forceinline__ int f2i_trunc(float f) {
    if (!(f >= -2147483648.0f && f < 2147483648.0f)) return (int)0x80000000;
    return (int)f;
}

// get_ray_hash (CoreRef.cpp:316-333)
__device__ __forceinline__ uint32_t ray_hash(f3 o, f3 d, const float *root_min, const float *cell) {
    int x = clampi(f2i_trunc((o.x - root_min[0]) / cell[0]), 0, 255);
    int y = clampi(f2i_trunc((o.y - root_min[1]) / cell[1]), 0, 255);
    int z = clampi(f2i_trunc((o.z - root_min[2]) / cell[2]), 0, 255);
    x = (int)spread3((uint32_t)x);
    y = (int)spread3((uint32_t)y);
    z = (int)spread3((uint32_t)z);
    const int om = (int)spread2((uint32_t)c_omega_table[clampi(f2i_trunc((1.0f + d.z) / 0.0625f), 0, 32)]);
    const int ph = (int)spread2(
        (uint32_t)c_phi_table[clampi(f2i_trunc((1.0f + d.y) / 0.125f), 0, 16)][clampi(f2i_trunc((1.0f + d.x) / 0.125f), 0, 16)]);
    return (uint32_t)((om << 25) | (ph << 24) | (y << 2) | (z << 1) | (x << 0));
}

// ------------------------------------------------------------------ ray I/O
__device__ __forceinline__ void store_ray(rh_ray *rays, rh_ray_aux *aux, uint32_t i, const RayFull &r) {
    float4 *p = reinterpret_cast<float4 *>(rays + i);
    p[0] = make_float4(r.o.x, r.o.y, r.o.z, __int_as_float(r.xy));
    p[1] = make_float4(r.d.x, r.d.y, r.d.z, __int_as_float(r.depth));
    float4 *q = reinterpret_cast<float4 *>(aux + i);
    q[0] = make_float4(r.c.x, r.c.y, r.c.z, r.ior);
    q[1] = make_float4(r.do_dx.x, r.do_dx.y, r.do_dx.z, r.dd_dx.x);
    q[2] = make_float4(r.dd_dx.y, r.dd_dx.z, r.do_dy.x, r.do_dy.y);
    q[3] = make_float4(r.do_dy.z, r.dd_dy.x, r.dd_dy.y, r.dd_dy.z);
}
__device__ __forceinline__ RayFull load_ray(const rh_ray *rays, const rh_ray_aux *aux, uint32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(rays + i);
    const float4 a = p[0], b = p[1];
    const float4 *q = reinterpret_cast<const float4 *>(aux + i);
    const float4 c0 = q[0], c1 = q[1], c2 = q[2], c3 = q[3];
    RayFull r;
    r.o = f3{a.x, a.y, a.z};
    r.xy = __float_as_int(a.w);
    r.d = f3{b.x, b.y, b.z};
    r.depth = __float_as_int(b.w);
    r.c = f3{c0.x, c0.y, c0.z};
    r.ior = c0.w;
    r.do_dx = f3{c1.x, c1.y, c1.z};
    r.dd_dx = f3{c1.w, c2.x, c2.y};
    r.do_dy = f3{c2.z, c2.w, c3.x};
    r.dd_dy = f3{c3.y, c3.z, c3.w};
    return r;
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// wave-aggregated append: one atomic per wave, slots in lane order
__device__ __forceinline__ int wave_append(bool pred, int *counter) {
    const uint64_t mask = __ballot(pred);
    if (!mask) return -1;
    const uint32_t lane = lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if ((int)lane == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader);
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    return pred ? base + __popcll(mask & lt) : -1;
}

// Block-aggregated append of per-lane predicates to `slots` queues laid out
// back to back in one array: entries of queue j come before those of queue
// j+1 within the block's range, lanes in order.  One atomic per block.  Every
// thread of the block must call it; s needs slots * waves + 1 ints of LDS.
// Returns the first index of the block's range; lane (j) goes to
// base + s[j * waves + wave] + (lanes below it in the wave with pred j).
__device__ __forceinline__ int block_append_base(int m, uint32_t slots, int *counter, int *s) {
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (uint32_t j = 0; j < slots; j++) {
        const uint64_t mask = __ballot((int)j < m);
        if (lane_id() == 0) s[j * nw + wave] = __popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < (int)slots * nw; k++) {
            const int c = s[k];
            s[k] = tot;
            tot += c;
        }
        s[slots * nw] = tot ? atomicAdd(counter, tot) : 0;
    }
    __syncthreads();
    return s[slots * nw];
}

__device__ __forceinline__ int lanes_below(uint64_t mask) {
    const uint32_t lane = lane_id();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    return __popcll(mask & lt);
}

__device__ __forceinline__ void wave_add_u64(unsigned long long *ctr, uint32_t v) {
    // reduce across the wave, one atomic
    uint32_t s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(ctr, (unsigned long long)s);
}

__device__ __forceinline__ void wave_add_i64(unsigned long long *ctr, int v) {
    // signed variant: the counter wraps modulo 2^64, so adding -k subtracts k
    int s = v;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane_id() == 0 && s) atomicAdd(ctr, (unsigned long long)(long long)s);
}

struct PassConst {
    rh_camera cam;
    int rect[4];
    int w, h;
    int iteration;
    float fov_k, aspect;
    float root_min[3], cell_size[3];
    float mix_k;
};

// ------------------------------------------------------------------ kernels

// GeneratePrimaryRays (CoreRef.cpp:427-505) for the pixel (x, y)
__device__ __forceinline__ RayFull gen_primary_ray(const PassConst &pc, const float *__restrict__ halton, int x, int y) {
    const rh_camera &cam = pc.cam;
    const f3 cam_origin = ld3(cam.origin), fwd = ld3(cam.fwd), side = ld3(cam.side), up = ld3(cam.up);
    const float focus_distance = cam.focus_distance;
    const float k = pc.aspect, fov_k = pc.fov_k;
    const float fw = (float)pc.w, fh = (float)pc.h;

    auto get_pix_dir = [&](float px, float py, f3 origin) {
        f3 p = f3{2 * fov_k * px / fw - fov_k, 2 * fov_k * (-py) / fh + fov_k, focus_distance};
        p = cam_origin + (k * p.x) * side + p.y * up + p.z * fwd;
        return normalize(p - origin);
    };

    const int index = y * pc.w + x;
    const int hi = (pc.iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT;
    float _x = (float)x, _y = (float)y;
    const int hash_val = hash(index);
    if (cam.filter == RH_FILTER_TENT) {
        float rx = fract_modf(halton[hi + 0] + construct_float((uint32_t)hash_val));
        if (rx < 0.5f) rx = __builtin_sqrtf(2.0f * rx) - 1.0f;
        else rx = 1.0f - __builtin_sqrtf(2.0f - 2 * rx);
        float ry = fract_modf(halton[hi + 1] + construct_float((uint32_t)hash(hash_val)));
        if (ry < 0.5f) ry = __builtin_sqrtf(2.0f * ry) - 1.0f;
        else ry = 1.0f - __builtin_sqrtf(2.0f - 2.0f * ry);
        _x += 0.5f + rx;
        _y += 0.5f + ry;
    } else {
        _x += fract_modf(halton[hi + 0] + construct_float((uint32_t)hash_val));
        _y += fract_modf(halton[hi + 1] + construct_float((uint32_t)hash(hash_val)));
    }
    const float ff1 = cam.focus_factor * (-0.5f + fract_modf(halton[hi + 2 + 0] + construct_float((uint32_t)hash_val)));
    const float ff2 = cam.focus_factor * (-0.5f + fract_modf(halton[hi + 2 + 1] + construct_float((uint32_t)hash(hash_val))));
    const f3 origin = cam_origin + side * ff1 + up * ff2;
    const f3 d = get_pix_dir(_x, _y, origin);
    const f3 dx = get_pix_dir(_x + 1, _y, origin), dy = get_pix_dir(_x, _y + 1, origin);

    RayFull r;
    r.o = origin;
    r.d = d;
    r.c = f3{1.0f, 1.0f, 1.0f};
    r.ior = 1.0f;
    r.do_dx = f3{0, 0, 0};
    r.dd_dx = dx - d;
    r.do_dy = f3{0, 0, 0};
    r.dd_dy = dy - d;
    r.xy = (x << 16) | y;
    r.depth = 0;
    return r;
}

// one lane per pixel of the region
__global__ void __launch_bounds__(kBlock) k_gen_primary(PassConst pc, const float *__restrict__ halton,
                                                        rh_ray *__restrict__ rays, rh_ray_aux *__restrict__ aux,
                                                        int *__restrict__ count) {
    const int rw = pc.rect[2], rh_ = pc.rect[3];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *count = rw * rh_;
    if (i >= rw * rh_) return;
    const int x = pc.rect[0] + i % rw, y = pc.rect[1] + i / rw;
    store_ray(rays, aux, (uint32_t)i, gen_primary_ray(pc, halton, x, y));
}

struct TraceArgs {
    Scene sc;
    const rh_ray *rays;
    const int *count; // device count (nullptr -> n)
    int n;
    const float *tmax; // optional per-ray tmax
    rh_hit *hits;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    unsigned long long *ray_counter; // optional: rays traced
    unsigned long long *overflow;    // stack overflow events
    unsigned long long *visits;      // Count mode: node visits, leaf triangle tests, instance tests, rays
};

template <bool AnyHit, bool Count> __global__ void __launch_bounds__(kBlock) k_trace(TraceArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const int n = a.count ? *a.count : a.n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = (int)i < n;
    if (i == 0 && a.ray_counter) atomicAdd(a.ray_counter, (unsigned long long)n);
    if (!active) return;

    const float4 *p = reinterpret_cast<const float4 *>(a.rays + i);
    const float4 r0 = p[0], r1 = p[1];
    Hit hit;
    hit.mask = 0;
    hit.obj = -1;
    hit.prim = -1;
    hit.t = a.tmax ? a.tmax[i] : RH_MAX_DIST;
    hit.u = hit.v = 0.0f;

    Stack st;
    st.lds = lds_stack;
    st.spill = a.spill + i;
    st.spill_stride = a.spill_stride;
    st.lds_depth = a.lds_depth;
    st.spill_depth = kSpillDepth;
    st.lane = threadIdx.x;
    st.block = blockDim.x;
    st.sp = 0;
    st.overflow = 0;
    TraceCounters tc = {0, 0, 0};
    if (a.sc.macro_root != RH_INVALID)
        traverse_fast<AnyHit>(a.sc, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, hit, st, Count ? &tc : nullptr);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    if (Count) {
        atomicAdd(a.visits + 0, (unsigned long long)tc.nodes);
        atomicAdd(a.visits + 1, (unsigned long long)tc.leaf_tris);
        atomicAdd(a.visits + 2, (unsigned long long)tc.instances);
        atomicAdd(a.visits + 3, 1ull);
    }

    float4 *h = reinterpret_cast<float4 *>(a.hits + i);
    h[0] = make_float4(__int_as_float(hit.mask), __int_as_float(hit.obj), __int_as_float(hit.prim), hit.t);
    h[1] = make_float4(hit.u, hit.v, r0.w, 0.0f);
}

struct ShadeArgs {
    Scene sc;
    const rh_ray *rays;
    const rh_ray_aux *aux;
    const rh_hit *hits;
    const int *count;
    int n;
    const rh_pass_info *per_ray_pi; // optional (parity entry point)
    rh_pass_info pi;                // pass-level info; index/rand_index derived from xy
    int w;
    const float *halton;
    float4 *temp; // framebuffer of this pass (null in the parity entry point)
    int primary;  // 1: set pixel, 0: add pixel (alpha 0)
    float4 *out_rgba; // parity entry point: per-ray colour
    rh_ray *out_rays;
    rh_ray_aux *out_aux;
    int *out_count;
    int *out_src;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    unsigned long long *shadow_counter;
    unsigned long long *overflow;
    unsigned long long *ray_counter; // path tail: secondary rays traced
    // deferred next-event estimation (null nee_ray: visibility traced inline)
    float4 *nee_ray;
    int2 *nee_aux;
    float4 *nee_term;
    uint32_t nee_slots; // slots per shaded ray (= light count)
    float4 *pend_c;     // ray throughput xyz, pixel index bits
    float4 *pend_alb;   // albedo xyz, flags bits (1: multiply by albedo, 2: primary)
    int *pend_count;    // deferred terms of the ray (0: nothing pending)
    uint32_t *shadow_queue; // compacted slot indices of the recorded light samples
    int *shadow_count;      // entries appended to shadow_queue
};

// Defer: direct light goes to the NEE stage (no traversal in this kernel).
template <bool Defer> __global__ void __launch_bounds__(kBlock) k_shade(ShadeArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const int n = a.count ? *a.count : a.n;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int)(blockIdx.x * blockDim.x) >= n) return; // whole block idle
    const bool active = (int)i < n;

    ShadeOut so;
    so.emit = false;
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    Stack st;
    st.lds = lds_stack;
    st.spill = a.spill + i;
    st.spill_stride = a.spill_stride;
    st.lds_depth = a.lds_depth;
    st.spill_depth = kSpillDepth;
    st.lane = threadIdx.x;
    st.block = blockDim.x;
    st.sp = 0;
    st.overflow = 0;
    int px = 0, py = 0;
    int nee_m = 0; // light samples recorded for this ray (deferred NEE)

    if (active) {
        const RayFull ray = load_ray(a.rays, a.aux, i);
        const float4 *hp = reinterpret_cast<const float4 *>(a.hits + i);
        const float4 h0 = hp[0], h1 = hp[1];
        Hit hit;
        hit.mask = __float_as_int(h0.x);
        hit.obj = __float_as_int(h0.y);
        hit.prim = __float_as_int(h0.z);
        hit.t = h0.w;
        hit.u = h1.x;
        hit.v = h1.y;

        px = (ray.xy >> 16) & 0xffff;
        py = ray.xy & 0xffff;
        rh_pass_info pi;
        if (a.per_ray_pi) {
            pi = a.per_ray_pi[i];
        } else {
            pi = a.pi;
            pi.index = py * a.w + px;
            pi.rand_index = pi.index;
            if (pi.settings.flags & RH_USE_COHERENT_SAMPLING) pi.rand_index = c_sampling_pattern[(py % 8) * 8 + (px % 8)];
        }
        NeeSink sink;
        NeeSink *sinkp = nullptr;
        if (Defer) {
            sink = NeeSink{a.nee_ray, a.nee_aux, a.nee_term, i * a.nee_slots, 0};
            sinkp = &sink;
        }
        shade_surface(a.sc, pi, hit, ray, a.halton, st, cnt, so, sinkp);
        if (Defer) {
            // the colour written below carries a zero direct term; the combine
            // stage adds ray.c * (sum of visible terms [* albedo]) on top
            const bool pending = so.nee_pending && sink.count > 0;
            nee_m = pending ? (int)sink.count : 0;
            a.pend_count[i] = nee_m;
            if (pending) {
                a.pend_c[i] = make_float4(ray.c.x, ray.c.y, ray.c.z, __int_as_float(py * a.w + px));
                a.pend_alb[i] = make_float4(so.albedo.x, so.albedo.y, so.albedo.z,
                                            __int_as_float(so.consider_albedo ? 1 : 0));
            }
        }

        if (a.temp) {
            float4 &dst = a.temp[(size_t)py * a.w + px];
            if (a.primary) {
                dst = make_float4(so.col.x, so.col.y, so.col.z, so.col.w);
            } else {
                float4 v = dst;
                v.x += so.col.x;
                v.y += so.col.y;
                v.z += so.col.z;
                v.w += 0.0f;
                dst = v;
            }
        }
        if (a.out_rgba) a.out_rgba[i] = make_float4(so.col.x, so.col.y, so.col.z, so.col.w);
    }
    __shared__ int s_app[kMaxDeferredLights * (kBlock / 64) + 1];
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (Defer) {
        // One shadow work item per recorded light sample.  Sample j of every
        // lane of the block goes into one run, so lanes that aim at the same
        // light sit next to each other in the queue.
        const int base = block_append_base(nee_m, a.nee_slots, a.shadow_count, s_app);
        for (uint32_t j = 0; j < a.nee_slots; j++) {
            const uint64_t mask = __ballot((int)j < nee_m);
            if ((int)j < nee_m) a.shadow_queue[base + s_app[j * nw + wave] + lanes_below(mask)] = i * a.nee_slots + j;
        }
        __syncthreads(); // s_app is reused below
    }
    {
        const bool emit = active && so.emit;
        const int base = block_append_base(emit ? 1 : 0, 1, a.out_count, s_app);
        const uint64_t mask = __ballot(emit);
        if (emit) {
            const int slot = base + s_app[wave] + lanes_below(mask);
            store_ray(a.out_rays, a.out_aux, (uint32_t)slot, so.next);
            if (a.out_src) a.out_src[slot] = (int)i;
        }
    }
    if (a.shadow_counter) wave_add_u64(a.shadow_counter, cnt.shadow_rays);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
}

// Path tail: the remaining bounces of each ray in one launch, one lane per
// path (trace, then ShadeSurface with visibility traced inline, then the
// next ray), from host bounce `first` up to `last` (exclusive).  Late bounces
// hold few rays whose cost is the latency of their longest traversals; here
// those latencies overlap across bounces instead of adding up launch after
// launch.  Every pixel still receives its path's contributions in bounce
// order, and each contribution is computed exactly as in the staged form, so
// the image is identical.
__global__ void __launch_bounds__(kBlock) k_path_tail(ShadeArgs a, int first, int last) {
    extern __shared__ uint32_t lds_stack[];
    const int n = *a.count;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int)(blockIdx.x * blockDim.x) >= n) return; // whole block idle
    const bool active = (int)i < n;
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    uint32_t traced = 0;
    Stack st;
    st.lds = lds_stack;
    st.spill = a.spill + i;
    st.spill_stride = a.spill_stride;
    st.lds_depth = a.lds_depth;
    st.spill_depth = kSpillDepth;
    st.lane = threadIdx.x;
    st.block = blockDim.x;
    st.sp = 0;
    st.overflow = 0;
    if (active) {
        RayFull ray = load_ray(a.rays, a.aux, i);
        const int px = (ray.xy >> 16) & 0xffff, py = ray.xy & 0xffff;
        rh_pass_info pi = a.pi;
        pi.index = py * a.w + px;
        pi.rand_index = pi.index;
        if (pi.settings.flags & RH_USE_COHERENT_SAMPLING) pi.rand_index = c_sampling_pattern[(py % 8) * 8 + (px % 8)];
        float4 &dst = a.temp[(size_t)py * a.w + px];
        for (int b = first; b < last; b++) {
            Hit hit;
            hit.mask = 0;
            hit.obj = -1;
            hit.prim = -1;
            hit.t = RH_MAX_DIST;
            hit.u = hit.v = 0.0f;
            st.sp = 0;
            if (a.sc.macro_root != RH_INVALID) traverse_fast<false>(a.sc, ray.o, ray.d, hit, st, nullptr);
            traced++;
            pi.bounce = b + 3;
            ShadeOut so;
            so.emit = false;
            shade_surface(a.sc, pi, hit, ray, a.halton, st, cnt, so, nullptr);
            float4 v = dst;
            v.x += so.col.x;
            v.y += so.col.y;
            v.z += so.col.z;
            v.w += 0.0f;
            dst = v;
            if (!so.emit) break;
            ray = so.next;
        }
    }
    if (a.ray_counter) wave_add_u64(a.ray_counter, traced);
    if (a.shadow_counter) wave_add_u64(a.shadow_counter, cnt.shadow_rays);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
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
    float4 *rec;        // per-lane vertex records, max_vertices x lanes, lane-minor
    uint32_t lanes;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    unsigned long long *ray_counter;
    unsigned long long *shadow_counter;
    unsigned long long *overflow;
};

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
// only run-to-run difference).
__global__ void __launch_bounds__(kBlock) k_adjoint(AdjointArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const int rw = a.pc.rect[2], rh_ = a.pc.rect[3];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = (int)i < rw * rh_;
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    uint32_t traced = 0;
    Stack st;
    st.lds = lds_stack;
    st.spill = a.spill + i;
    st.spill_stride = a.spill_stride;
    st.lds_depth = a.lds_depth;
    st.spill_depth = kSpillDepth;
    st.lane = threadIdx.x;
    st.block = blockDim.x;
    st.sp = 0;
    st.overflow = 0;
    if (active) {
        const int x = a.pc.rect[0] + (int)i % rw, y = a.pc.rect[1] + (int)i / rw;
        const int p = y * a.pc.w + x;
        const float4 dl = a.adj[p];
        if (dl.x != 0.0f || dl.y != 0.0f || dl.z != 0.0f) {
            RayFull ray = gen_primary_ray(a.pc, a.halton, x, y);
            rh_pass_info pi = a.pi;
            pi.index = p;
            pi.rand_index = p;
            if (pi.settings.flags & RH_USE_COHERENT_SAMPLING) pi.rand_index = c_sampling_pattern[(y % 8) * 8 + (x % 8)];
            const int last = pi.settings.max_total_depth;
            const bool skip_indirect = (pi.settings.flags & RH_SKIP_INDIRECT_LIGHT) != 0;
            int nv = 0;
            for (int b = -1;;) {
                Hit hit;
                hit.mask = 0;
                hit.obj = -1;
                hit.prim = -1;
                hit.t = RH_MAX_DIST;
                hit.u = hit.v = 0.0f;
                st.sp = 0;
                if (a.sc.macro_root != RH_INVALID) traverse_fast<false>(a.sc, ray.o, ray.d, hit, st, nullptr);
                traced++;
                pi.bounce = b < 0 ? 2 : b + 3;
                ShadeOut so;
                so.emit = false;
                shade_surface(a.sc, pi, hit, ray, a.halton, st, cnt, so, nullptr);
                a.rec[(size_t)nv * a.lanes + i] = make_float4(so.col.x, so.col.y, so.col.z, __int_as_float(so.alb_mat));
                nv++;
                b++;
                if (!so.emit || skip_indirect || b >= last) break;
                ray = so.next;
            }
            f3 S = f3{0.0f, 0.0f, 0.0f};
            for (int k = nv - 1; k >= 0; k--) {
                const float4 r = a.rec[(size_t)k * a.lanes + i];
                S = S + f3{r.x, r.y, r.z};
                const int m = __float_as_int(r.w);
                if (m < 0) continue;
                const float *mc = a.sc.materials[m].main_color;
                const float g[3] = {dl.x * S.x, dl.y * S.y, dl.z * S.z};
#pragma unroll
                for (int ch = 0; ch < 3; ch++) {
                    if (mc[ch] != 0.0f && g[ch] != 0.0f) atomicAdd(&a.grad[4 * m + ch], a.scale * (g[ch] / mc[ch]));
                }
            }
        }
    }
    if (a.ray_counter) wave_add_u64(a.ray_counter, traced);
    if (a.shadow_counter) wave_add_u64(a.shadow_counter, cnt.shadow_rays);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
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

struct ShadowArgs {
    Scene sc;
    const uint32_t *queue; // slot indices of the recorded light samples
    const int *count;      // entries in queue (device)
    const float4 *nee_ray;
    const int2 *nee_aux;
    uint8_t *nee_vis;
    const float *halton;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    unsigned long long *shadow_counter;
    unsigned long long *overflow;
};

// Visibility of each deferred light sample: ComputeVisibility
// (CoreRef.cpp:1801-1896), any-hit traversal stepping through transparent
// surfaces.  One lane per queue entry; the grid covers every slot and lanes
// past the device count exit at once.
__global__ void __launch_bounds__(kBlock) k_shadow(ShadowArgs a) {
    extern __shared__ uint32_t lds_stack[];
    const int n = *a.count;
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if ((int)(blockIdx.x * blockDim.x) >= n) return; // whole block idle
    if (q == 0 && a.shadow_counter) atomicAdd(a.shadow_counter, (unsigned long long)n);
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    if ((int)q < n) {
        const uint32_t k = a.queue[q];
        const float4 r = a.nee_ray[k];
        const int2 x = a.nee_aux[k];
        const rh_light &l = a.sc.lights[__float_as_uint(r.w)];
        Stack st;
        st.lds = lds_stack;
        st.spill = a.spill + q;
        st.spill_stride = a.spill_stride;
        st.lds_depth = a.lds_depth;
        st.spill_depth = kSpillDepth;
        st.lane = threadIdx.x;
        st.block = blockDim.x;
        st.sp = 0;
        st.overflow = 0;
        const float vis = compute_visibility(a.sc, f3{r.x, r.y, r.z}, ld3(l.pos), a.halton, x.x, x.y, st, cnt);
        a.nee_vis[k] = vis != 0.0f ? 1 : 0;
        if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    }
    // one segment per entry is counted above; add the difference (segments
    // past the first through transparent surfaces, or none at all when the
    // light is within HIT_EPS)
    const int extra = (int)q < n ? (int)cnt.shadow_rays - 1 : 0;
    if (a.shadow_counter && __ballot(extra != 0)) wave_add_i64(a.shadow_counter, extra);
}

// Adds each shaded ray's deferred direct light: ray.c * (sum of visible terms
// in light order [* albedo]), to its pixel (or to the per-ray output).
__global__ void __launch_bounds__(kBlock) k_nee_combine(const int *count, int n_static, uint32_t slots,
                                                        const int *__restrict__ pend_count,
                                                        const float4 *__restrict__ pend_c,
                                                        const float4 *__restrict__ pend_alb,
                                                        const float4 *__restrict__ nee_term,
                                                        const uint8_t *__restrict__ nee_vis, float4 *temp,
                                                        float4 *out_rgba) {
    const int n = count ? *count : n_static;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = pend_count[i];
    if (!m) return;
    f3 col = f3{0.0f, 0.0f, 0.0f};
    const uint32_t base = (uint32_t)i * slots;
    for (int j = 0; j < m; j++) {
        if (nee_vis[base + j]) {
            const float4 t = nee_term[base + j];
            col = col + f3{t.x, t.y, t.z};
        }
    }
    const float4 pc = pend_c[i], pa = pend_alb[i];
    if (__float_as_int(pa.w) & 1) col = col * f3{pa.x, pa.y, pa.z};
    const f3 add = f3{pc.x * col.x, pc.y * col.y, pc.z * col.z};
    float4 *dst = temp ? &temp[__float_as_int(pc.w)] : &out_rgba[i];
    float4 v = *dst;
    v.x += add.x;
    v.y += add.y;
    v.z += add.z;
    *dst = v;
}

// MixWith + tonemap (FramebufferRef.cpp:80-86, RendererRef.cpp:301-324)
__global__ void k_mix_tonemap(const float4 *__restrict__ temp, float4 *__restrict__ clean, float4 *__restrict__ final_,
                              int x0, int y0, int rw, int rh_, int w, float k, int dtype, float gamma, uint32_t flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rw * rh_) return;
    const int x = x0 + i % rw, y = y0 + i / rw;
    const size_t p = (size_t)y * w + x;
    const float4 t = temp[p];
    float4 c = clean[p];
    c.x += (t.x - c.x) * k;
    c.y += (t.y - c.y) * k;
    c.z += (t.z - c.z) * k;
    c.w += (t.w - c.w) * k;
    clean[p] = c;
    float v[4] = {c.x, c.y, c.z, c.w};
    if (dtype == RH_DEVICE_SRGB) {
        for (int j = 0; j < 3; j++) {
            if (v[j] > 0.0031308f) v[j] = rpow(1.055f * v[j], (1.0f / 2.4f)) - 0.055f;
            else v[j] = 12.92f * v[j];
        }
    }
    if (gamma != 1.0f) {
        const float g = 1.0f / gamma;
        for (int j = 0; j < 4; j++) v[j] = rpow(v[j], g);
    }
    if (flags & RH_CLAMP) {
        for (int j = 0; j < 4; j++) v[j] = v[j] > 1.0f ? 1.0f : (v[j] < 0.0f ? 0.0f : v[j]);
    }
    final_[p] = make_float4(v[0], v[1], v[2], v[3]);
}

__global__ void k_fill(float4 *buf, size_t n, float4 v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) buf[i] = v;
}

__global__ void k_set_int(int *p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void k_copy_int(int *dst, const int *src) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
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

// Radix pass 1/3: per-tile digit histogram, stored digit-major.
__global__ void __launch_bounds__(kBlock) k_radix_hist(const uint32_t *__restrict__ keys, const int *__restrict__ count,
                                                       int n_static, int shift, uint32_t *__restrict__ hist,
                                                       int num_tiles) {
    __shared__ uint32_t h[256];
    const int n = count ? *count : n_static;
    h[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kSortTile;
    if (base < n) {
        for (int r = 0; r < kSortItems; r++) {
            const int i = base + r * kBlock + threadIdx.x;
            if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
        }
    }
    __syncthreads();
    hist[threadIdx.x * num_tiles + blockIdx.x] = h[threadIdx.x];
}

// Wave-level inclusive scan (64 lanes) with cross-lane shuffles.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const uint32_t lane = lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(v, off);
        if (lane >= (uint32_t)off) v += u;
    }
    return v;
}

// Radix pass 2a/3: one block per digit scans that digit's row of per-tile
// counts (exclusive, in place) and records the digit's total.
__global__ void __launch_bounds__(kBlock) k_radix_scan_rows(uint32_t *__restrict__ hist, int num_tiles,
                                                            uint32_t *__restrict__ totals) {
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t carry;
    uint32_t *row = hist + (size_t)blockIdx.x * num_tiles;
    const int t = threadIdx.x, wave = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < num_tiles; base += kBlock) {
        const int i = base + t;
        const uint32_t v = i < num_tiles ? row[i] : 0u;
        const uint32_t inc = wave_inclusive_scan(v);
        if ((t & 63) == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wave; w++) before += wsum[w];
        if (i < num_tiles) row[i] = before + inc - v;
        __syncthreads();
        if (t == kBlock - 1) carry = before + inc;
        __syncthreads();
    }
    if (t == 0) totals[blockIdx.x] = carry;
}

// Radix pass 2b/3: adds each digit's base (sum of the smaller digits' totals).
__global__ void __launch_bounds__(kBlock) k_radix_scan_bases(uint32_t *__restrict__ hist, int num_tiles,
                                                             const uint32_t *__restrict__ totals) {
    __shared__ uint32_t base;
    const int d = blockIdx.x;
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < d; k++) s += totals[k];
        base = s;
    }
    __syncthreads();
    uint32_t *row = hist + (size_t)d * num_tiles;
    for (int i = threadIdx.x; i < num_tiles; i += kBlock) row[i] += base;
}

// Radix pass 3/3: stable scatter.  Within a round of 256 keys, each wave finds
// the lanes sharing its digit with eight ballots (one per digit bit), ranks
// them by lane order, and the per-wave digit counts are prefix-summed across
// the block's four waves in LDS; rounds are processed in order, so equal keys
// keep their input order.
__global__ void __launch_bounds__(kBlock) k_radix_scatter(const uint32_t *__restrict__ keys_in,
                                                          const uint32_t *__restrict__ vals_in,
                                                          uint32_t *__restrict__ keys_out,
                                                          uint32_t *__restrict__ vals_out, const int *__restrict__ count,
                                                          int n_static, int shift, const uint32_t *__restrict__ hist,
                                                          int num_tiles) {
    __shared__ uint32_t run[256];          // running digit offsets of this tile
    __shared__ uint32_t wcount[4][256];    // per-wave digit counts of this round
    const int n = count ? *count : n_static;
    const int base = blockIdx.x * kSortTile;
    if (base >= n) return;
    const int t = threadIdx.x, wave = t >> 6;
    const uint32_t lane = lane_id();
    run[t] = hist[t * num_tiles + blockIdx.x];
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

    for (int r = 0; r < kSortItems; r++) {
        const int i = base + r * kBlock + t;
        const bool active = i < n;
        uint32_t key = 0, val = 0, digit = 0;
        if (active) {
            key = keys_in[i];
            val = vals_in[i];
            digit = (key >> shift) & 0xffu;
        }
        wcount[0][t] = wcount[1][t] = wcount[2][t] = wcount[3][t] = 0;
        __syncthreads();
        uint64_t peers = __ballot(active);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (active && rank == 0) wcount[wave][digit] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t pos = 0;
        if (active) {
            uint32_t before = 0;
            for (int w2 = 0; w2 < wave; w2++) before += wcount[w2][digit];
            pos = run[digit] + before + rank;
        }
        __syncthreads();
        // advance the running offsets by this round's totals
        run[t] += wcount[0][t] + wcount[1][t] + wcount[2][t] + wcount[3][t];
        if (active) {
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock) k_gather(const rh_ray *__restrict__ rays_in, const rh_ray_aux *__restrict__ aux_in,
                                                   const uint32_t *__restrict__ perm, const int *__restrict__ count,
                                                   int n_static, rh_ray *__restrict__ rays_out,
                                                   rh_ray_aux *__restrict__ aux_out) {
    const int n = count ? *count : n_static;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    rays_out[i] = rays_in[j];
    aux_out[i] = aux_in[j];
}

} // namespace

// ====================================================================== host

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct rh_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int w = 0, h = 0;
    size_t max_rays = 0;

    // scene
    std::vector<DevBuf> scene_bufs;
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
    float *color_stage = nullptr;
    size_t color_stage_cap = 0; // floats
    uint32_t n_materials = 0;
    int *counters = nullptr; // [0], [1]: ray queue counts, [2]: shadow queue count
    unsigned long long *stats_dev = nullptr; // [0] primary [1] secondary [2] shadow [3] sorted [4] overflow
    float *halton = nullptr;

    // profiling: event pairs around trace / shade / sort launches, read back on demand
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<std::pair<int, std::pair<size_t, size_t>>> ev_marks; // kind, (start, stop)
    unsigned long long *visits = nullptr;                             // [nodes, tris, instances, rays]

    // options
    int kernel_timing = 0;
    int count_visits = 0;
    int sort_rays = 1;
    int sort_bits = 32; // key bits the in-pass ray sort orders by (top bits)
    int path_tail = 2;  // first bounce run by the path-tail kernel (-1: none)
    int stage_timing = 0;
    int lds_depth = 24;
    rh_stats stats = {};
};

namespace {

int fail(hipError_t e, const char *what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return -(int)e;
}
#define RH_CHECK(call)                                                                                                 \
    do {                                                                                                               \
        hipError_t _e = (call);                                                                                        \
        if (_e != hipSuccess) return fail(_e, #call);                                                                  \
    } while (0)

int arg_error(const char *msg) {
    g_err = msg;
    return -1001;
}

void free_ptr(void *p) {
    if (p) (void)hipFree(p);
}

int alloc_frame(rh_context *c, int w, int h) {
    const size_t np = (size_t)w * h;
    free_ptr(c->temp);
    free_ptr(c->clean);
    free_ptr(c->final_);
    for (int k = 0; k < 2; k++) {
        free_ptr(c->rays[k]);
        free_ptr(c->aux[k]);
        free_ptr(c->keys[k]);
        free_ptr(c->vals[k]);
    }
    free_ptr(c->hits);
    free_ptr(c->hist);
    free_ptr(c->spill);
    c->temp = c->clean = c->final_ = nullptr;
    c->hits = nullptr;
    c->hist = nullptr;
    c->spill = nullptr;
    for (int k = 0; k < 2; k++) c->rays[k] = nullptr, c->aux[k] = nullptr, c->keys[k] = nullptr, c->vals[k] = nullptr;
    c->w = w;
    c->h = h;
    c->max_rays = std::max<size_t>(np, 1);
    RH_CHECK(hipMalloc(&c->temp, np * sizeof(float4)));
    RH_CHECK(hipMalloc(&c->clean, np * sizeof(float4)));
    RH_CHECK(hipMalloc(&c->final_, np * sizeof(float4)));
    RH_CHECK(hipMemsetAsync(c->temp, 0, np * sizeof(float4), c->stream));
    RH_CHECK(hipMemsetAsync(c->clean, 0, np * sizeof(float4), c->stream));
    RH_CHECK(hipMemsetAsync(c->final_, 0, np * sizeof(float4), c->stream));
    return 0;
}

int ensure_nee(rh_context *c);

int ensure_rays(rh_context *c, size_t n) {
    if (n <= c->max_rays && c->rays[0]) return ensure_nee(c);
    c->max_rays = std::max(n, c->max_rays);
    const size_t m = c->max_rays;
    for (int k = 0; k < 2; k++) {
        free_ptr(c->rays[k]);
        free_ptr(c->aux[k]);
        free_ptr(c->keys[k]);
        free_ptr(c->vals[k]);
        RH_CHECK(hipMalloc(&c->rays[k], m * sizeof(rh_ray)));
        RH_CHECK(hipMalloc(&c->aux[k], m * sizeof(rh_ray_aux)));
        RH_CHECK(hipMalloc(&c->keys[k], m * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&c->vals[k], m * sizeof(uint32_t)));
    }
    free_ptr(c->hits);
    free_ptr(c->hist);
    RH_CHECK(hipMalloc(&c->hits, m * sizeof(rh_hit)));
    const size_t tiles = (m + kSortTile - 1) / kSortTile;
    // per-tile digit counts (digit-major) followed by the 256 digit totals
    RH_CHECK(hipMalloc(&c->hist, (256 * tiles + 256) * sizeof(uint32_t)));
    return ensure_nee(c);
}

// Spill slices are indexed by lane (ray or shadow slot): the array holds
// kSpillDepth entries for each of spill_cap lanes, stride spill_cap.
int ensure_spill(rh_context *c, size_t lanes) {
    if (lanes <= c->spill_cap && c->spill) return 0;
    free_ptr(c->spill);
    c->spill = nullptr;
    c->spill_cap = std::max(lanes, c->spill_cap);
    RH_CHECK(hipMalloc(&c->spill, (size_t)kSpillDepth * c->spill_cap * sizeof(uint32_t)));
    return 0;
}

// Buffers of the deferred NEE stage for the current scene and frame size.
int ensure_nee(rh_context *c) {
    const size_t m = c->max_rays;
    if (!c->nee_slots) return ensure_spill(c, m);
    const size_t cap = m * c->nee_slots;
    if (cap > c->nee_cap || !c->nee_ray) {
        free_ptr(c->nee_ray);
        free_ptr(c->nee_term);
        free_ptr(c->nee_aux);
        free_ptr(c->nee_vis);
        free_ptr(c->pend_c);
        free_ptr(c->pend_alb);
        free_ptr(c->pend_count);
        free_ptr(c->shadow_queue);
        c->nee_cap = cap;
        RH_CHECK(hipMalloc(&c->nee_ray, cap * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->nee_term, cap * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->nee_aux, cap * sizeof(int2)));
        RH_CHECK(hipMalloc(&c->nee_vis, cap * sizeof(uint8_t)));
        RH_CHECK(hipMalloc(&c->pend_c, m * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->pend_alb, m * sizeof(float4)));
        RH_CHECK(hipMalloc(&c->pend_count, m * sizeof(int)));
        RH_CHECK(hipMalloc(&c->shadow_queue, cap * sizeof(uint32_t)));
    }
    return ensure_spill(c, std::max(m, cap));
}

template <typename T> int upload(rh_context *c, const T *src, size_t count, const T **dst) {
    *dst = nullptr;
    if (!src || !count) return 0;
    DevBuf b;
    b.bytes = count * sizeof(T);
    RH_CHECK(hipMalloc(&b.p, b.bytes));
    RH_CHECK(hipMemcpyAsync(b.p, src, b.bytes, hipMemcpyHostToDevice, c->stream));
    c->scene_bufs.push_back(b);
    *dst = static_cast<const T *>(b.p);
    return 0;
}

size_t lds_bytes(const rh_context *c) { return (size_t)c->lds_depth * kBlock * sizeof(uint32_t); }

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
            c->ev_marks.push_back({kind, {a, b}});
        }
    }
};

int launch_trace(rh_context *c, const TraceArgs &a, int grid_rays, bool any_hit) {
    const int blocks = std::max(1, (grid_rays + kBlock - 1) / kBlock);
    KernelMark km(c, 0);
    const bool count = c->count_visits && a.visits;
    if (any_hit) {
        if (count) hipLaunchKernelGGL((k_trace<true, true>), dim3(blocks), dim3(kBlock), lds_bytes(c), c->stream, a);
        else hipLaunchKernelGGL((k_trace<true, false>), dim3(blocks), dim3(kBlock), lds_bytes(c), c->stream, a);
    } else {
        if (count) hipLaunchKernelGGL((k_trace<false, true>), dim3(blocks), dim3(kBlock), lds_bytes(c), c->stream, a);
        else hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kBlock), lds_bytes(c), c->stream, a);
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
    s.shadow_counter = c->stats_dev + 2;
    s.overflow = c->stats_dev + 4;
    const size_t lanes = (size_t)grid_rays * a.nee_slots;
    {
        KernelMark km(c, 3);
        hipLaunchKernelGGL(k_shadow, dim3((unsigned)std::max<size_t>(1, (lanes + kBlock - 1) / kBlock)), dim3(kBlock),
                           lds_bytes(c), c->stream, s);
        RH_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_nee_combine, dim3(blocks), dim3(kBlock), 0, c->stream, a.count, a.n, a.nee_slots,
                       a.pend_count, a.pend_c, a.pend_alb, a.nee_term, c->nee_vis, a.temp, a.out_rgba);
    RH_CHECK(hipGetLastError());
    return 0;
}

// Stable LSB radix sort of (key, val) pairs in keys[0]/vals[0] by the top
// `bits` bits of the key (8, 16, 24 or 32: one 8-bit pass each).  Returns the
// buffer index (0 or 1) that holds the sorted pairs.
int radix_sort(rh_context *c, const int *count, int max_n, int bits, int *result) {
    const int tiles = std::max(1, (max_n + kSortTile - 1) / kSortTile);
    KernelMark km(c, 2);
    int src = 0;
    for (int shift = 32 - bits; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, c->stream, c->keys[src], count, max_n, shift,
                           c->hist, tiles);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(kBlock), 0, c->stream, c->hist, tiles,
                           c->hist + (size_t)256 * tiles);
        hipLaunchKernelGGL(k_radix_scan_bases, dim3(256), dim3(kBlock), 0, c->stream, c->hist, tiles,
                           c->hist + (size_t)256 * tiles);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, c->stream, c->keys[src], c->vals[src],
                           c->keys[src ^ 1], c->vals[src ^ 1], count, max_n, shift, c->hist, tiles);
        src ^= 1;
    }
    RH_CHECK(hipGetLastError());
    *result = src;
    return 0;
}

PassConst make_pass_const(rh_context *c, const rh_pass_desc *p) {
    PassConst pc;
    memset(&pc, 0, sizeof(pc));
    pc.cam = p->cam;
    memcpy(pc.rect, p->rect, sizeof(pc.rect));
    pc.w = c->w;
    pc.h = c->h;
    pc.iteration = p->iteration;
    // host-side float math as GeneratePrimaryRays (CoreRef.cpp:430-431)
    pc.aspect = float(c->w) / c->h;
    pc.fov_k = std::tan(0.5f * p->cam.fov * RH_PI / 180.0f) * p->cam.focus_distance;
    memcpy(pc.root_min, p->root_min, sizeof(pc.root_min));
    memcpy(pc.cell_size, p->cell_size, sizeof(pc.cell_size));
    pc.mix_k = 1.0f / p->iteration;
    return pc;
}

struct StageTimer {
    rh_context *c;
    hipEvent_t ev[8];
    int n = 0;
    explicit StageTimer(rh_context *ctx) : c(ctx) {
        if (c->stage_timing)
            for (auto &e : ev) (void)hipEventCreate(&e);
    }
    void mark() {
        if (c->stage_timing && n < 8) (void)hipEventRecord(ev[n++], c->stream);
    }
    float between(int a, int b) {
        float ms = 0;
        if (c->stage_timing && b < n) (void)hipEventElapsedTime(&ms, ev[a], ev[b]);
        return ms;
    }
    ~StageTimer() {
        if (c->stage_timing)
            for (auto &e : ev) (void)hipEventDestroy(e);
    }
};

} // namespace

extern "C" {

const char *rh_last_error(void) { return g_err.c_str(); }

int rh_create(int device, rh_context **out) {
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        g_err = "no HIP device available";
        return -1002;
    }
    if (device < 0) {
        RH_CHECK(hipGetDevice(&device));
    }
    if (device >= ndev) return arg_error("device index out of range");
    RH_CHECK(hipSetDevice(device));
    auto *c = new rh_context;
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        g_err = "hipStreamCreate failed";
        return -1003;
    }
    RH_CHECK(hipMalloc(&c->counters, 16 * sizeof(int)));
    RH_CHECK(hipMalloc(&c->stats_dev, 8 * sizeof(unsigned long long)));
    RH_CHECK(hipMemsetAsync(c->counters, 0, 16 * sizeof(int), c->stream));
    RH_CHECK(hipMemsetAsync(c->stats_dev, 0, 8 * sizeof(unsigned long long), c->stream));
    RH_CHECK(hipMalloc(&c->visits, 4 * sizeof(unsigned long long)));
    RH_CHECK(hipMemsetAsync(c->visits, 0, 4 * sizeof(unsigned long long), c->stream));
    RH_CHECK(hipMalloc(&c->halton, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float)));
    RH_CHECK(hipMemsetAsync(c->halton, 0, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float), c->stream));
    *out = c;
    return 0;
}

void rh_destroy(rh_context *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : c->scene_bufs) free_ptr(b.p);
    free_ptr(c->temp);
    free_ptr(c->clean);
    free_ptr(c->final_);
    for (int k = 0; k < 2; k++) {
        free_ptr(c->rays[k]);
        free_ptr(c->aux[k]);
        free_ptr(c->keys[k]);
        free_ptr(c->vals[k]);
    }
    free_ptr(c->hits);
    free_ptr(c->hist);
    free_ptr(c->spill);
    free_ptr(c->nee_ray);
    free_ptr(c->nee_term);
    free_ptr(c->nee_aux);
    free_ptr(c->nee_vis);
    free_ptr(c->pend_c);
    free_ptr(c->pend_alb);
    free_ptr(c->pend_count);
    free_ptr(c->shadow_queue);
    free_ptr(c->counters);
    free_ptr(c->stats_dev);
    free_ptr(c->halton);
    free_ptr(c->visits);
    for (auto &e : c->ev_pool) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int rh_resize(rh_context *c, int w, int h) {
    if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return arg_error("framebuffer size must be 1..65535 per side");
    RH_CHECK(hipSetDevice(c->device));
    int rc = alloc_frame(c, w, h);
    if (rc) return rc;
    return ensure_rays(c, (size_t)w * h);
}

int rh_clear(rh_context *c, const float rgba[4]) {
    const size_t n = (size_t)c->w * c->h;
    if (!n) return 0;
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, c->clean, n,
                       make_float4(rgba[0], rgba[1], rgba[2], rgba[3]));
    RH_CHECK(hipGetLastError());
    return 0;
}

int rh_upload_scene(rh_context *c, const rh_scene_view *v) {
    RH_CHECK(hipSetDevice(c->device));
    RH_CHECK(hipStreamSynchronize(c->stream));
    for (auto &b : c->scene_bufs) free_ptr(b.p);
    c->scene_bufs.clear();
    Scene s = {};
    int rc = 0;
    rc |= upload(c, v->nodes, v->nodes_count, &s.nodes);
    rc |= upload(c, v->tris, v->tris_count, &s.tris);
    rc |= upload(c, v->tri_indices, v->tri_indices_count, &s.tri_indices);
    rc |= upload(c, v->meshes, v->meshes_count, &s.meshes);
    rc |= upload(c, v->mesh_instances, v->mesh_instances_count, &s.mesh_instances);
    rc |= upload(c, v->mi_indices, v->mi_indices_count, &s.mi_indices);
    rc |= upload(c, v->transforms, v->transforms_count, &s.transforms);
    rc |= upload(c, v->vertices, v->vertices_count, &s.vertices);
    rc |= upload(c, v->vtx_indices, v->vtx_indices_count, &s.vtx_indices);
    rc |= upload(c, v->materials, v->materials_count, &s.materials);
    rc |= upload(c, v->textures, v->textures_count, &s.textures);
    rc |= upload(c, v->tex_mips, (size_t)v->textures_count * RH_NUM_MIP_LEVELS, &s.tex_mips);
    rc |= upload(c, v->texels, v->texels_count, &s.texels);
    rc |= upload(c, v->lights, v->lights_count, &s.lights);
    rc |= upload(c, v->li_indices, v->li_indices_count, &s.li_indices);
    if (rc) return rc;
    if (v->env) s.env = *v->env;
    s.macro_root = v->macro_nodes_root;
    s.light_root = v->light_nodes_root;
    // light-tree leaves in the reference's depth-first order (right child first)
    std::vector<rh_bvh_node> leaves;
    if (s.light_root != RH_INVALID && v->nodes && s.light_root < v->nodes_count) {
        std::vector<uint32_t> st{s.light_root};
        while (!st.empty()) {
            const uint32_t cur = st.back();
            st.pop_back();
            if (cur >= v->nodes_count) return arg_error("light tree node outside the node array");
            const rh_bvh_node &n = v->nodes[cur];
            if (n.prim_index & RH_LEAF_NODE_BIT) {
                leaves.push_back(n);
            } else {
                st.push_back(n.prim_index);
                st.push_back(n.prim_count & RH_RIGHT_CHILD_BITS);
            }
        }
    }
    rc = upload(c, leaves.data(), leaves.size(), &s.light_leaves);
    if (rc) return rc;
    s.light_leaf_count = (uint32_t)leaves.size();
    // deferred NEE keeps one slot per light for every shaded ray; scenes with
    // many lights trace visibility inline instead
    c->nee_slots = (v->lights_count >= 1 && v->lights_count <= kMaxDeferredLights) ? v->lights_count : 0;
    if (s.macro_root != RH_INVALID && (!v->nodes || s.macro_root >= v->nodes_count))
        return arg_error("macro tree root outside the node array");
    c->sc = s;
    c->n_materials = (uint32_t)v->materials_count;
    c->have_scene = true;
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_set_halton(rh_context *c, const float *seq, size_t count) {
    if (count < (size_t)RH_HALTON_COUNT * RH_HALTON_SEQ_LEN) return arg_error("Halton table needs 256 x 24 floats");
    RH_CHECK(hipMemcpyAsync(c->halton, seq, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float), hipMemcpyHostToDevice,
                            c->stream));
    // the host table may change right after this call returns
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_render_pass(rh_context *c, const rh_pass_desc *p) {
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

    StageTimer tm(c);
    tm.mark(); // 0
    int *cnt[2] = {c->counters + 0, c->counters + 1};
    hipLaunchKernelGGL(k_gen_primary, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, pc, c->halton,
                       c->rays[0], c->aux[0], cnt[0]);
    RH_CHECK(hipGetLastError());
    tm.mark(); // 1
    rc = launch_trace(c, trace_args(c, c->rays[0], cnt[0], n, c->hits, nullptr, c->stats_dev + 0), n, false);
    if (rc) return rc;
    tm.mark(); // 2

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
    sa.temp = c->temp;
    sa.primary = 1;
    sa.out_rays = c->rays[1];
    sa.out_aux = c->aux[1];
    sa.out_count = cnt[1];
    rc = launch_shade(c, sa, n);
    if (rc) return rc;
    tm.mark(); // 3

    float t_sort = 0, t_trace = 0, t_shade = 0;
    int q = 1; // queue holding the rays to trace next
    const bool skip_indirect = (pi.settings.flags & RH_SKIP_INDIRECT_LIGHT) != 0;
    const int sort_on = c->sort_rays && p->sort_rays;
    for (int bounce = 0; bounce < pi.settings.max_total_depth && !skip_indirect; bounce++) {
        if (c->path_tail >= 0 && bounce >= c->path_tail) {
            // the remaining bounces, one lane per path
            ShadeArgs tb = sa;
            tb.rays = c->rays[q];
            tb.aux = c->aux[q];
            tb.count = cnt[q];
            tb.primary = 0;
            tb.nee_ray = nullptr;
            tb.nee_aux = nullptr;
            tb.nee_term = nullptr;
            tb.ray_counter = c->stats_dev + 1;
            KernelMark km(c, 4);
            hipLaunchKernelGGL(k_path_tail, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), lds_bytes(c), c->stream, tb,
                               bounce, (int)pi.settings.max_total_depth);
            RH_CHECK(hipGetLastError());
            break;
        }
        StageTimer bt(c);
        bt.mark();
        if (sort_on) {
            hipLaunchKernelGGL(k_ray_keys, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->rays[q],
                               cnt[q], n, pc, c->keys[0], c->vals[0]);
            int sorted = 0;
            rc = radix_sort(c, cnt[q], n, c->sort_bits, &sorted);
            if (rc) return rc;
            hipLaunchKernelGGL(k_gather, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, c->rays[q],
                               c->aux[q], c->vals[sorted], cnt[q], n, c->rays[q ^ 1], c->aux[q ^ 1]);
            RH_CHECK(hipGetLastError());
            // the sorted rays now sit in the other buffer; move the count along
            hipLaunchKernelGGL(k_copy_int, dim3(1), dim3(1), 0, c->stream, cnt[q ^ 1], cnt[q]);
            q ^= 1;
        }
        bt.mark();
        rc = launch_trace(c, trace_args(c, c->rays[q], cnt[q], n, c->hits, nullptr, c->stats_dev + 1), n, false);
        if (rc) return rc;
        bt.mark();
        hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, c->stream, cnt[q ^ 1], 0);
        ShadeArgs sb = sa;
        sb.rays = c->rays[q];
        sb.aux = c->aux[q];
        sb.count = cnt[q];
        sb.pi.bounce = bounce + 3;
        sb.primary = 0;
        sb.out_rays = c->rays[q ^ 1];
        sb.out_aux = c->aux[q ^ 1];
        sb.out_count = cnt[q ^ 1];
        rc = launch_shade(c, sb, n);
        if (rc) return rc;
        bt.mark();
        if (c->stage_timing) {
            RH_CHECK(hipStreamSynchronize(c->stream));
            t_sort += bt.between(0, 1);
            t_trace += bt.between(1, 2);
            t_shade += bt.between(2, 3);
        }
        q ^= 1;
    }

    hipLaunchKernelGGL(k_mix_tonemap, dim3((n + 255) / 256), dim3(256), 0, c->stream, c->temp, c->clean, c->final_, rx,
                       ry, rw, rhh, c->w, pc.mix_k, p->cam.dtype, p->cam.gamma, p->cam.pass_settings.flags);
    RH_CHECK(hipGetLastError());
    if (c->stage_timing) {
        RH_CHECK(hipStreamSynchronize(c->stream));
        c->stats.time_primary_ray_gen_us += (unsigned long long)(tm.between(0, 1) * 1000.0f);
        c->stats.time_primary_trace_us += (unsigned long long)(tm.between(1, 2) * 1000.0f);
        c->stats.time_primary_shade_us += (unsigned long long)(tm.between(2, 3) * 1000.0f);
        c->stats.time_secondary_sort_us += (unsigned long long)(t_sort * 1000.0f);
        c->stats.time_secondary_trace_us += (unsigned long long)(t_trace * 1000.0f);
        c->stats.time_secondary_shade_us += (unsigned long long)(t_shade * 1000.0f);
    }
    return 0;
}

int rh_read_pixels(rh_context *c, float *rgba, size_t count) {
    const size_t n = (size_t)c->w * c->h * 4;
    if (count < n) return arg_error("output buffer too small");
    RH_CHECK(hipMemcpyAsync(rgba, c->final_, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_synchronize(rh_context *c) {
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_get_stats(rh_context *c, rh_stats *out) {
    *out = c->stats;
    return 0;
}

int rh_reset_stats(rh_context *c) {
    memset(&c->stats, 0, sizeof(c->stats));
    RH_CHECK(hipMemsetAsync(c->stats_dev, 0, 8 * sizeof(unsigned long long), c->stream));
    return 0;
}

int rh_get_ray_counts(rh_context *c, rh_ray_counts *out) {
    unsigned long long v[8];
    RH_CHECK(hipMemcpyAsync(v, c->stats_dev, sizeof(v), hipMemcpyDeviceToHost, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    out->primary = v[0];
    out->secondary = v[1];
    out->shadow = v[2];
    out->sorted = v[3];
    out->adjoint = v[5];
    out->adjoint_shadow = v[6];
    if (v[4]) {
        g_err = "traversal stack overflow: " + std::to_string(v[4]) + " pushes dropped";
    }
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
    else if (n == "defer_nee") *value = c->defer_nee;
    else if (n == "stack_overflows") {
        unsigned long long v[8];
        RH_CHECK(hipMemcpyAsync(v, c->stats_dev, sizeof(v), hipMemcpyDeviceToHost, c->stream));
        RH_CHECK(hipStreamSynchronize(c->stream));
        *value = (int)v[4];
    } else return arg_error("unknown option");
    return 0;
}

int rh_device_synchronize(rh_context *c) {
    RH_CHECK(hipSetDevice(c->device));
    RH_CHECK(hipDeviceSynchronize());
    return 0;
}

int rh_get_kernel_profile(rh_context *c, rh_kernel_profile *out) {
    memset(out, 0, sizeof(*out));
    RH_CHECK(hipStreamSynchronize(c->stream));
    for (const auto &m : c->ev_marks) {
        float ms = 0;
        RH_CHECK(hipEventElapsedTime(&ms, c->ev_pool[m.second.first], c->ev_pool[m.second.second]));
        out->ms[m.first] += ms;
        out->launches[m.first]++;
    }
    unsigned long long v[4];
    RH_CHECK(hipMemcpy(v, c->visits, sizeof(v), hipMemcpyDeviceToHost));
    out->node_visits = v[0];
    out->counted_rays = v[3];
    out->tri_tests = v[1];
    out->instance_tests = v[2];
    return 0;
}

int rh_reset_kernel_profile(rh_context *c) {
    RH_CHECK(hipStreamSynchronize(c->stream));
    c->ev_marks.clear();
    c->ev_used = 0;
    RH_CHECK(hipMemsetAsync(c->visits, 0, 4 * sizeof(unsigned long long), c->stream));
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
    RH_CHECK(hipMemcpyAsync(c->halton, halton, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float),
                            hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_gen_primary, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, c->stream, pc, c->halton,
                       c->rays[0], c->aux[0], c->counters + 0);
    RH_CHECK(hipGetLastError());
    return download_packets(c, out, (size_t)n, 0);
}

int rh_trace_rays(rh_context *c, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax, rh_hit *out) {
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
    if (!c->have_scene) return arg_error("no scene uploaded");
    if (n <= 0) return 0;
    if ((size_t)n > c->max_rays) return arg_error("n exceeds the context's ray capacity (resize first)");
    TraceArgs a = trace_args(c, d_rays, nullptr, n, d_hits, nullptr, c->stats_dev + 1);
    const int blocks = (n + kBlock - 1) / kBlock;
    hipStream_t s = stream ? (hipStream_t)stream : c->stream;
    if (any_hit) {
        hipLaunchKernelGGL((k_trace<true, false>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
    } else {
        hipLaunchKernelGGL((k_trace<false, false>), dim3(blocks), dim3(kBlock), lds_bytes(c), s, a);
    }
    RH_CHECK(hipGetLastError());
    return 0;
}

int rh_ray_hashes(rh_context *c, const rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3],
                  uint32_t *out) {
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
    int sorted = 0;
    rc = radix_sort(c, nullptr, (int)n, 32, &sorted); // the full key, as SortRays_CPU
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream, c->rays[0],
                       c->aux[0], c->vals[sorted], nullptr, (int)n, c->rays[1], c->aux[1]);
    RH_CHECK(hipGetLastError());
    return download_packets(c, rays, n, 1);
}

int rh_shade_rays(rh_context *c, const rh_pass_info *pi, const rh_ray_packet *rays, const rh_hit *hits, size_t n,
                  const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_count, int *out_src) {
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
    RH_CHECK(hipMemcpyAsync(c->halton, halton, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float),
                            hipMemcpyHostToDevice, c->stream));
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

int rh_adjoint_pass(rh_context *c, const rh_pass_desc *p, const float *d_adj_rgba, float *d_grad, float scale) {
    if (!c->have_scene) return arg_error("no scene uploaded");
    const int rx = p->rect[0], ry = p->rect[1], rw = p->rect[2], rhh = p->rect[3];
    if (rw <= 0 || rhh <= 0 || rx < 0 || ry < 0 || rx + rw > c->w || ry + rhh > c->h)
        return arg_error("region outside the framebuffer");
    if (p->iteration < 1) return arg_error("iteration must be >= 1");
    if (!d_adj_rgba || !d_grad) return arg_error("adjoint image and gradient buffers are required");
    const int depth = std::min<int>(p->cam.pass_settings.max_total_depth, RH_MAX_BOUNCES);
    const size_t lanes = (size_t)rw * rhh;
    const size_t max_vertices = (size_t)depth + 1;
    if (lanes > c->adj_rec_cap || !c->adj_rec) {
        free_ptr(c->adj_rec);
        c->adj_rec = nullptr;
        c->adj_rec_cap = lanes;
        RH_CHECK(hipMalloc(&c->adj_rec, (RH_MAX_BOUNCES + 1) * lanes * sizeof(float4)));
    }
    (void)max_vertices;
    int rc = ensure_spill(c, lanes);
    if (rc) return rc;
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
    a.rec = c->adj_rec;
    a.lanes = (uint32_t)c->adj_rec_cap;
    a.spill = c->spill;
    a.spill_stride = (uint32_t)c->spill_cap;
    a.lds_depth = (uint32_t)c->lds_depth;
    a.ray_counter = c->stats_dev + 5;
    a.shadow_counter = c->stats_dev + 6;
    a.overflow = c->stats_dev + 4;
    KernelMark km(c, 4);
    hipLaunchKernelGGL(k_adjoint, dim3((unsigned)((lanes + kBlock - 1) / kBlock)), dim3(kBlock), lds_bytes(c),
                       c->stream, a);
    RH_CHECK(hipGetLastError());
    return 0;
}

int rh_set_material_colors(rh_context *c, uint32_t first, uint32_t n, const float *rgb) {
    if (!c->have_scene) return arg_error("no scene uploaded");
    if ((size_t)first + n > c->n_materials) return arg_error("material range outside the scene");
    if (!n) return 0;
    const size_t floats = (size_t)3 * n;
    if (floats > c->color_stage_cap || !c->color_stage) {
        free_ptr(c->color_stage);
        c->color_stage = nullptr;
        c->color_stage_cap = floats;
        RH_CHECK(hipMalloc(&c->color_stage, floats * sizeof(float)));
    }
    RH_CHECK(hipMemcpyAsync(c->color_stage, rgb, floats * sizeof(float), hipMemcpyDefault, c->stream));
    hipLaunchKernelGGL(k_set_material_colors, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       const_cast<rh_material *>(c->sc.materials), c->color_stage, first, n);
    RH_CHECK(hipGetLastError());
    // the caller's array may change right after this call returns
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

int rh_read_accum(rh_context *c, float *dst, size_t count) {
    const size_t n = (size_t)c->w * c->h * 4;
    if (count < n) return arg_error("output buffer too small");
    RH_CHECK(hipMemcpyAsync(dst, c->clean, n * sizeof(float), hipMemcpyDefault, c->stream));
    RH_CHECK(hipStreamSynchronize(c->stream));
    return 0;
}

} // extern "C"
