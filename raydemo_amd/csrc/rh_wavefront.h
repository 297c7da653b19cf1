// Wavefront building blocks shared by the single-scene pipeline
// (rh_kernels.hip) and the many-scene batch pipeline (rh_batch.hip): launch
// constants, the ray-hash tables, ray records, block/wave-aggregated queue
// appends, per-pass constants, primary-ray generation, and the body of every
// stage kernel.  A stage body processes one 256-lane tile [base, base + 256)
// of its queue with scene-local arguments; the single-scene kernels call it
// with base = blockIdx.x * 256, the batch kernels for each (scene, tile) work
// item they take, with every pointer already offset to the scene's segment.
#pragma once

#include <hip/hip_runtime.h>

#include "rh_hip.h"
#include "rh_shade.h"

using namespace rh;

namespace {

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

template <bool AnyHit, bool Count, int Tree = kTreeAny>
__device__ __forceinline__ void trace_body(const TraceArgs &a, uint32_t base, uint32_t *lds_stack) {
    const int n = a.count ? *a.count : a.n;
    const uint32_t i = base + threadIdx.x;
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
    st.init(lds_stack, a.spill + i, a.spill_stride, a.lds_depth, kSpillDepth);
    TraceCounters tc = {0, 0, 0};
    if (a.sc.macro_root != RH_INVALID)
        traverse<AnyHit, Tree>(a.sc, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, hit, st, Count ? &tc : nullptr);
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

// trace_body for 8-wide trees with G lanes per ray (traverse_wide's group
// form): thread t walks ray t / G together with the other lanes of its
// group, which hold identical copies of the ray's state; the group's first
// lane writes the hit and the counts.  The block's LDS holds one stack per
// ray (blockDim.x / G of them).
template <bool AnyHit, bool Count, int G>
__device__ __forceinline__ void trace_group_body(const TraceArgs &a, uint32_t *lds_stack) {
    const int n = a.count ? *a.count : a.n;
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x, i = t / G;
    if (t == 0 && a.ray_counter) atomicAdd(a.ray_counter, (unsigned long long)n);
    if ((int)i >= n) return; // the whole group

    const float4 *p = reinterpret_cast<const float4 *>(a.rays + i);
    const float4 r0 = p[0], r1 = p[1];
    Hit hit;
    hit.mask = 0;
    hit.obj = -1;
    hit.prim = -1;
    hit.t = a.tmax ? a.tmax[i] : RH_MAX_DIST;
    hit.u = hit.v = 0.0f;

    Stack st;
    st.init_ray(lds_stack, a.spill + i, a.spill_stride, a.lds_depth, kSpillDepth, threadIdx.x / G, blockDim.x / G);
    TraceCounters tc = {0, 0, 0};
    if (a.sc.macro_root != RH_INVALID)
        traverse_wide<AnyHit, G>(a.sc, f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, hit, st, Count ? &tc : nullptr);
    if (threadIdx.x & (G - 1)) return;
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

// Closest-hit trace with lane refill: a persistent grid walks the queue and a
// lane whose ray is finished takes the next one while the rest of its wave is
// still traversing, so the wave does not idle until its slowest ray ends.
// The traversal of every ray is traverse_fast's, cut into single steps (one
// interior-node pair, one instance, one leaf or one pop sequence per loop
// iteration), so each ray visits the same nodes in the same order and gets
// the same hit.  Lanes fetch new rays in wave batches (one atomic per batch)
// once a quarter of the wave is idle, or when every lane is.
enum : uint32_t { kStepMacro = 0, kStepInstances = 1, kStepMicro = 2, kStepIdle = 3 };

template <bool Count>
__device__ __forceinline__ void trace_refill_body(const TraceArgs &a, int *next_ray, uint32_t *lds_stack) {
    const int n = a.count ? *a.count : a.n;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.ray_counter) atomicAdd(a.ray_counter, (unsigned long long)n);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x; // persistent lane: owns spill slice gl
    Stack st;
    st.init(lds_stack, a.spill + gl, a.spill_stride, a.lds_depth, kSpillDepth);
    TraceCounters tc = {0, 0, 0};
    TraceCounters *cnt = Count ? &tc : nullptr;

    int r = -1;                  // ray of this lane
    uint32_t step = kStepIdle;
    f3 o = f3{0, 0, 0}, d = o, inv_d = o;
    float xy = 0.0f;
    Hit hit;
    hit.mask = 0;
    hit.obj = hit.prim = -1;
    hit.t = 0.0f;
    hit.u = hit.v = 0.0f;
    NodeBox node;                // entered node (macro or micro)
    bool have = false;           // micro: node is entered (else pop next)
    uint32_t leaf_cur = 0, leaf_end = 0, base = 0;
    f3 mo = o, md = o, minv = o;
    int mi_index = 0;
    bool finite = true;
    bool exhausted = false;      // the queue has no rays left for this wave

    for (;;) {
        // ---- refill
        const uint64_t idle = __ballot(step == kStepIdle);
        const uint64_t live = __ballot(true);
        if (!exhausted && idle && (__popcll(idle) >= 16 || idle == live)) {
            int first = 0;
            const int want = __popcll(idle);
            if (lane_id() == (uint32_t)(__ffsll((unsigned long long)idle) - 1)) first = atomicAdd(next_ray, want);
            first = __shfl(first, __ffsll((unsigned long long)idle) - 1);
            if (first + want >= n) exhausted = true;
            if (step == kStepIdle) {
                r = first + lanes_below(idle);
                if (r < n) {
                    const float4 *p = reinterpret_cast<const float4 *>(a.rays + r);
                    const float4 r0 = p[0], r1 = p[1];
                    o = f3{r0.x, r0.y, r0.z};
                    d = f3{r1.x, r1.y, r1.z};
                    xy = r0.w;
                    inv_d = safe_invert(d);
                    hit.mask = 0;
                    hit.obj = -1;
                    hit.prim = -1;
                    hit.t = a.tmax ? a.tmax[r] : RH_MAX_DIST;
                    hit.u = hit.v = 0.0f;
                    st.sp = 0;
                    step = kStepMacro;
                    if (a.sc.macro_root == RH_INVALID) {
                        step = kStepIdle; // empty scene: the miss is written below
                    } else {
                        node = load_node(a.sc.nodes, a.sc.macro_root);
                        if (cnt) cnt->nodes++;
                        if (!box_hit<true>(o, inv_d, hit.t, node.bmin, node.bmax)) step = kStepIdle;
                    }
                    if (step == kStepIdle) {
                        float4 *h = reinterpret_cast<float4 *>(a.hits + r);
                        h[0] = make_float4(__int_as_float(hit.mask), __int_as_float(hit.obj), __int_as_float(hit.prim),
                                           hit.t);
                        h[1] = make_float4(hit.u, hit.v, xy, 0.0f);
                        if (Count) atomicAdd(a.visits + 3, 1ull);
                    }
                }
            }
        }
        if (!__ballot(step != kStepIdle)) {
            if (exhausted) break;
            continue;
        }
        // ---- one traversal step
        bool finished = false;
        if (step == kStepMacro) {
            if (!(node.a & RH_LEAF_NODE_BIT)) {
                if (!enter_children<true>(a.sc, node, o, d, inv_d, hit.t, st, node, cnt) &&
                    !pop_entered<true>(a.sc, 0, o, inv_d, hit.t, st, node, cnt))
                    finished = true;
            } else {
                leaf_cur = node.a & RH_PRIM_INDEX_BITS;
                leaf_end = leaf_cur + node.b;
                step = kStepInstances;
            }
        } else if (step == kStepInstances) {
            if (leaf_cur < leaf_end) {
                const InstanceRec mi = load_instance(a.sc, leaf_cur++);
                mi_index = (int)mi.mi_index;
                if (cnt) cnt->instances++;
                if (box_hit<true>(o, inv_d, hit.t, mi.bmin, mi.bmax)) {
                    transform_ray(mi, o, d, mo, md);
                    minv = f3{1.0f / md.x, 1.0f / md.y, 1.0f / md.z};
                    finite = __builtin_isfinite(minv.x) && __builtin_isfinite(minv.y) && __builtin_isfinite(minv.z);
                    base = st.sp;
                    node = load_node(a.sc.nodes, mi.root);
                    if (cnt) cnt->nodes++;
                    have = finite ? box_hit<true>(mo, minv, hit.t, node.bmin, node.bmax)
                                  : box_hit<false>(mo, minv, hit.t, node.bmin, node.bmax);
                    step = kStepMicro;
                }
            } else if (pop_entered<true>(a.sc, 0, o, inv_d, hit.t, st, node, cnt)) {
                step = kStepMacro;
            } else {
                finished = true;
            }
        } else if (step == kStepMicro) {
            if (have) {
                if (!(node.a & RH_LEAF_NODE_BIT)) {
                    have = finite ? enter_children<true>(a.sc, node, mo, md, minv, hit.t, st, node, cnt)
                                  : enter_children<false>(a.sc, node, mo, md, minv, hit.t, st, node, cnt);
                } else {
                    intersect_leaf<false>(a.sc, node, mo, md, mi_index, hit, cnt);
                    have = false;
                }
            } else {
                have = finite ? pop_entered<true>(a.sc, base, mo, minv, hit.t, st, node, cnt)
                              : pop_entered<false>(a.sc, base, mo, minv, hit.t, st, node, cnt);
                if (!have) step = kStepInstances;
            }
        }
        if (finished) {
            float4 *h = reinterpret_cast<float4 *>(a.hits + r);
            h[0] = make_float4(__int_as_float(hit.mask), __int_as_float(hit.obj), __int_as_float(hit.prim), hit.t);
            h[1] = make_float4(hit.u, hit.v, xy, 0.0f);
            if (Count) atomicAdd(a.visits + 3, 1ull);
            step = kStepIdle;
        }
    }
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    if (Count) {
        atomicAdd(a.visits + 0, (unsigned long long)tc.nodes);
        atomicAdd(a.visits + 1, (unsigned long long)tc.leaf_tris);
        atomicAdd(a.visits + 2, (unsigned long long)tc.instances);
    }
}

// trace_refill_body for 8-wide trees: traverse_wide's closest-hit walk cut
// into single steps (one wide node, one instance, one leaf or one pop per loop
// iteration), so each ray makes the same wide_step / wide_pop / intersect_leaf
// calls in the same order as in traverse_wide and gets the same hit.  The
// node step is shared by the two levels: (co, cd, cinv) is the ray of the
// level the lane is in (the macro ray, or the instance's transformed one), so
// lanes at different levels run one wide_step together.
template <bool Count>
__device__ __forceinline__ void trace_refill_wide_body(const TraceArgs &a, int *next_ray, uint32_t *lds_stack) {
    const int n = a.count ? *a.count : a.n;
    if (blockIdx.x == 0 && threadIdx.x == 0 && a.ray_counter) atomicAdd(a.ray_counter, (unsigned long long)n);
    const uint32_t gl = blockIdx.x * blockDim.x + threadIdx.x; // persistent lane: owns spill slice gl
    Stack st;
    st.init(lds_stack, a.spill + gl, a.spill_stride, a.lds_depth, kSpillDepth);
    TraceCounters tc = {0, 0, 0};
    TraceCounters *cnt = Count ? &tc : nullptr;
    const float4 *wn = a.sc.wnodes;

    int r = -1;                   // ray of this lane
    uint32_t step = kStepIdle;    // kStepMacro / kStepMicro: a node level; kStepInstances: a macro leaf's instances
    f3 o = f3{0, 0, 0}, d = o, inv_d = o;  // the ray
    f3 co = o, cd = o, cinv = o;           // the ray at the current level
    float xy = 0.0f;
    Hit hit;
    hit.mask = 0;
    hit.obj = hit.prim = -1;
    hit.t = 0.0f;
    hit.u = hit.v = 0.0f;
    WideCur cur = {0, 0};         // node or leaf to process when `have`
    bool have = false;            // else the next step pops
    uint32_t leaf_cur = 0, leaf_end = 0, base = 0;
    int mi_index = 0;
    bool finite = true;
    bool exhausted = false;       // the queue has no rays left for this wave

    for (;;) {
        // ---- refill
        const uint64_t idle = __ballot(step == kStepIdle);
        const uint64_t live = __ballot(true);
        if (!exhausted && idle && (__popcll(idle) >= 16 || idle == live)) {
            int first = 0;
            const int want = __popcll(idle);
            if (lane_id() == (uint32_t)(__ffsll((unsigned long long)idle) - 1)) first = atomicAdd(next_ray, want);
            first = __shfl(first, __ffsll((unsigned long long)idle) - 1);
            if (first + want >= n) exhausted = true;
            if (step == kStepIdle) {
                r = first + lanes_below(idle);
                if (r < n) {
                    const float4 *p = reinterpret_cast<const float4 *>(a.rays + r);
                    const float4 r0 = p[0], r1 = p[1];
                    o = f3{r0.x, r0.y, r0.z};
                    d = f3{r1.x, r1.y, r1.z};
                    xy = r0.w;
                    inv_d = safe_invert(d);
                    co = o;
                    cd = d;
                    cinv = inv_d;
                    finite = true;
                    base = 0;
                    hit.mask = 0;
                    hit.obj = -1;
                    hit.prim = -1;
                    hit.t = a.tmax ? a.tmax[r] : RH_MAX_DIST;
                    hit.u = hit.v = 0.0f;
                    st.sp = 0;
                    if (a.sc.macro_root == RH_INVALID) {
                        // empty scene: the miss is written now
                        float4 *h = reinterpret_cast<float4 *>(a.hits + r);
                        h[0] = make_float4(__int_as_float(hit.mask), __int_as_float(hit.obj), __int_as_float(hit.prim),
                                           hit.t);
                        h[1] = make_float4(hit.u, hit.v, xy, 0.0f);
                        if (Count) atomicAdd(a.visits + 3, 1ull);
                    } else {
                        cur = WideCur{a.sc.macro_word, a.sc.macro_count};
                        have = true;
                        step = kStepMacro;
                    }
                }
            }
        }
        if (!__ballot(step != kStepIdle)) {
            if (exhausted) break;
            continue;
        }
        // ---- one traversal step, in three phases so that a lane can enter an
        // instance, or pop, and process the node it arrives at in the same
        // iteration (the wave runs each phase once for all its lanes)
        bool finished = false;
        // (1) the next instance of a macro leaf
        if (step == kStepInstances) {
            if (leaf_cur < leaf_end) {
                const InstanceRec mi = load_instance(a.sc, leaf_cur++);
                mi_index = (int)mi.mi_index;
                if (cnt) cnt->instances++;
                if (box_hit<true>(o, inv_d, hit.t, mi.bmin, mi.bmax)) {
                    transform_ray(mi, o, d, co, cd);
                    cinv = f3{1.0f / cd.x, 1.0f / cd.y, 1.0f / cd.z};
                    finite = __builtin_isfinite(cinv.x) && __builtin_isfinite(cinv.y) && __builtin_isfinite(cinv.z);
                    base = st.sp;
                    cur = wide_root(wn, mi.root);
                    have = true;
                    step = kStepMicro;
                }
            } else {
                step = kStepMacro; // the leaf is done: pop
                have = false;
            }
        }
        // (2) pop
        if ((step == kStepMacro || step == kStepMicro) && !have) {
            have = wide_pop(wn, base, hit.t, st, cur, a.sc.wide_leaf_counts);
            if (!have) {
                if (step == kStepMicro) {
                    // the instance is done: back to the macro leaf's next one
                    co = o;
                    cd = d;
                    cinv = inv_d;
                    finite = true;
                    base = 0;
                    step = kStepInstances;
                } else {
                    finished = true;
                }
            }
        }
        // (3) the node or leaf the lane is at
        if ((step == kStepMacro || step == kStepMicro) && have) {
            if (!(cur.word & RH_LEAF_NODE_BIT)) {
                if (cnt) cnt->nodes++;
                have = finite ? wide_step<true>(wn, cur.word, co, cinv, hit.t, st, cur)
                              : wide_step<false>(wn, cur.word, co, cinv, hit.t, st, cur);
            } else if (step == kStepMicro) {
                NodeBox leaf;
                leaf.a = cur.word;
                leaf.b = cur.count;
                intersect_leaf<false>(a.sc, leaf, co, cd, mi_index, hit, cnt);
                have = false;
            } else {
                leaf_cur = cur.word & RH_PRIM_INDEX_BITS;
                leaf_end = leaf_cur + cur.count;
                step = kStepInstances;
            }
        }
        if (finished) {
            float4 *h = reinterpret_cast<float4 *>(a.hits + r);
            h[0] = make_float4(__int_as_float(hit.mask), __int_as_float(hit.obj), __int_as_float(hit.prim), hit.t);
            h[1] = make_float4(hit.u, hit.v, xy, 0.0f);
            if (Count) atomicAdd(a.visits + 3, 1ull);
            step = kStepIdle;
        }
    }
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    if (Count) {
        atomicAdd(a.visits + 0, (unsigned long long)tc.nodes);
        atomicAdd(a.visits + 1, (unsigned long long)tc.leaf_tris);
        atomicAdd(a.visits + 2, (unsigned long long)tc.instances);
    }
}

// Per-lane vertex records of the path-replay adjoint (lane-minor: record k of
// lane i at k * lanes + i): rec = (C_k, material whose main_color scales it);
// rec_tex (when texel gradients are taken) = 3 records per vertex: the
// footprint's 4 texel indices, (kx, ky, raw r, raw g), (raw b, sRGB flag);
// (raw b, sRGB flag, texture id bits); nverts = vertices recorded per lane.
struct AdjRecords {
    float4 *rec;
    float4 *rec_tex;
    int *nverts;
    uint32_t lanes;
    int vertex;
    int rect[3]; // x, y, w of the pass rectangle
};

__device__ __forceinline__ void write_vertex_record(const AdjRecords &r, uint32_t lane, int k, f3 cj,
                                                    const ShadeOut &so) {
    r.rec[(size_t)k * r.lanes + lane] = make_float4(cj.x, cj.y, cj.z, __int_as_float(so.alb_mat));
    if (r.rec_tex && so.alb_mat >= 0) {
        float4 *rt = r.rec_tex + (size_t)k * 3 * r.lanes + lane;
        rt[0] = make_float4(__uint_as_float(so.alb_foot.idx[0]), __uint_as_float(so.alb_foot.idx[1]),
                            __uint_as_float(so.alb_foot.idx[2]), __uint_as_float(so.alb_foot.idx[3]));
        rt[r.lanes] = make_float4(so.alb_foot.kx, so.alb_foot.ky, so.alb_raw.x, so.alb_raw.y);
        rt[2 * (size_t)r.lanes] = make_float4(so.alb_raw.z, so.alb_srgb ? 1.0f : 0.0f, __uint_as_float(so.alb_tex),
                                              __int_as_float(so.alb_foot.lod));
    }
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
    unsigned long long *visits;      // path tail, Count mode: node visits, triangle tests, instance tests,
                                     // segments, shadow rays
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
    // Vertex records of a staged adjoint replay (null: write pixels).  A ray
    // of pixel (x, y) is lane (y - rec_rect[1]) * rec_rect[2] + (x - rec_rect[0])
    // of the pass rectangle; this stage's vertex index is rec_vertex.
    AdjRecords rec;
};

// Defer: direct light goes to the NEE stage (no traversal in this kernel).
template <bool Defer>
__device__ __forceinline__ void shade_body(const ShadeArgs &a, uint32_t base, uint32_t *lds_stack, int *s_app) {
    const int n = a.count ? *a.count : a.n;
    const uint32_t i = base + threadIdx.x;
    if ((int)base >= n) return; // whole block idle
    const bool active = (int)i < n;

    ShadeOut so;
    so.emit = false;
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    Stack st;
    st.init(lds_stack, a.spill + i, a.spill_stride, a.lds_depth, kSpillDepth);
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

        if (a.rec.rec) {
            // staged adjoint replay: the vertex's colour goes to its record (the
            // combine stage adds the visible direct light to it)
            const uint32_t lane = (uint32_t)((py - a.rec.rect[1]) * a.rec.rect[2] + (px - a.rec.rect[0]));
            write_vertex_record(a.rec, lane, a.rec.vertex, f3{so.col.x, so.col.y, so.col.z}, so);
            a.rec.nverts[lane] = a.rec.vertex + 1;
            if (Defer && nee_m) {
                // the combine stage finds the record through pend_c.w
                a.pend_c[i].w = __int_as_float((int)lane);
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
// Defer: the direct light is recorded in the lane's NEE slots as in the staged
// deferred form and resolved after ShadeSurface returns (visibility, visible
// terms summed in light order, albedo, throughput: the combine stage's
// arithmetic), so the shading state is dead while the shadow rays traverse
// and fewer registers are live at the peak.
//
// Refill (next_path set): the grid is persistent and a lane whose path ends
// takes the next unstarted path of the queue (a wave fetches for its idle
// lanes with one atomic once a quarter of it is idle, or all of it), so a
// wave no longer idles until its longest path ends.  Each path is still
// followed by one lane in bounce order, so every pixel receives the same
// values in the same order.  Without refill each lane follows the one path
// of its index.  A lane's stack spill slice belongs to its grid position; the
// NEE slots of a path belong to its queue index.
template <bool Defer, int Tree = kTreeAny, bool Count = false, bool Solid = false>
__device__ __forceinline__ void path_tail_body(const ShadeArgs &a, int first, int last, uint32_t base,
                                               uint32_t *lds_stack, int *next_path = nullptr) {
    const int n = *a.count;
    const uint32_t gl = base + threadIdx.x;
    if (first >= last || (!next_path && (int)base >= n)) return; // no bounce left, or the whole block idle
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    uint32_t traced = 0;
    Stack st;
    st.init(lds_stack, a.spill + gl, a.spill_stride, a.lds_depth, kSpillDepth);
    TraceCounters tc = {0, 0, 0};
    TraceCounters *tcp = Count ? &tc : nullptr;

    int path = -1; // queue index of the lane's path (-1: none)
    int b = first; // host bounce of the lane's next segment
    RayFull ray;
    rh_pass_info pi = a.pi;
    size_t pixel = 0;
    auto begin = [&](int p) {
        path = p;
        b = first;
        ray = load_ray(a.rays, a.aux, (uint32_t)p);
        const int px = (ray.xy >> 16) & 0xffff, py = ray.xy & 0xffff;
        pi.index = py * a.w + px;
        pi.rand_index = pi.index;
        if (pi.settings.flags & RH_USE_COHERENT_SAMPLING) pi.rand_index = c_sampling_pattern[(py % 8) * 8 + (px % 8)];
        pixel = (size_t)py * a.w + px;
    };
    bool exhausted = next_path == nullptr; // no more paths to take
    if (!next_path && (int)gl < n) begin((int)gl);

    for (;;) {
        if (!exhausted) {
            const uint64_t idle = __ballot(path < 0), live = __ballot(true);
            if (idle && (__popcll(idle) >= 16 || idle == live)) {
                const int want = __popcll(idle), leader = __ffsll((unsigned long long)idle) - 1;
                int first_path = 0;
                if ((int)lane_id() == leader) first_path = atomicAdd(next_path, want);
                first_path = __shfl(first_path, leader);
                if (first_path + want >= n) exhausted = true;
                if (path < 0) {
                    const int p = first_path + lanes_below(idle);
                    if (p < n) begin(p);
                }
            }
        }
        if (!__ballot(path >= 0)) {
            if (exhausted) break;
            continue;
        }
        if (path < 0) continue;
        Hit hit;
        hit.mask = 0;
        hit.obj = -1;
        hit.prim = -1;
        hit.t = RH_MAX_DIST;
        hit.u = hit.v = 0.0f;
        st.sp = 0;
        if (a.sc.macro_root != RH_INVALID) traverse<false, Tree>(a.sc, ray.o, ray.d, hit, st, tcp);
        traced++;
        pi.bounce = b + 3;
        ShadeOut so;
        so.emit = false;
        NeeSink sink;
        if (Defer) sink = NeeSink{a.nee_ray, a.nee_aux, a.nee_term, (uint32_t)path * a.nee_slots, 0};
        shade_surface<Tree>(a.sc, pi, hit, ray, a.halton, st, cnt, so, Defer ? &sink : nullptr);
        float4 &dst = a.temp[pixel];
        float4 v = dst;
        v.x += so.col.x;
        v.y += so.col.y;
        v.z += so.col.z;
        v.w += 0.0f;
        dst = v;
        if (Defer && so.nee_pending && sink.count > 0) {
            const f3 add = resolve_light_samples<Tree, Solid>(a.sc, a.halton, sink, so, ray.c, st, cnt, tcp);
            float4 w = dst;
            w.x += add.x;
            w.y += add.y;
            w.z += add.z;
            dst = w;
        }
        b++;
        if (!so.emit || b >= last) path = -1;
        else ray = so.next;
    }
    if (a.ray_counter) wave_add_u64(a.ray_counter, traced);
    if (a.shadow_counter) wave_add_u64(a.shadow_counter, cnt.shadow_rays);
    if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    if (Count && a.visits) {
        wave_add_u64(a.visits + 0, tc.nodes);
        wave_add_u64(a.visits + 1, tc.leaf_tris);
        wave_add_u64(a.visits + 2, tc.instances);
        wave_add_u64(a.visits + 3, traced);
        wave_add_u64(a.visits + 4, cnt.shadow_rays);
    }
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
    unsigned long long *visits; // Count mode: node visits, triangle tests, instance tests, entries, segments
};

// Visibility of each deferred light sample: ComputeVisibility
// (CoreRef.cpp:1801-1896), any-hit traversal stepping through transparent
// surfaces.  One lane per queue entry; the grid covers every slot and lanes
// past the device count exit at once.
template <int Tree = kTreeAny, bool Count = false, bool Solid = false>
__device__ __forceinline__ void shadow_body(const ShadowArgs &a, uint32_t base, uint32_t *lds_stack) {
    const int n = *a.count;
    const uint32_t q = base + threadIdx.x;
    if ((int)base >= n) return; // whole block idle
    if (q == 0 && a.shadow_counter) atomicAdd(a.shadow_counter, (unsigned long long)n);
    ShadeCounters cnt;
    cnt.shadow_rays = 0;
    TraceCounters tc = {0, 0, 0};
    if ((int)q < n) {
        const uint32_t k = a.queue[q];
        const float4 r = a.nee_ray[k];
        const int2 x = a.nee_aux[k];
        const rh_light &l = a.sc.lights[__float_as_uint(r.w)];
        Stack st;
        st.init(lds_stack, a.spill + q, a.spill_stride, a.lds_depth, kSpillDepth);
        const float vis =
            compute_visibility<Tree, Solid>(a.sc, f3{r.x, r.y, r.z}, ld3(l.pos), a.halton, x.x, x.y, st, cnt,
                                            Count ? &tc : nullptr);
        a.nee_vis[k] = vis != 0.0f ? 1 : 0;
        if (st.overflow && a.overflow) atomicAdd(a.overflow, (unsigned long long)st.overflow);
    }
    // one segment per entry is counted above; add the difference (segments
    // past the first through transparent surfaces, or none at all when the
    // light is within HIT_EPS)
    const int extra = (int)q < n ? (int)cnt.shadow_rays - 1 : 0;
    if (a.shadow_counter && __ballot(extra != 0)) wave_add_i64(a.shadow_counter, extra);
    if (Count && a.visits) {
        wave_add_u64(a.visits + 0, tc.nodes);
        wave_add_u64(a.visits + 1, tc.leaf_tris);
        wave_add_u64(a.visits + 2, tc.instances);
        wave_add_u64(a.visits + 3, (int)q < n ? 1u : 0u);
        wave_add_u64(a.visits + 4, cnt.shadow_rays);
    }
}

// Adds each shaded ray's deferred direct light: ray.c * (sum of visible terms
// in light order [* albedo]), to its pixel (or to the per-ray output).
struct CombineArgs {
    const int *count; // shaded rays (device); nullptr -> n
    int n;
    uint32_t slots;
    const int *pend_count;
    const float4 *pend_c;
    const float4 *pend_alb;
    const float4 *nee_term;
    const uint8_t *nee_vis;
    float4 *temp;     // pixels (pend_c.w holds the pixel index), or
    float4 *out_rgba; // per-ray output (parity entry point), or
    AdjRecords rec;   // vertex records (pend_c.w holds the lane), and the
                      // pixels too when temp is set (the adjoint tape)
    int frame_w;      // framebuffer width (pixel of a lane, with records and pixels)
};

__device__ __forceinline__ void combine_body(const CombineArgs &a, uint32_t base) {
    const int n = a.count ? *a.count : a.n;
    const int i = (int)(base + threadIdx.x);
    if (i >= n) return;
    const int m = a.pend_count[i];
    if (!m) return;
    f3 col = f3{0.0f, 0.0f, 0.0f};
    const uint32_t s0 = (uint32_t)i * a.slots;
    for (int j = 0; j < m; j++) {
        if (a.nee_vis[s0 + j]) {
            const float4 t = a.nee_term[s0 + j];
            col = col + f3{t.x, t.y, t.z};
        }
    }
    const float4 pc = a.pend_c[i], pa = a.pend_alb[i];
    if (__float_as_int(pa.w) & 1) col = col * f3{pa.x, pa.y, pa.z};
    const f3 add = f3{pc.x * col.x, pc.y * col.y, pc.z * col.z};
    float4 *dst = a.rec.rec ? &a.rec.rec[(size_t)a.rec.vertex * a.rec.lanes + (uint32_t)__float_as_int(pc.w)]
                  : a.temp  ? &a.temp[__float_as_int(pc.w)]
                            : &a.out_rgba[i];
    float4 v = *dst;
    v.x += add.x;
    v.y += add.y;
    v.z += add.z;
    *dst = v;
    if (a.rec.rec && a.temp) {
        // the tape: the pixel too, found from the lane
        const int lane = __float_as_int(pc.w);
        float4 &px = a.temp[(size_t)(a.rec.rect[1] + lane / a.rec.rect[2]) * a.frame_w + a.rec.rect[0] +
                            lane % a.rec.rect[2]];
        float4 u = px;
        u.x += add.x;
        u.y += add.y;
        u.z += add.z;
        px = u;
    }
}

// MixWith + tonemap of one pixel (FramebufferRef.cpp:80-86, RendererRef.cpp:301-324):
// clean += (temp - clean) * k, then final = tonemap(clean)
__device__ __forceinline__ void mix_tonemap_pixel(const float4 *__restrict__ temp, float4 *__restrict__ clean,
                                                  float4 *__restrict__ final_, size_t p, float k, int dtype,
                                                  float gamma, uint32_t flags) {
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

} // namespace
