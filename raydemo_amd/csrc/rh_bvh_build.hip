// The fast BVH build (mesh_desc_t::use_fast_bvh_build) on the device: the
// bottom half of PreprocessPrims_HLBVH (Core.cpp:532-596) and EmitLBVH
// (Core.cpp:342-429) as kernels.  The host joins the treelets with the SAH
// top tree (Core.cpp:598-680) over at most 4096 boxes.
//
// Morton codes and their stable sort give every primitive its place; the
// treelets are the runs sharing the codes' top 12 bits.  The reference emits
// each treelet depth first, appending a split node's two children together,
// so the children of the k-th interior node in depth-first order are the
// treelet's nodes 2k+1 and 2k+2.  Here the trees are built level by level
// instead (all treelets at once, one launch per level, at most 19 levels:
// each split takes one of the bits 17..0), then the subtrees' interior counts
// are summed bottom-up and the depth-first ranks handed down, which gives
// every node the index the reference's order gives it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rh_hip.h"
#include "rh_radix.h"
#include "rh_runtime.h"

using rh_rt::arg_error;
using rh_rt::fail;
using rh_rt::free_ptr;

namespace {

constexpr int kTreeletBits = 12;             // top Morton bits naming a treelet
constexpr int kTreeletShift = 18;            // Morton bits 18..29
constexpr int kTreelets = 1 << kTreeletBits; // at most 4096
constexpr int kFirstBit = kTreeletShift - 1; // treelets split from bit 17 down
constexpr int kLevels = kFirstBit + 2;       // root level + one per bit
constexpr int kLeafBelow = 8;                // ranges of fewer primitives are leaves
constexpr int kPartials = 1024;              // blocks of the bounds reduction

struct Task {
    uint32_t first, count; // range of the code-sorted primitives
    int32_t bit;           // split bit, -1 for a leaf
    uint32_t child;        // task of the lower half (the upper half is child + 1)
    uint32_t treelet;
    uint32_t interior;     // interior nodes in the subtree, itself included
    uint32_t rank;         // depth-first rank among its treelet's interior nodes
    uint32_t local;        // node index within the treelet
    uint32_t swapped;      // children stored upper half first
    float bmin[3], bmax[3];
};

__device__ __forceinline__ float min_std(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_std(float a, float b) { return (a < b) ? b : a; }

// float -> uint32_t as the reference's x86-64 build converts it: a 64-bit
// truncation whose low half is kept (NaN and out-of-range values give 0)
__device__ __forceinline__ uint32_t f2u32_x86(float f) {
    if (!(f > -9.2233715e18f && f < 9.2233715e18f)) return 0u;
    return (uint32_t)(int64_t)f;
}

__device__ __forceinline__ uint32_t spread_bits3(uint32_t x) {
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// Mesh bounds, pass 1: per-block min/max of the primitive boxes (6 floats a
// primitive: bbox_min xyz, bbox_max xyz).
__global__ void __launch_bounds__(kBlock) k_lbvh_bounds(const float *__restrict__ prims, uint32_t n,
                                                       float *__restrict__ partial) {
    __shared__ float s[6][kBlock];
    const float big = 3.402823466e+38f;
    float lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const float *p = prims + 6 * (size_t)i;
        for (int a = 0; a < 3; a++) {
            lo[a] = min_std(lo[a], p[a]);
            hi[a] = max_std(hi[a], p[3 + a]);
        }
    }
    for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = min_std(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = max_std(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

// Mesh bounds, pass 2 (one block): the box and the Morton scale
// 1024 / extent per axis (infinite for a flat axis, as in the reference).
__global__ void __launch_bounds__(kBlock) k_lbvh_scale(const float *__restrict__ partial, int blocks,
                                                      float *__restrict__ box) {
    __shared__ float s[6][kBlock];
    const float big = 3.402823466e+38f;
    float lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
    for (int b = threadIdx.x; b < blocks; b += kBlock)
        for (int a = 0; a < 3; a++) {
            lo[a] = min_std(lo[a], partial[6 * b + a]);
            hi[a] = max_std(hi[a], partial[6 * b + 3 + a]);
        }
    for (int a = 0; a < 3; a++) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; a++) {
                s[a][threadIdx.x] = min_std(s[a][threadIdx.x], s[a][threadIdx.x + w]);
                s[3 + a][threadIdx.x] = max_std(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int a = 0; a < 3; a++) {
            box[a] = s[a][0];
            box[3 + a] = 1024.0f / (s[3 + a][0] - s[a][0]);
        }
}

// Morton code of each primitive's box centre (Core.cpp:552-562), with its
// index as the sort value.
__global__ void __launch_bounds__(kBlock) k_lbvh_codes(const float *__restrict__ prims, uint32_t n,
                                                      const float *__restrict__ box, uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float *p = prims + 6 * (size_t)i;
    uint32_t c[3];
    for (int a = 0; a < 3; a++) {
        const float centre = 0.5f * (p[a] + p[3 + a]);
        c[a] = f2u32_x86((centre - box[a]) * box[3 + a]);
    }
    keys[i] = (spread_bits3(c[2]) << 2) + (spread_bits3(c[1]) << 1) + spread_bits3(c[0]);
    vals[i] = i;
}

// First sorted position of each treelet (head[top 12 bits]).
__global__ void __launch_bounds__(kBlock) k_lbvh_heads(const uint32_t *__restrict__ codes, uint32_t n,
                                                      uint32_t *__restrict__ head) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const uint32_t t = codes[i] >> kTreeletShift;
    if (i == 0 || (codes[i - 1] >> kTreeletShift) != t) head[t] = i;
}

// One block: the non-empty treelets in code order, each a root task of
// level 0.  meta[0] = task count, meta[1] = treelet count, meta[2 + 2 L] and
// meta[3 + 2 L] = level L's task range.
__global__ void __launch_bounds__(kBlock) k_lbvh_treelets(const uint32_t *__restrict__ head, uint32_t n,
                                                         Task *__restrict__ tasks, uint32_t *__restrict__ meta) {
    constexpr int kPer = kTreelets / kBlock; // 16 buckets per thread
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t firsts[kTreelets + 1];
    const int t = threadIdx.x;
    uint32_t cnt = 0;
    for (int k = 0; k < kPer; k++) cnt += head[t * kPer + k] != 0xffffffffu;
    const uint32_t inc = wave_inclusive_scan(cnt);
    if ((t & 63) == 63) wsum[t >> 6] = inc;
    __syncthreads();
    uint32_t before = inc - cnt;
    for (int w = 0; w < (t >> 6); w++) before += wsum[w];
    uint32_t total = 0;
    for (int w = 0; w < kBlock / 64; w++) total += wsum[w];
    for (int k = 0; k < kPer; k++) {
        const uint32_t h = head[t * kPer + k];
        if (h != 0xffffffffu) firsts[before++] = h;
    }
    __syncthreads();
    if (t == 0) firsts[total] = n;
    __syncthreads();
    for (uint32_t j = t; j < total; j += kBlock) {
        Task r;
        memset(&r, 0, sizeof(r));
        r.first = firsts[j];
        r.count = firsts[j + 1] - firsts[j];
        r.bit = kFirstBit;
        r.treelet = j;
        tasks[j] = r;
    }
    if (t == 0) {
        meta[0] = total;
        meta[1] = total;
        meta[2] = 0;
        meta[3] = total;
    }
}

// One level: each task becomes a leaf or splits where its highest differing
// bit turns from 0 to 1 (EmitLBVH's test and binary search); a split appends
// its two halves to the next level.
__global__ void __launch_bounds__(kBlock) k_lbvh_split(Task *__restrict__ tasks, const uint32_t *__restrict__ codes,
                                                      uint32_t *__restrict__ meta, int level) {
    const uint32_t lb = meta[2 + 2 * level], le = meta[3 + 2 * level];
    const uint32_t j = lb + blockIdx.x * kBlock + threadIdx.x;
    if (j >= le) return;
    Task &k = tasks[j];
    const uint32_t first = k.first, count = k.count, last = first + count - 1;
    int bit = k.bit;
    while (bit >= 0 && count >= (uint32_t)kLeafBelow && ((codes[first] >> bit) & 1u) == ((codes[last] >> bit) & 1u))
        bit--;
    if (bit < 0 || count < (uint32_t)kLeafBelow) {
        k.bit = -1;
        return;
    }
    // the range agrees above `bit` and is sorted: 0...0 1...1 in it
    uint32_t lo = first, hi = last; // bit(lo) = 0, bit(hi) = 1
    while (lo + 1 != hi) {
        const uint32_t mid = (lo + hi) / 2;
        if (((codes[mid] >> bit) & 1u) == 0u) lo = mid;
        else hi = mid;
    }
    const uint32_t c = atomicAdd(&meta[0], 2u);
    Task a, b;
    memset(&a, 0, sizeof(a));
    memset(&b, 0, sizeof(b));
    a.first = first;
    a.count = hi - first;
    b.first = hi;
    b.count = last + 1 - hi;
    a.bit = b.bit = bit - 1;
    a.treelet = b.treelet = k.treelet;
    tasks[c] = a;
    tasks[c + 1] = b;
    k.bit = bit;
    k.child = c;
}

__global__ void k_lbvh_close_level(uint32_t *__restrict__ meta, int level) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        meta[2 + 2 * (level + 1)] = meta[3 + 2 * level];
        meta[3 + 2 * (level + 1)] = meta[0];
    }
}

// Bottom-up over one level: leaf boxes over their primitives in order
// (EmitLBVH's loop), interior boxes from the two halves (lower first, as the
// reference combines left and right before its swap), the swap test on the
// split axis, and the subtree's interior count.
__global__ void __launch_bounds__(kBlock) k_lbvh_up(Task *__restrict__ tasks, const float *__restrict__ prims,
                                                   const uint32_t *__restrict__ order,
                                                   const uint32_t *__restrict__ meta, int level) {
    const uint32_t lb = meta[2 + 2 * level], le = meta[3 + 2 * level];
    const uint32_t j = lb + blockIdx.x * kBlock + threadIdx.x;
    if (j >= le) return;
    Task &k = tasks[j];
    if (k.bit < 0) {
        const float big = 3.402823466e+38f;
        float lo[3] = {big, big, big}, hi[3] = {-big, -big, -big};
        for (uint32_t i = k.first; i < k.first + k.count; i++) {
            const float *p = prims + 6 * (size_t)order[i];
            for (int a = 0; a < 3; a++) {
                lo[a] = min_std(lo[a], p[a]);
                hi[a] = max_std(hi[a], p[3 + a]);
            }
        }
        for (int a = 0; a < 3; a++) {
            k.bmin[a] = lo[a];
            k.bmax[a] = hi[a];
        }
        k.interior = 0;
        return;
    }
    const Task &l = tasks[k.child], &r = tasks[k.child + 1];
    for (int a = 0; a < 3; a++) {
        k.bmin[a] = min_std(l.bmin[a], r.bmin[a]);
        k.bmax[a] = max_std(l.bmax[a], r.bmax[a]);
    }
    const int axis = k.bit % 3;
    k.swapped = l.bmin[axis] > r.bmin[axis];
    k.interior = 1 + l.interior + r.interior;
}

// One block: each treelet's first node in the treelets' node array (a
// treelet of I interior nodes has 2 I + 1 nodes); meta[4 + 2 kLevels] = total.
__global__ void __launch_bounds__(kBlock) k_lbvh_bases(const Task *__restrict__ tasks, uint32_t *__restrict__ meta,
                                                      uint32_t *__restrict__ base) {
    __shared__ uint32_t wsum[kBlock / 64];
    __shared__ uint32_t carry;
    const uint32_t T = meta[1];
    const int t = threadIdx.x;
    if (t == 0) carry = 0;
    __syncthreads();
    for (uint32_t b0 = 0; b0 < T; b0 += kBlock) {
        const uint32_t i = b0 + t;
        const uint32_t v = i < T ? 2 * tasks[i].interior + 1 : 0u;
        const uint32_t inc = wave_inclusive_scan(v);
        if ((t & 63) == 63) wsum[t >> 6] = inc;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < (t >> 6); w++) before += wsum[w];
        if (i < T) base[i] = before + inc - v;
        __syncthreads();
        if (t == kBlock - 1) carry = before + inc;
        __syncthreads();
    }
    if (t == 0) meta[4 + 2 * kLevels] = carry;
}

// Top-down over one level: the depth-first ranks.  The children of the
// interior node of rank r are the treelet's nodes 2 r + 1 (lower half) and
// 2 r + 2; the lower half's rank follows r, the upper half's follows the
// lower half's subtree.
__global__ void __launch_bounds__(kBlock) k_lbvh_down(Task *__restrict__ tasks, const uint32_t *__restrict__ meta,
                                                     int level) {
    const uint32_t lb = meta[2 + 2 * level], le = meta[3 + 2 * level];
    const uint32_t j = lb + blockIdx.x * kBlock + threadIdx.x;
    if (j >= le) return;
    Task &k = tasks[j];
    if (level == 0) {
        k.rank = 0;
        k.local = 0;
    }
    if (k.bit < 0) return;
    Task &l = tasks[k.child], &r = tasks[k.child + 1];
    l.local = 2 * k.rank + 1;
    r.local = 2 * k.rank + 2;
    l.rank = k.rank + 1;
    r.rank = k.rank + 1 + l.interior;
}

// Node records in the reference's layout (interior child links relative to
// the treelets' node array; leaves point into the scene's index array).
__global__ void __launch_bounds__(kBlock) k_lbvh_write(const Task *__restrict__ tasks,
                                                      const uint32_t *__restrict__ meta,
                                                      const uint32_t *__restrict__ base, uint32_t index_offset,
                                                      rh_bvh_node *__restrict__ nodes) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= meta[0]) return;
    const Task &k = tasks[j];
    const uint32_t tb = base[k.treelet];
    rh_bvh_node n;
    for (int a = 0; a < 3; a++) {
        n.bbox_min[a] = k.bmin[a];
        n.bbox_max[a] = k.bmax[a];
    }
    if (k.bit < 0) {
        n.prim_index = RH_LEAF_NODE_BIT + k.first + index_offset;
        n.prim_count = k.count;
    } else {
        uint32_t lo = tb + tasks[k.child].local, hi = tb + tasks[k.child + 1].local;
        if (k.swapped) {
            const uint32_t x = lo;
            lo = hi;
            hi = x;
        }
        n.prim_index = lo;
        n.prim_count = hi + ((uint32_t)(k.bit % 3) << 30);
    }
    nodes[tb + k.local] = n;
}

// Treelet records for the host: (first, count, root node), root boxes follow
// in the nodes.
__global__ void __launch_bounds__(kBlock) k_lbvh_treelet_out(const Task *__restrict__ tasks,
                                                            const uint32_t *__restrict__ meta,
                                                            const uint32_t *__restrict__ base,
                                                            uint32_t *__restrict__ out) {
    const uint32_t j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= meta[1]) return;
    out[3 * j + 0] = tasks[j].first;
    out[3 * j + 1] = tasks[j].count;
    out[3 * j + 2] = base[j];
}

// ------------------------------------------------------------------ SAH build
// PreprocessPrims_SAH (Core.cpp:431-530) with SplitPrimitives_SAH's object
// split (BVHSplit.cpp:126-218, no spatial splits) over n boxes, on the device,
// node for node equal to the host's.  The host takes its lists from a deque in
// creation order, so the nodes come out breadth first and a split node's
// children take the next two slots: the k-th interior node in that order has
// its children at 2k+1 and 2k+2, and the leaves' primitives follow one
// another in node order.  Here each level of the tree is one launch with one
// workgroup per node: per axis it sorts the node's primitives by bbox_max,
// scans their boxes from both ends and evaluates every split position; a
// second launch numbers the level's children and leaves in node order.
//
// The host orders each axis with std::sort, whose order of equal keys this
// sort cannot reproduce: a node whose keys tie along any axis raises a flag
// and the caller builds on the host instead.
constexpr int kSahThreads = 1024;
constexpr uint32_t kSahLdsKeys = 4096; // nodes of more primitives sort in global scratch

struct SahTask {
    uint32_t first, count; // range of the level's primitive order
    uint32_t node;         // node index
    float bmin[3], bmax[3]; // the node's box as the parent's split computed it
};

struct SahSplit {
    int32_t axis;   // -1: a leaf
    uint32_t index; // primitives going left
    float lmin[3], lmax[3], rmin[3], rmax[3];
};

struct Box6 {
    float v[6]; // min xyz, max xyz
};

__device__ __forceinline__ Box6 box_identity() {
    Box6 b;
    for (int k = 0; k < 3; k++) {
        b.v[k] = 3.402823466e+38F;
        b.v[3 + k] = -3.402823466e+38F;
    }
    return b;
}
// the running box `acc` grown by `x`, with std::min/std::max's choice between
// equal values (the running value stays): associative, so scans may group
__device__ __forceinline__ Box6 box_join(const Box6 &acc, const Box6 &x) {
    Box6 r;
    for (int k = 0; k < 3; k++) {
        r.v[k] = min_std(acc.v[k], x.v[k]);
        r.v[3 + k] = max_std(acc.v[3 + k], x.v[3 + k]);
    }
    return r;
}
__device__ __forceinline__ Box6 load_box(const float *boxes, uint32_t p) {
    Box6 b;
    for (int k = 0; k < 6; k++) b.v[k] = boxes[(size_t)p * 6 + k];
    return b;
}
// bbox_t::surface_area (BVHSplit.cpp:8-23): 2 (dx + dy + dz)
__device__ __forceinline__ float half_area(const float *mn, const float *mx) {
    const float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
    return 2 * (d0 + d1 + d2);
}

// ascending sort of m (key, value) pairs at key/val (room for the next power
// of two), equal keys ordered by value; every thread of the block calls it
__device__ void block_bitonic(float *key, uint32_t *val, uint32_t m) {
    uint32_t np = 1;
    while (np < m) np <<= 1;
    for (uint32_t i = m + threadIdx.x; i < np; i += blockDim.x) {
        key[i] = __builtin_inff();
        val[i] = 0xffffffffu;
    }
    __syncthreads();
    for (uint32_t k = 2; k <= np; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const float ki = key[i], kl = key[l];
                    const uint32_t vi = val[i], vl = val[l];
                    const bool gt = ki > kl || (ki == kl && vi > vl);
                    const bool up = (i & k) == 0;
                    if (gt == up) {
                        key[i] = kl;
                        key[l] = ki;
                        val[i] = vl;
                        val[l] = vi;
                    }
                }
            }
            __syncthreads();
        }
    }
}

// Hillis-Steele scan of one Box6 per thread in LDS (6 floats each): forward,
// each entry becomes the join of itself and all before it; backward, of
// itself and all after it (the later entries' box being the running one)
__device__ void block_scan_boxes(float *s, bool backward) {
    const uint32_t t = threadIdx.x, T = blockDim.x;
    for (uint32_t d = 1; d < T; d <<= 1) {
        Box6 other, mine;
        bool have = false;
        const uint32_t o = backward ? t + d : t - d;
        if (backward ? (t + d < T) : (t >= d)) {
            for (int k = 0; k < 6; k++) other.v[k] = s[6 * o + k];
            have = true;
        }
        for (int k = 0; k < 6; k++) mine.v[k] = s[6 * t + k];
        __syncthreads();
        if (have) {
            // forward the earlier entries' box runs, backward the later ones'
            const Box6 r = box_join(other, mine);
            for (int k = 0; k < 6; k++) s[6 * t + k] = r.v[k];
        }
        __syncthreads();
    }
}

// One workgroup per task of the level: the best object split of its range of
// order_in (sorted orders, left/right bounds per axis in scratch), the chosen
// axis's order (or, for a leaf, the incoming one) written to order_out.
// flags[0] is set where keys tie.
__global__ void __launch_bounds__(kSahThreads) k_sah_split(const float *__restrict__ boxes,
                                                          const SahTask *__restrict__ tasks,
                                                          const uint32_t *__restrict__ ntasks,
                                                          const uint32_t *__restrict__ order_in,
                                                          uint32_t *__restrict__ order_out, float *gkey,
                                                          uint32_t *gval, uint32_t *sorted, float *lb, float *rb,
                                                          size_t n, float oversplit, float cost,
                                                          SahSplit *__restrict__ splits, uint32_t *flags) {
    if (blockIdx.x >= *ntasks) return;
    __shared__ float s_key[kSahLdsKeys];
    __shared__ uint32_t s_val[kSahLdsKeys];
    __shared__ float s_box[6 * kSahThreads];
    __shared__ float s_sah[kSahThreads];
    __shared__ uint32_t s_idx[kSahThreads];
    __shared__ int s_tie;

    const SahTask task = tasks[blockIdx.x];
    const uint32_t first = task.first, m = task.count, t = threadIdx.x, T = blockDim.x;
    const float sa_whole = half_area(task.bmin, task.bmax);
    // a leaf unless a split beats this (the host's initial res_sah)
    float best = oversplit * sa_whole * (float)m;
    int best_axis = -1;
    uint32_t best_index = 0;

    const bool in_lds = m <= kSahLdsKeys;
    float *key = in_lds ? s_key : gkey + 2 * (size_t)first;
    uint32_t *val = in_lds ? s_val : gval + 2 * (size_t)first;
    const uint32_t K = (m + T - 1) / T; // consecutive positions per thread
    const uint32_t c0 = min(m, t * K), c1 = min(m, c0 + K);

    for (int axis = 0; axis < 3 && m > 1; axis++) {
        if (t == 0) s_tie = 0;
        for (uint32_t i = t; i < m; i += T) {
            const uint32_t p = order_in[first + i];
            key[i] = boxes[(size_t)p * 6 + 3 + axis];
            val[i] = p;
        }
        __syncthreads();
        block_bitonic(key, val, m);
        uint32_t *srt = sorted + (size_t)axis * n + first;
        for (uint32_t i = t; i < m; i += T) {
            srt[i] = val[i];
            if (i + 1 < m && key[i] == key[i + 1]) s_tie = 1;
        }
        __syncthreads();
        if (s_tie) {
            // the host builds this tree; the node is written as a leaf so that
            // the level's bookkeeping stays within its arrays
            if (t == 0) atomicOr(flags, 1u);
            best_axis = -1;
            break;
        }
        // forward: left_i = join of sorted[0 .. i-1]
        Box6 acc = box_identity();
        for (uint32_t i = c0; i < c1; i++) acc = box_join(acc, load_box(boxes, srt[i]));
        for (int k = 0; k < 6; k++) s_box[6 * t + k] = acc.v[k];
        __syncthreads();
        block_scan_boxes(s_box, false);
        Box6 run = box_identity();
        if (t > 0)
            for (int k = 0; k < 6; k++) run.v[k] = s_box[6 * (t - 1) + k];
        float *lba = lb + ((size_t)axis * n + first) * 6, *rba = rb + ((size_t)axis * n + first) * 6;
        for (uint32_t i = c0; i < c1; i++) {
            if (i >= 1)
                for (int k = 0; k < 6; k++) lba[(size_t)i * 6 + k] = run.v[k];
            run = box_join(run, load_box(boxes, srt[i]));
        }
        __syncthreads();
        // backward: right(i) = join of sorted[m-1 .. i], the later ones running
        acc = box_identity();
        for (uint32_t i = c1; i > c0; i--) acc = box_join(acc, load_box(boxes, srt[i - 1]));
        for (int k = 0; k < 6; k++) s_box[6 * t + k] = acc.v[k];
        __syncthreads();
        block_scan_boxes(s_box, true);
        run = box_identity();
        if (t + 1 < T)
            for (int k = 0; k < 6; k++) run.v[k] = s_box[6 * (t + 1) + k];
        for (uint32_t i = c1; i > c0; i--) {
            run = box_join(run, load_box(boxes, srt[i - 1]));
            if (i - 1 >= 1)
                for (int k = 0; k < 6; k++) rba[(size_t)(i - 1) * 6 + k] = run.v[k];
        }
        // the split cost at every position: the first minimum wins
        float my = __builtin_inff();
        uint32_t my_i = 0xffffffffu;
        for (uint32_t i = max(c0, 1u); i < c1; i++) {
            const float *l = lba + (size_t)i * 6, *r = rba + (size_t)i * 6;
            const float sah = cost * sa_whole + half_area(l, l + 3) * (float)i + half_area(r, r + 3) * (float)(m - i);
            if (sah < my) {
                my = sah;
                my_i = i;
            }
        }
        s_sah[t] = my;
        s_idx[t] = my_i;
        __syncthreads();
        for (uint32_t d = T / 2; d > 0; d >>= 1) {
            if (t < d) {
                const float a = s_sah[t], b = s_sah[t + d];
                const uint32_t ia = s_idx[t], ib = s_idx[t + d];
                if (b < a || (b == a && ib < ia)) {
                    s_sah[t] = b;
                    s_idx[t] = ib;
                }
            }
            __syncthreads();
        }
        if (s_idx[0] != 0xffffffffu && s_sah[0] < best) {
            best = s_sah[0];
            best_axis = axis;
            best_index = s_idx[0];
        }
        __syncthreads();
    }

    // the children's order: the chosen axis's; a leaf keeps the incoming one
    const uint32_t *src = best_axis >= 0 ? sorted + (size_t)best_axis * n + first : order_in + first;
    for (uint32_t i = t; i < m; i += T) order_out[first + i] = src[i];
    if (t == 0) {
        SahSplit sp;
        sp.axis = best_axis;
        sp.index = best_index;
        if (best_axis >= 0) {
            const float *l = lb + ((size_t)best_axis * n + first + best_index) * 6;
            const float *r = rb + ((size_t)best_axis * n + first + best_index) * 6;
            for (int k = 0; k < 3; k++) {
                sp.lmin[k] = l[k];
                sp.lmax[k] = l[3 + k];
                sp.rmin[k] = r[k];
                sp.rmax[k] = r[3 + k];
            }
        } else {
            for (int k = 0; k < 3; k++) {
                sp.lmin[k] = task.bmin[k];
                sp.lmax[k] = task.bmax[k];
                sp.rmin[k] = sp.rmax[k] = 0.0f;
            }
        }
        splits[blockIdx.x] = sp;
    }
}

// One workgroup: the level's nodes in node order.  meta[0] = interior nodes
// so far (children of the k-th at 2k+1, 2k+2), meta[1] = leaf primitives so
// far, meta[2] = this level's task count, meta[3] = the next level's.
__global__ void __launch_bounds__(kSahThreads) k_sah_emit(const SahTask *__restrict__ tasks,
                                                         const SahSplit *__restrict__ splits,
                                                         const uint32_t *__restrict__ order_in,
                                                         SahTask *__restrict__ next, rh_bvh_node *__restrict__ nodes,
                                                         uint32_t *__restrict__ out_indices, uint32_t *meta) {
    __shared__ uint32_t s_int[kSahThreads], s_leaf[kSahThreads];
    const uint32_t nt = meta[2], t = threadIdx.x, T = blockDim.x;
    const uint32_t K = (nt + T - 1) / T;
    const uint32_t c0 = min(nt, t * K), c1 = min(nt, c0 + K);
    uint32_t ni = 0, nl = 0;
    for (uint32_t i = c0; i < c1; i++) {
        if (splits[i].axis >= 0) ni++;
        else nl += tasks[i].count;
    }
    s_int[t] = ni;
    s_leaf[t] = nl;
    __syncthreads();
    for (uint32_t d = 1; d < T; d <<= 1) {
        const uint32_t a = t >= d ? s_int[t - d] : 0u, b = t >= d ? s_leaf[t - d] : 0u;
        __syncthreads();
        s_int[t] += a;
        s_leaf[t] += b;
        __syncthreads();
    }
    const uint32_t int_base = meta[0], leaf_base = meta[1];
    uint32_t ri = (t ? s_int[t - 1] : 0u), rl = (t ? s_leaf[t - 1] : 0u);
    for (uint32_t i = c0; i < c1; i++) {
        const SahTask &tk = tasks[i];
        const SahSplit &sp = splits[i];
        rh_bvh_node nd;
        if (sp.axis < 0) {
            const uint32_t off = leaf_base + rl;
            nd.prim_index = RH_LEAF_NODE_BIT + off;
            nd.prim_count = tk.count;
            for (int k = 0; k < 3; k++) {
                nd.bbox_min[k] = tk.bmin[k];
                nd.bbox_max[k] = tk.bmax[k];
            }
            for (uint32_t j = 0; j < tk.count; j++) out_indices[off + j] = order_in[tk.first + j];
            rl += tk.count;
        } else {
            const uint32_t k2 = int_base + ri;
            const uint32_t left = 2 * k2 + 1, right = 2 * k2 + 2;
            float cl[3], cr[3], dist[3];
            for (int k = 0; k < 3; k++) {
                cl[k] = (sp.lmin[k] + sp.lmax[k]) / 2;
                cr[k] = (sp.rmin[k] + sp.rmax[k]) / 2;
                dist[k] = __builtin_fabsf(cl[k] - cr[k]);
            }
            uint32_t space_axis;
            if (dist[0] > dist[1] && dist[0] > dist[2]) space_axis = 0;
            else if (dist[1] > dist[0] && dist[1] > dist[2]) space_axis = 1;
            else space_axis = 2;
            nd.prim_index = left;
            nd.prim_count = (space_axis << 30) + right;
            for (int k = 0; k < 3; k++) {
                nd.bbox_min[k] = min_std(sp.lmin[k], sp.rmin[k]);
                nd.bbox_max[k] = max_std(sp.lmax[k], sp.rmax[k]);
            }
            SahTask a, b;
            a.first = tk.first;
            a.count = sp.index;
            a.node = left;
            b.first = tk.first + sp.index;
            b.count = tk.count - sp.index;
            b.node = right;
            for (int k = 0; k < 3; k++) {
                a.bmin[k] = sp.lmin[k];
                a.bmax[k] = sp.lmax[k];
                b.bmin[k] = sp.rmin[k];
                b.bmax[k] = sp.rmax[k];
            }
            next[2 * ri] = a;
            next[2 * ri + 1] = b;
            ri++;
        }
        nodes[tk.node] = nd;
    }
    __syncthreads();
    if (t == T - 1) {
        meta[0] = int_base + s_int[T - 1];
        meta[1] = leaf_base + s_leaf[T - 1];
        meta[3] = 2 * s_int[T - 1];
    }
}

// next level's task count becomes the current one
__global__ void k_sah_advance(uint32_t *meta) {
    meta[2] = meta[3];
    meta[3] = 0;
}

} // namespace

struct rh_builder {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t cap = 0; // primitives the buffers hold
    float *prims = nullptr, *partial = nullptr, *box = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr}, *hist = nullptr;
    uint32_t *head = nullptr, *meta = nullptr, *base = nullptr, *treelets = nullptr;
    Task *tasks = nullptr;
    rh_bvh_node *nodes = nullptr;
    // the SAH build's buffers, for up to sah_cap primitives
    size_t sah_cap = 0;
    float *sah_boxes = nullptr, *sah_key = nullptr, *sah_lb = nullptr, *sah_rb = nullptr;
    uint32_t *sah_order[2] = {nullptr, nullptr}, *sah_val = nullptr, *sah_sorted = nullptr, *sah_indices = nullptr;
    uint32_t *sah_meta = nullptr; // 4 counters and the tie flag
    SahTask *sah_tasks[2] = {nullptr, nullptr};
    SahSplit *sah_splits = nullptr;
    rh_bvh_node *sah_nodes = nullptr;
};

static void free_sah_buffers(rh_builder *b) {
    for (void *p : {(void *)b->sah_boxes, (void *)b->sah_key, (void *)b->sah_lb, (void *)b->sah_rb,
                    (void *)b->sah_order[0], (void *)b->sah_order[1], (void *)b->sah_val, (void *)b->sah_sorted,
                    (void *)b->sah_indices, (void *)b->sah_meta, (void *)b->sah_tasks[0], (void *)b->sah_tasks[1],
                    (void *)b->sah_splits, (void *)b->sah_nodes})
        free_ptr(p);
    b->sah_boxes = b->sah_key = b->sah_lb = b->sah_rb = nullptr;
    b->sah_order[0] = b->sah_order[1] = b->sah_val = b->sah_sorted = b->sah_indices = b->sah_meta = nullptr;
    b->sah_tasks[0] = b->sah_tasks[1] = nullptr;
    b->sah_splits = nullptr;
    b->sah_nodes = nullptr;
    b->sah_cap = 0;
}

static void free_buffers(rh_builder *b) {
    for (void *p : {(void *)b->prims, (void *)b->keys[0], (void *)b->keys[1], (void *)b->vals[0], (void *)b->vals[1],
                    (void *)b->hist, (void *)b->tasks, (void *)b->nodes})
        free_ptr(p);
    b->prims = nullptr;
    b->keys[0] = b->keys[1] = b->vals[0] = b->vals[1] = b->hist = nullptr;
    b->tasks = nullptr;
    b->nodes = nullptr;
    b->cap = 0;
}

extern "C" {

int rh_builder_create(int device, rh_builder **out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        rh_rt::err() = "no HIP device available";
        return -1002;
    }
    if (device < 0) RH_CHECK(hipGetDevice(&device));
    if (device >= ndev) return arg_error("device index out of range");
    RH_CHECK(hipSetDevice(device));
    auto *b = new rh_builder;
    b->device = device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        rh_rt::err() = "hipStreamCreate failed";
        return -1003;
    }
    RH_CHECK(hipMalloc(&b->partial, 6 * kPartials * sizeof(float)));
    RH_CHECK(hipMalloc(&b->box, 6 * sizeof(float)));
    RH_CHECK(hipMalloc(&b->head, kTreelets * sizeof(uint32_t)));
    RH_CHECK(hipMalloc(&b->meta, (8 + 2 * kLevels) * sizeof(uint32_t)));
    RH_CHECK(hipMalloc(&b->base, kTreelets * sizeof(uint32_t)));
    RH_CHECK(hipMalloc(&b->treelets, 3 * kTreelets * sizeof(uint32_t)));
    *out = b;
    return 0;
}

void rh_builder_destroy(rh_builder *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_buffers(b);
    free_sah_buffers(b);
    for (void *p : {(void *)b->partial, (void *)b->box, (void *)b->head, (void *)b->meta, (void *)b->base,
                    (void *)b->treelets})
        free_ptr(p);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

int rh_build_lbvh(rh_builder *b, const float *prims, size_t n, uint32_t index_offset, uint32_t *out_order,
                  rh_bvh_node *out_nodes, size_t nodes_cap, uint32_t *out_node_count, uint32_t *out_treelets,
                  uint32_t *out_treelet_count) {
    *out_node_count = 0;
    *out_treelet_count = 0;
    if (!n) return 0;
    if (n > 0x7fffffffu / 2) return arg_error("too many primitives for one tree");
    if (nodes_cap < 2 * n) return arg_error("node output must hold 2 n nodes");
    RH_CHECK(hipSetDevice(b->device));
    hipStream_t s = b->stream;
    const int tiles = (int)((n + kSortTile - 1) / kSortTile);
    if (n > b->cap) {
        free_buffers(b);
        const size_t cap = std::max(n, (size_t)4096);
        const size_t tiles_cap = (cap + kSortTile - 1) / kSortTile;
        RH_CHECK(hipMalloc(&b->prims, 6 * cap * sizeof(float)));
        for (int k = 0; k < 2; k++) {
            RH_CHECK(hipMalloc(&b->keys[k], cap * sizeof(uint32_t)));
            RH_CHECK(hipMalloc(&b->vals[k], cap * sizeof(uint32_t)));
        }
        RH_CHECK(hipMalloc(&b->hist, (256 * tiles_cap + 256) * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->tasks, 2 * cap * sizeof(Task)));
        RH_CHECK(hipMalloc(&b->nodes, 2 * cap * sizeof(rh_bvh_node)));
        b->cap = cap;
    }
    const unsigned grid = (unsigned)((n + kBlock - 1) / kBlock);
    RH_CHECK(hipMemcpyAsync(b->prims, prims, 6 * n * sizeof(float), hipMemcpyHostToDevice, s));
    const int pblocks = (int)std::min<size_t>(kPartials, grid);
    hipLaunchKernelGGL(k_lbvh_bounds, dim3(pblocks), dim3(kBlock), 0, s, b->prims, (uint32_t)n, b->partial);
    hipLaunchKernelGGL(k_lbvh_scale, dim3(1), dim3(kBlock), 0, s, b->partial, pblocks, b->box);
    hipLaunchKernelGGL(k_lbvh_codes, dim3(grid), dim3(kBlock), 0, s, b->prims, (uint32_t)n, b->box, b->keys[0],
                       b->vals[0]);
    // stable sort by the whole code: equal codes keep the primitive order
    int src = 0;
    for (int shift = 0; shift < 32; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(tiles), dim3(kBlock), 0, s, b->keys[src], nullptr, (int)n, shift,
                           b->hist, tiles);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(kBlock), 0, s, b->hist, tiles,
                           b->hist + (size_t)256 * tiles);
        hipLaunchKernelGGL(k_radix_scan_bases, dim3(256), dim3(kBlock), 0, s, b->hist, tiles,
                           b->hist + (size_t)256 * tiles);
        hipLaunchKernelGGL(k_radix_scatter, dim3(tiles), dim3(kBlock), 0, s, b->keys[src], b->vals[src],
                           b->keys[src ^ 1], b->vals[src ^ 1], nullptr, (int)n, shift, b->hist, tiles);
        src ^= 1;
    }
    const uint32_t *codes = b->keys[src], *order = b->vals[src];
    RH_CHECK(hipMemsetAsync(b->head, 0xff, kTreelets * sizeof(uint32_t), s));
    RH_CHECK(hipMemsetAsync(b->meta, 0, (8 + 2 * kLevels) * sizeof(uint32_t), s));
    hipLaunchKernelGGL(k_lbvh_heads, dim3(grid), dim3(kBlock), 0, s, codes, (uint32_t)n, b->head);
    hipLaunchKernelGGL(k_lbvh_treelets, dim3(1), dim3(kBlock), 0, s, b->head, (uint32_t)n, b->tasks, b->meta);
    // a level holds disjoint ranges: at most n tasks
    for (int level = 0; level < kLevels; level++) {
        hipLaunchKernelGGL(k_lbvh_split, dim3(grid), dim3(kBlock), 0, s, b->tasks, codes, b->meta, level);
        if (level + 1 < kLevels) hipLaunchKernelGGL(k_lbvh_close_level, dim3(1), dim3(1), 0, s, b->meta, level);
    }
    for (int level = kLevels - 1; level >= 0; level--)
        hipLaunchKernelGGL(k_lbvh_up, dim3(grid), dim3(kBlock), 0, s, b->tasks, b->prims, order, b->meta, level);
    hipLaunchKernelGGL(k_lbvh_bases, dim3(1), dim3(kBlock), 0, s, b->tasks, b->meta, b->base);
    for (int level = 0; level < kLevels; level++)
        hipLaunchKernelGGL(k_lbvh_down, dim3(grid), dim3(kBlock), 0, s, b->tasks, b->meta, level);
    hipLaunchKernelGGL(k_lbvh_write, dim3(2 * grid), dim3(kBlock), 0, s, b->tasks, b->meta, b->base, index_offset,
                       b->nodes);
    hipLaunchKernelGGL(k_lbvh_treelet_out, dim3(kTreelets / kBlock), dim3(kBlock), 0, s, b->tasks, b->meta, b->base,
                       b->treelets);
    RH_CHECK(hipGetLastError());
    uint32_t meta[8 + 2 * kLevels];
    RH_CHECK(hipMemcpyAsync(meta, b->meta, sizeof(meta), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipStreamSynchronize(s));
    const uint32_t total = meta[4 + 2 * kLevels], T = meta[1];
    if (meta[0] > 2 * n || total != meta[0]) return arg_error("fast build: inconsistent node count");
    RH_CHECK(hipMemcpyAsync(out_order, order, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipMemcpyAsync(out_nodes, b->nodes, total * sizeof(rh_bvh_node), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipMemcpyAsync(out_treelets, b->treelets, 3 * T * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipStreamSynchronize(s));
    *out_node_count = total;
    *out_treelet_count = T;
    return 0;
}

int rh_build_sah(rh_builder *b, const float *boxes, size_t n, float oversplit, float cost, rh_bvh_node *out_nodes,
                 size_t nodes_cap, uint32_t *out_node_count, uint32_t *out_indices, int *host_needed) {
    *out_node_count = 0;
    *host_needed = 0;
    if (!n) return 0;
    if (n > 0x3fffffffu / 2) return arg_error("too many primitives for one tree");
    if (nodes_cap < 2 * n) return arg_error("node output must hold 2 n nodes");
    // the sort pads with +inf: boxes must be finite (else the host builds)
    for (size_t i = 0; i < 6 * n; i++)
        if (!std::isfinite(boxes[i])) {
            *host_needed = 1;
            return 0;
        }
    RH_CHECK(hipSetDevice(b->device));
    hipStream_t s = b->stream;
    if (n > b->sah_cap) {
        free_sah_buffers(b);
        const size_t cap = std::max(n, (size_t)1024);
        RH_CHECK(hipMalloc(&b->sah_boxes, 6 * cap * sizeof(float)));
        RH_CHECK(hipMalloc(&b->sah_key, 2 * cap * sizeof(float)));
        RH_CHECK(hipMalloc(&b->sah_val, 2 * cap * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_lb, 18 * cap * sizeof(float)));
        RH_CHECK(hipMalloc(&b->sah_rb, 18 * cap * sizeof(float)));
        RH_CHECK(hipMalloc(&b->sah_order[0], cap * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_order[1], cap * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_sorted, 3 * cap * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_indices, cap * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_meta, 8 * sizeof(uint32_t)));
        RH_CHECK(hipMalloc(&b->sah_tasks[0], cap * sizeof(SahTask)));
        RH_CHECK(hipMalloc(&b->sah_tasks[1], cap * sizeof(SahTask)));
        RH_CHECK(hipMalloc(&b->sah_splits, cap * sizeof(SahSplit)));
        RH_CHECK(hipMalloc(&b->sah_nodes, 2 * cap * sizeof(rh_bvh_node)));
        b->sah_cap = cap;
    }
    // the root: every primitive in index order, the box they span
    std::vector<uint32_t> iota(n);
    SahTask root;
    root.first = 0;
    root.count = (uint32_t)n;
    root.node = 0;
    for (int k = 0; k < 3; k++) {
        root.bmin[k] = 3.402823466e+38F;
        root.bmax[k] = -3.402823466e+38F;
    }
    for (size_t i = 0; i < n; i++) {
        iota[i] = (uint32_t)i;
        for (int k = 0; k < 3; k++) {
            const float lo = boxes[6 * i + k], hi = boxes[6 * i + 3 + k];
            root.bmin[k] = (lo < root.bmin[k]) ? lo : root.bmin[k];
            root.bmax[k] = (root.bmax[k] < hi) ? hi : root.bmax[k];
        }
    }
    const uint32_t meta0[8] = {0, 0, 1, 0, 0, 0, 0, 0};
    RH_CHECK(hipMemcpyAsync(b->sah_boxes, boxes, 6 * n * sizeof(float), hipMemcpyHostToDevice, s));
    RH_CHECK(hipMemcpyAsync(b->sah_order[0], iota.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    RH_CHECK(hipMemcpyAsync(b->sah_tasks[0], &root, sizeof(root), hipMemcpyHostToDevice, s));
    RH_CHECK(hipMemcpyAsync(b->sah_meta, meta0, sizeof(meta0), hipMemcpyHostToDevice, s));
    uint32_t ntasks = 1;
    int cur = 0;
    uint32_t meta[8];
    while (ntasks) {
        hipLaunchKernelGGL(k_sah_split, dim3(ntasks), dim3(kSahThreads), 0, s, b->sah_boxes, b->sah_tasks[cur],
                           b->sah_meta + 2, b->sah_order[cur], b->sah_order[cur ^ 1], b->sah_key, b->sah_val,
                           b->sah_sorted, b->sah_lb, b->sah_rb, n, oversplit, cost, b->sah_splits, b->sah_meta + 4);
        hipLaunchKernelGGL(k_sah_emit, dim3(1), dim3(kSahThreads), 0, s, b->sah_tasks[cur], b->sah_splits,
                           b->sah_order[cur], b->sah_tasks[cur ^ 1], b->sah_nodes, b->sah_indices, b->sah_meta);
        RH_CHECK(hipGetLastError());
        RH_CHECK(hipMemcpyAsync(meta, b->sah_meta, sizeof(meta), hipMemcpyDeviceToHost, s));
        RH_CHECK(hipStreamSynchronize(s));
        if (meta[4]) {
            *host_needed = 1; // equal keys: the host's std::sort order decides
            return 0;
        }
        if (2 * meta[0] + 1 > 2 * n || meta[1] > n) return arg_error("SAH build: inconsistent node count");
        ntasks = meta[3];
        if (ntasks) hipLaunchKernelGGL(k_sah_advance, dim3(1), dim3(1), 0, s, b->sah_meta);
        cur ^= 1;
    }
    if (meta[1] != n) return arg_error("SAH build: primitives lost");
    const uint32_t total = 2 * meta[0] + 1;
    RH_CHECK(hipMemcpyAsync(out_nodes, b->sah_nodes, total * sizeof(rh_bvh_node), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipMemcpyAsync(out_indices, b->sah_indices, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    RH_CHECK(hipStreamSynchronize(s));
    *out_node_count = total;
    return 0;
}

} // extern "C"
