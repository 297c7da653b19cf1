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
};

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

} // extern "C"
