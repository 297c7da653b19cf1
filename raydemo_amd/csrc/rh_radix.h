// Stable LSB radix sort passes over (u32 key, u32 value) pairs, 8 bits per
// pass, shared by the ray sort and the BVH build.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rh_wavefront.h"

namespace {

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


// ---------------------------------------------------------------------------
// One-sweep radix sort of the ray queue by its hash (get_ray_hash,
// CoreRef.cpp:316-333; the result is the stable sort by key that
// SortRays_CPU, CoreRef.cpp:615-657, produces).  Per sort: one clear, one
// kernel that writes the keys and counts the digits of every pass over the
// whole queue, and one scatter per 8-bit digit, the last of which moves the
// rays themselves.  A scatter block takes a tile ticket in start order,
// counts its tile's digits, publishes them, and sums the counts of the
// tiles before it from their published words (decoupled look-back: a tile
// whose prefix is known publishes it, so a walk stops there).  A block only
// ever waits for tiles whose tickets were taken before its own, by blocks
// already running, so every wait ends; the walk is also bounded
// (kLookbackSpins polls of one word) and counts a give-up in `stalls`.
//
// Scratch layout (32-bit words), sort_scratch_words(max_n) in all:
//   [kSortTotals, +4 x 256)  digit counts of the whole queue per digit position
//   [kSortTickets, +4)        tile tickets of the scatter passes
//   [kSortStatus, ...)        per pass, per tile, per digit: the tile's count
//                             (flag kTileAggregate) or the count of every
//                             tile up to and including it (kTilePrefix)
constexpr int kSweepItems = 4;                   // keys per thread of a one-sweep tile
constexpr int kSweepTile = kBlock * kSweepItems; // keys per tile (4 waves x 256)
constexpr uint32_t kSortTotals = 0, kSortTickets = 1024, kSortStatus = 1280;
constexpr uint32_t kTileAggregate = 1u << 30, kTilePrefix = 2u << 30, kTileValue = kTileAggregate - 1u;
constexpr int kLookbackSpins = 1 << 20;
// largest queue the 30-bit counts of the status words hold
constexpr long long kSortMaxRays = (1ll << 30) - 1;

__host__ __device__ constexpr size_t sort_tiles(size_t max_n) { return (max_n + kSweepTile - 1) / kSweepTile; }
__host__ __device__ constexpr size_t sort_scratch_words(size_t max_n) {
    return kSortStatus + 4 * 256 * (sort_tiles(max_n) > 0 ? sort_tiles(max_n) : 1);
}

// Zeroes the totals, the tickets and the status rows of this sort's passes
// for the tiles of the device count.
__global__ void __launch_bounds__(256) k_sort_clear(uint32_t *__restrict__ scr, const int *__restrict__ count,
                                                    int n_static, int npass, int tiles_static) {
    const int n = count ? *count : n_static;
    const size_t tiles = sort_tiles((size_t)(n > 0 ? n : 0));
    const size_t plane = tiles * 256, total = kSortStatus + (size_t)npass * plane;
    for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < total; w += (size_t)gridDim.x * blockDim.x) {
        if (w < kSortStatus) {
            scr[w] = 0u;
        } else {
            const size_t r = w - kSortStatus, pass = r / plane, o = r % plane;
            scr[kSortStatus + pass * (size_t)tiles_static * 256 + o] = 0u;
        }
    }
}

// Keys (the ray hash) and the identity permutation of the queue, and the
// digit counts of every pass (shift0 + 8 p, p < npass) over the whole queue.
__global__ void __launch_bounds__(kBlock) k_sort_keys(const rh_ray *__restrict__ rays, const int *__restrict__ count,
                                                      int n_static, PassConst pc, uint32_t *__restrict__ keys,
                                                      uint32_t *__restrict__ vals, uint32_t *__restrict__ scr,
                                                      int shift0, int npass) {
    __shared__ uint32_t h[4][256];
    const int n = count ? *count : n_static;
    const int base = blockIdx.x * kSweepTile;
    if (base >= n) return;
    const int t = threadIdx.x;
#pragma unroll
    for (int p = 0; p < 4; p++) h[p][t] = 0u;
    float4 a[kSweepItems], b[kSweepItems];
#pragma unroll
    for (int r = 0; r < kSweepItems; r++) { // the loads first, all in flight together
        const int i = base + r * kBlock + t;
        if (i < n) {
            const float4 *q = reinterpret_cast<const float4 *>(rays + i);
            a[r] = q[0];
            b[r] = q[1];
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSweepItems; r++) {
        const int i = base + r * kBlock + t;
        if (i < n) {
            const uint32_t k = ray_hash(f3{a[r].x, a[r].y, a[r].z}, f3{b[r].x, b[r].y, b[r].z}, pc.root_min, pc.cell_size);
            keys[i] = k;
            vals[i] = (uint32_t)i;
            for (int p = 0; p < npass; p++) atomicAdd(&h[p][(k >> (shift0 + 8 * p)) & 0xffu], 1u);
        }
    }
    __syncthreads();
    for (int p = 0; p < npass; p++) {
        const uint32_t v = h[p][t];
        if (v) atomicAdd(scr + kSortTotals + 256 * p + t, v);
    }
}

// One stable scatter pass by the digit at `shift` (pass index `pass`).
// Gather: the last pass, which writes the rays (rays_in[value]) at the
// sorted positions instead of the keys and values.  Wave w of a block owns
// the w-th quarter of the tile, 16 rounds of 64 keys in index order; it ranks
// each key among its wave's earlier keys of the same digit (peers found with
// eight ballots, one per digit bit; its own running digit counts in LDS), so
// no block-wide barrier is needed per round.  The waves' counts then give
// the tile's digit counts and each wave's offset within a digit.
template <bool Gather>
__global__ void __launch_bounds__(kBlock) k_sort_scatter(const uint32_t *__restrict__ keys_in,
                                                         const uint32_t *__restrict__ vals_in,
                                                         uint32_t *__restrict__ keys_out, uint32_t *__restrict__ vals_out,
                                                         const int *__restrict__ count, int n_static, int shift, int pass,
                                                         uint32_t *__restrict__ scr, int tiles_static,
                                                         const rh_ray *__restrict__ rays_in,
                                                         const rh_ray_aux *__restrict__ aux_in,
                                                         rh_ray *__restrict__ rays_out, rh_ray_aux *__restrict__ aux_out,
                                                         int *__restrict__ count_out,
                                                         unsigned long long *__restrict__ stalls) {
    constexpr int kWaves = kBlock / 64, kWaveKeys = kSweepTile / kWaves;
    __shared__ uint32_t s_tile;
    __shared__ uint32_t run[256];       // the digit's first position in the output, for this tile
    __shared__ uint32_t wc[kWaves][256]; // per-wave digit counts, then each wave's offset within the digit
    __shared__ uint32_t wsum[kWaves];
    const int n = count ? *count : n_static;
    const int tiles = (int)sort_tiles((size_t)(n > 0 ? n : 0));
    if (Gather && count_out && blockIdx.x == 0 && threadIdx.x == 0) *count_out = n; // the sorted queue's count
    if ((int)blockIdx.x >= tiles) return; // exactly `tiles` blocks take tickets
    const int t = threadIdx.x, wave = t >> 6;
    const uint32_t lane = lane_id();
    if (t == 0) s_tile = atomicAdd(scr + kSortTickets + pass, 1u);
#pragma unroll
    for (int w = 0; w < kWaves; w++) wc[w][t] = 0u;
    __syncthreads();
    const uint32_t tile = s_tile;
    const int first = (int)tile * kSweepTile + wave * kWaveKeys; // this wave's first key

    uint32_t k[kSweepItems], v[kSweepItems], local[kSweepItems];
#pragma unroll
    for (int r = 0; r < kSweepItems; r++) {
        const int i = first + r * 64 + (int)lane;
        k[r] = 0u;
        v[r] = 0u;
        if (i < n) {
            k[r] = keys_in[i];
            v[r] = vals_in[i];
        }
    }
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
    for (int r = 0; r < kSweepItems; r++) {
        const bool active = first + r * 64 + (int)lane < n;
        const uint32_t digit = (k[r] >> shift) & 0xffu;
        uint64_t peers = __ballot(active);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        const uint32_t before = active ? wc[wave][digit] : 0u;
        local[r] = before + rank;
        // the wave's lanes read the count before its first lane of the digit adds to it
        if (active && rank == 0) wc[wave][digit] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // digit t: the tile's count, and each wave's offset among the tile's keys of the digit
    uint32_t mine = 0u;
#pragma unroll
    for (int w = 0; w < kWaves; w++) {
        const uint32_t c = wc[w][t];
        wc[w][t] = mine;
        mine += c;
    }

    // publish, then sum the tiles before this one (thread t: digit t)
    uint32_t *status = scr + kSortStatus + (size_t)pass * tiles_static * 256;
    __hip_atomic_store(status + (size_t)tile * 256 + t, (tile == 0 ? kTilePrefix : kTileAggregate) | mine,
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    uint32_t excl = 0u;
    if (tile > 0) {
        int j = (int)tile - 1;
        int spins = 0;
        for (;;) {
            const uint32_t w = __hip_atomic_load(status + (size_t)j * 256 + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t flag = w & ~kTileValue;
            if (flag == 0u) {
                if (++spins > kLookbackSpins) {
                    if (stalls) atomicAdd(stalls, 1ull);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
                continue;
            }
            excl += w & kTileValue;
            if (flag == kTilePrefix) break;
            --j; // tile 0 publishes a prefix, so j stays >= 0
        }
        __hip_atomic_store(status + (size_t)tile * 256 + t, kTilePrefix | (excl + mine), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }

    // the digit's first position: counts of the smaller digits over the whole
    // queue, plus this digit's count in the earlier tiles
    const uint32_t tot = scr[kSortTotals + 256 * pass + t];
    const uint32_t inc = wave_inclusive_scan(tot);
    if ((t & 63) == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t below = 0u;
    for (int w = 0; w < wave; w++) below += wsum[w];
    run[t] = below + inc - tot + excl;
    __syncthreads();

    if (Gather) {
        // every ray of the thread loaded before any is stored
        rh_ray rr[kSweepItems];
        rh_ray_aux ra[kSweepItems];
#pragma unroll
        for (int r = 0; r < kSweepItems; r++) {
            if (first + r * 64 + (int)lane < n) {
                rr[r] = rays_in[v[r]];
                ra[r] = aux_in[v[r]];
            }
        }
#pragma unroll
        for (int r = 0; r < kSweepItems; r++) {
            if (first + r * 64 + (int)lane >= n) continue;
            const uint32_t digit = (k[r] >> shift) & 0xffu;
            const uint32_t pos = run[digit] + wc[wave][digit] + local[r];
            rays_out[pos] = rr[r];
            aux_out[pos] = ra[r];
        }
    } else {
#pragma unroll
        for (int r = 0; r < kSweepItems; r++) {
            if (first + r * 64 + (int)lane >= n) continue;
            const uint32_t digit = (k[r] >> shift) & 0xffu;
            const uint32_t pos = run[digit] + wc[wave][digit] + local[r];
            keys_out[pos] = k[r];
            vals_out[pos] = v[r];
        }
    }
}

} // namespace
