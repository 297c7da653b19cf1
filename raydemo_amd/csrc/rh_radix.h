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

} // namespace
