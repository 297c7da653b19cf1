// Many scenes in one wavefront: the batch pipeline of the HIP backend.
//
// Every scene of a batch owns a contiguous segment of the shared queues
// (rays, hits, framebuffers, next-event-estimation slots) and its own queue
// counters.  Each stage is one persistent kernel over the whole batch: its
// blocks take (scene, 256-ray tile) work items from an atomic counter, the
// tiles of scene s following those of scene s-1, and run the same stage body
// as the single-scene pipeline (rh_wavefront.h) with the scene's arguments, so
// every pixel of every scene is computed exactly as a lone RenderScene pass
// computes it.  Small and large scenes share the GPU instead of leaving it
// idle in turn.  Per-pass state (iteration, active scenes, counters) lives on
// the device and is set by the first kernel of the pass, so the launch
// sequence of a pass is fixed and is captured once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>

#include "rh_runtime.h"

using rh_rt::arg_error;
using rh_rt::DevBuf;
using rh_rt::fail;
using rh_rt::free_ptr;
using rh_rt::upload_halton;

namespace {

constexpr int kMaxBatchScenes = 256;
// tile counters of one pass: generate; per staged bounce five, and with the
// sort two more plus two per radix pass; two tails; mix
constexpr int kMaxStages = 15 * (RH_MAX_BOUNCES_EXT + 1) + 4;

// one scene of the batch (device table)
struct BatchEntry {
    Scene sc;
    PassConst pc;      // camera and frame; iteration and mix_k are set per pass on the device
    rh_pass_info pi;   // pass settings (bounce set per stage)
    uint32_t offset;   // first pixel of the scene in the framebuffers
    uint32_t q_offset; // first ray of the scene in the queue arrays (within its group)
    uint32_t pixels;   // w * h
    uint32_t nee_slots;  // light samples per shaded ray (0: visibility inline)
    uint32_t nee_offset; // first slot of the scene in the NEE arrays (within its group)
    int spp;             // passes the scene takes part in
    int depth;           // max_total_depth (clamped)
};

struct BatchDev {
    const BatchEntry *__restrict__ table;
    int n_scenes;
    int s_begin, s_end; // the scenes this launch works on (a group: their rays share the queues)
    int *pass;          // passes started so far (device)
    int *counts;        // [(RH_MAX_BOUNCES_EXT + 2) * n_scenes]: rays entering stage b (b = 0: primary)
    int *shadow_counts; // [(RH_MAX_BOUNCES_EXT + 2) * n_scenes]
    int *work;          // [kMaxStages] tile counters
    rh_ray *rays[2];
    rh_ray_aux *aux[2];
    rh_hit *hits;
    float4 *temp, *clean, *final_;
    float4 *nee_ray, *nee_term;
    int2 *nee_aux;
    uint8_t *nee_vis;
    float4 *pend_c, *pend_alb;
    int *pend_count;
    uint32_t *shadow_queue;
    uint32_t *spill;
    uint32_t spill_stride;
    uint32_t lds_depth;
    const float *halton;
    unsigned long long *stats; // [0] primary [1] secondary [2] shadow [4] overflow
    uint32_t *keys[2], *vals[2]; // ray sort: keys and ray indices (ping-pong), indexed as the rays
    uint32_t *hist;              // ray sort: digit counts, 256 per tile of kSortTile keys
};

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Table entry s (s uniform) read through the constant address space: the
// loads become scalar and the scene's pointers stay in SGPRs, as the kernel
// arguments of the single-scene kernels do.
typedef const __attribute__((address_space(4))) BatchEntry ConstEntry;
__device__ __forceinline__ BatchEntry entry(const BatchDev &b, int s) {
#if defined(__HIP_DEVICE_COMPILE__)
    return ((ConstEntry *)b.table)[s];
#else
    return b.table[s]; // the host pass only parses device functions
#endif
}

// s_pre[s] = the first global tile of scene s, s_pre[S] = all tiles, for
// tiles of TileRays rays of each active scene's counts[s] rays (S <= 256: one
// scene per thread of the 256-thread block).  Every thread calls it.
template <int TileRays, typename Active>
__device__ __forceinline__ void scene_tile_prefix(const BatchDev &b, const int *counts, Active active, int *s_pre) {
    const int S = b.n_scenes;
    int tiles = 0;
    if ((int)threadIdx.x < S) {
        const int s = threadIdx.x;
        if (s >= b.s_begin && s < b.s_end && active(s)) tiles = (counts[s] + TileRays - 1) / TileRays;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = tiles;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off);
        if (lane >= off) incl += v;
    }
    if (lane == 63) s_pre[wave] = incl; // wave totals (temporary)
    __syncthreads();
    int wave_base = 0;
    for (int w = 0; w < wave; w++) wave_base += s_pre[w];
    __syncthreads();
    if ((int)threadIdx.x < S) s_pre[threadIdx.x] = wave_base + incl - tiles;
    if ((int)threadIdx.x == S - 1) s_pre[S] = wave_base + incl;
    __syncthreads();
}

// Runs body(scene, tile base, global tile) for every TileRays-ray tile of
// every scene's queue, count_s = counts[s] rays (0 for scenes that skip this
// stage), the tiles numbered scene after scene.  Persistent: blocks take
// runs of consecutive tiles from *work until none is left.  One counter
// hands them out and same-address atomics serialise (a 4K frame is 32k
// tiles a stage), so a block takes about a quarter of its fair share per
// grab, and one tile at a time once there are few tiles per block.
template <int TileRays = kBlock, typename Active, typename Body>
__device__ __forceinline__ void for_each_tile(const BatchDev &b, const int *counts, int *work, Active active,
                                              Body body) {
    __shared__ int s_pre[kMaxBatchScenes + 1];
    __shared__ int s_t;
    const int S = b.n_scenes;
    scene_tile_prefix<TileRays>(b, counts, active, s_pre);
    const int total = s_pre[S];
    const int chunk = max(1, total / (int)(4 * gridDim.x));
    for (;;) {
        if (threadIdx.x == 0) s_t = atomicAdd(work, chunk);
        __syncthreads();
        const int t0 = uniform(s_t);
        __syncthreads();
        if (t0 >= total) break;
        const int t1 = min(t0 + chunk, total);
        for (int t = t0; t < t1; t++) {
            int lo = 0, hi = S; // s_pre[lo] <= t < s_pre[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= t) lo = mid;
                else hi = mid;
            }
            const int s = uniform(lo);
            if constexpr (std::is_invocable_v<Body, int, uint32_t, int>) body(s, (uint32_t)(t - s_pre[s]) * TileRays, t);
            else body(s, (uint32_t)(t - s_pre[s]) * TileRays);
            __syncthreads(); // the body's LDS is reused by the next tile
        }
    }
}

// for_each_tile for bodies whose lanes work alone (the traversal kernels:
// per-lane LDS stack columns and wave-level counters, no block-wide LDS or
// barrier): each wave takes runs of 64-ray tiles by itself, so a wave whose
// rays are done takes new ones while its block's other waves still trace.
// The body gets base = the tile's first ray minus 64 x the wave's index
// within the block (unsigned arithmetic), so that base + threadIdx.x is the
// lane's ray.
template <typename Active, typename Body>
__device__ __forceinline__ void for_each_wave_tile(const BatchDev &b, const int *counts, int *work, Active active,
                                                   Body body) {
    __shared__ int s_pre[kMaxBatchScenes + 1];
    const int S = b.n_scenes;
    scene_tile_prefix<64>(b, counts, active, s_pre);
    const int total = s_pre[S];
    const int waves = (int)(gridDim.x * (blockDim.x >> 6));
    const int chunk = max(1, total / (4 * waves));
    const uint32_t wave_first = threadIdx.x & ~63u;
    for (;;) {
        int t0 = 0;
        if (lane_id() == 0) t0 = atomicAdd(work, chunk);
        t0 = uniform(__shfl(t0, 0));
        if (t0 >= total) break;
        const int t1 = min(t0 + chunk, total);
        for (int t = t0; t < t1; t++) {
            int lo = 0, hi = S; // s_pre[lo] <= t < s_pre[hi]
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_pre[mid] <= t) lo = mid;
                else hi = mid;
            }
            const int s = uniform(lo);
            body(s, (uint32_t)(t - s_pre[s]) * 64u - wave_first);
        }
    }
}

// First kernel of a pass: advances the pass counter, sets every scene's
// iteration and mixing weight, clears the queue and tile counters, and
// activates the primary queue of scenes still rendering.
// stride: passes between two passes of this slot's pass counter (2 with two
// passes in flight, each slot taking every other pass).
__global__ void k_b_begin_pass(BatchDev b, BatchEntry *table, int n_stages, int stride) {
    const int pass = *b.pass + stride; // every thread reads the old value
    __syncthreads();
    const int S = b.n_scenes;
    const int nc = (RH_MAX_BOUNCES_EXT + 2) * S;
    for (int k = threadIdx.x; k < nc; k += blockDim.x) {
        b.counts[k] = (k < S && pass <= table[k].spp) ? (int)table[k].pixels : 0;
        b.shadow_counts[k] = 0;
    }
    for (int k = threadIdx.x; k < n_stages; k += blockDim.x) b.work[k] = 0;
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
        table[s].pc.iteration = pass;
        table[s].pi.iteration = pass;
        table[s].pc.mix_k = 1.0f / pass;
    }
    __syncthreads();
    if (threadIdx.x == 0) *b.pass = pass;
}

__global__ void __launch_bounds__(kBlock) k_b_gen(BatchDev b, int work) {
    for_each_tile(b, b.counts, b.work + work, [](int) { return true; }, [&](int s, uint32_t base) {
        const BatchEntry e = entry(b, s);
        const int n = b.counts[s];
        const int i = (int)(base + threadIdx.x);
        if (i >= n) return;
        const int w = e.pc.w;
        if (i == 0) atomicAdd(b.stats + 0, (unsigned long long)n);
        store_ray(b.rays[0] + e.q_offset, b.aux[0] + e.q_offset, (uint32_t)i, gen_primary_ray(e.pc, b.halton, i % w, i / w));
    });
}

// stage b: 0 = primary, 1 + k = bounce k
__device__ __forceinline__ bool stage_active(const BatchEntry &e, int stage) {
    if (stage == 0) return true;
    if (e.pi.settings.flags & RH_SKIP_INDIRECT_LIGHT) return false;
    return stage - 1 < e.depth;
}

// Tree: the trees of the batch's scenes (kTreeAny when they are mixed)
template <int Tree> __global__ void __launch_bounds__(kBlock) k_b_trace(BatchDev b, int stage, int q, int work) {
    extern __shared__ uint32_t lds_stack[];
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_wave_tile(b, counts, b.work + work, [&](int s) { return stage_active(b.table[s], stage); },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      TraceArgs a;
                      a.sc = e.sc;
                      a.rays = b.rays[q] + e.q_offset;
                      a.count = counts + s;
                      a.n = 0;
                      a.tmax = nullptr;
                      a.hits = b.hits + e.q_offset;
                      a.spill = b.spill + e.q_offset;
                      a.spill_stride = b.spill_stride;
                      a.lds_depth = b.lds_depth;
                      a.ray_counter = b.stats + (stage == 0 ? 0 : 1);
                      a.overflow = b.stats + 4;
                      a.visits = nullptr;
                      if (stage == 0) a.ray_counter = nullptr; // primary rays are counted at generation
                      trace_body<false, false, Tree>(a, base, lds_stack);
                  });
}

__device__ __forceinline__ ShadeArgs batch_shade_args(const BatchDev &b, const BatchEntry &e, int s, int stage, int q,
                                                      bool defer) {
    ShadeArgs a = {};
    a.sc = e.sc;
    a.rays = b.rays[q] + e.q_offset;
    a.aux = b.aux[q] + e.q_offset;
    a.hits = b.hits + e.q_offset;
    a.count = b.counts + stage * b.n_scenes + s;
    a.pi = e.pi;
    a.pi.bounce = stage == 0 ? 2 : stage + 2;
    a.w = e.pc.w;
    a.halton = b.halton;
    a.temp = b.temp + e.offset;
    a.primary = stage == 0 ? 1 : 0;
    a.out_rays = b.rays[q ^ 1] + e.q_offset;
    a.out_aux = b.aux[q ^ 1] + e.q_offset;
    a.out_count = b.counts + (stage + 1) * b.n_scenes + s;
    a.spill = b.spill + e.q_offset;
    a.spill_stride = b.spill_stride;
    a.lds_depth = b.lds_depth;
    a.shadow_counter = b.stats + 2;
    a.overflow = b.stats + 4;
    a.ray_counter = b.stats + 1;
    if (defer && e.nee_slots) {
        a.nee_ray = b.nee_ray + e.nee_offset;
        a.nee_aux = b.nee_aux + e.nee_offset;
        a.nee_term = b.nee_term + e.nee_offset;
        a.nee_slots = e.nee_slots;
        a.pend_c = b.pend_c + e.q_offset;
        a.pend_alb = b.pend_alb + e.q_offset;
        a.pend_count = b.pend_count + e.q_offset;
        a.shadow_queue = b.shadow_queue + e.nee_offset;
        a.shadow_count = b.shadow_counts + stage * b.n_scenes + s;
    }
    return a;
}

// Shading with next-event estimation deferred to the shadow stage for scenes
// with 1..8 lights; scenes with more trace visibility inline (Defer = false).
template <bool Defer> __global__ void __launch_bounds__(kBlock) k_b_shade(BatchDev b, int stage, int q, int work) {
    extern __shared__ uint32_t lds_stack[];
    __shared__ int s_app[kMaxDeferredLights * (kBlock / 64) + 1];
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile(b, counts, b.work + work,
                  [&](int s) { return stage_active(b.table[s], stage) && ((b.table[s].nee_slots != 0) == Defer); },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      const ShadeArgs a = batch_shade_args(b, e, s, stage, q, Defer);
                      shade_body<Defer>(a, base, lds_stack, s_app);
                  });
}

// Solid: every scene of the batch has solid triangles only (Scene::all_solid)
template <int Tree, bool Solid> __global__ void __launch_bounds__(kBlock) k_b_shadow(BatchDev b, int stage, int work) {
    extern __shared__ uint32_t lds_stack[];
    const int *counts = b.shadow_counts + stage * b.n_scenes;
    for_each_wave_tile(b, counts, b.work + work, [&](int s) { return b.table[s].nee_slots != 0; },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      ShadowArgs a;
                      a.sc = e.sc;
                      a.queue = b.shadow_queue + e.nee_offset;
                      a.count = counts + s;
                      a.nee_ray = b.nee_ray + e.nee_offset;
                      a.nee_aux = b.nee_aux + e.nee_offset;
                      a.nee_vis = b.nee_vis + e.nee_offset;
                      a.halton = b.halton;
                      a.spill = b.spill + e.nee_offset;
                      a.spill_stride = b.spill_stride;
                      a.lds_depth = b.lds_depth;
                      a.shadow_counter = b.stats + 2;
                      a.overflow = b.stats + 4;
                      a.visits = nullptr;
                      shadow_body<Tree, false, Solid>(a, base, lds_stack);
                  });
}

__global__ void __launch_bounds__(kBlock) k_b_combine(BatchDev b, int stage, int work) {
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile(b, counts, b.work + work,
                  [&](int s) { return stage_active(b.table[s], stage) && b.table[s].nee_slots != 0; },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      CombineArgs a;
                      a.count = counts + s;
                      a.n = 0;
                      a.slots = e.nee_slots;
                      a.pend_count = b.pend_count + e.q_offset;
                      a.pend_c = b.pend_c + e.q_offset;
                      a.pend_alb = b.pend_alb + e.q_offset;
                      a.nee_term = b.nee_term + e.nee_offset;
                      a.nee_vis = b.nee_vis + e.nee_offset;
                      a.temp = b.temp + e.offset;
                      // not read while temp is set; a real address keeps the
                      // per-ray form's address arithmetic off a null pointer
                      a.out_rgba = a.temp;
                      a.rec = AdjRecords{};
                      combine_body(a, base);
                  });
}

// ---- ray sort per scene (the single scene's sort on each scene's segment) ----
// Before a staged bounce is traced, every scene's queue is sorted stably by
// the get_ray_hash key of its rays (CoreRef.cpp:316-333, over the scene's own
// hash grid) or by the key's top sort_bits bits: keys and identity indices,
// then 8-bit LSB radix passes (per-tile digit counts, per-scene scan, stable
// scatter), then the rays gathered into the other queue.  Tiles of kSortTile
// keys never straddle scenes, and the digit counts are tile-major
// (hist[tile * 256 + digit]).  The order of a queue changes no pixel.

__device__ __forceinline__ bool sort_active(const BatchDev &b, int s, int stage) {
    return stage >= 1 && stage_active(b.table[s], stage);
}

__global__ void __launch_bounds__(kBlock) k_b_sort_keys(BatchDev b, int stage, int q, int work) {
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile(b, counts, b.work + work, [&](int s) { return sort_active(b, s, stage); },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      const int i = (int)(base + threadIdx.x);
                      if (i >= counts[s]) return;
                      const float4 *p = reinterpret_cast<const float4 *>(b.rays[q] + e.q_offset + i);
                      const float4 r0 = p[0], r1 = p[1];
                      b.keys[0][e.q_offset + i] =
                          ray_hash(f3{r0.x, r0.y, r0.z}, f3{r1.x, r1.y, r1.z}, e.pc.root_min, e.pc.cell_size);
                      b.vals[0][e.q_offset + i] = (uint32_t)i;
                  });
}

// digit counts of each kSortTile-key tile
__global__ void __launch_bounds__(kBlock) k_b_sort_hist(BatchDev b, int stage, int src, int shift, int work) {
    __shared__ uint32_t h[256];
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile<kSortTile>(b, counts, b.work + work, [&](int s) { return sort_active(b, s, stage); },
                             [&](int s, uint32_t base, int t) {
                                 const BatchEntry e = entry(b, s);
                                 const int n = counts[s];
                                 const uint32_t *keys = b.keys[src] + e.q_offset;
                                 h[threadIdx.x] = 0;
                                 __syncthreads();
                                 for (int r = 0; r < kSortItems; r++) {
                                     const int i = (int)base + r * kBlock + (int)threadIdx.x;
                                     if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
                                 }
                                 __syncthreads();
                                 b.hist[(size_t)t * 256 + threadIdx.x] = h[threadIdx.x];
                             });
}

// One block per scene: thread d turns digit d's counts over the scene's tiles
// into each tile's first output position of the digit (the digit's base = the
// scene's keys of smaller digits, plus the counts of earlier tiles).
__global__ void __launch_bounds__(256) k_b_sort_scan(BatchDev b, int stage) {
    __shared__ int s_pre[kMaxBatchScenes + 1];
    __shared__ uint32_t s_tot[256];
    const int *counts = b.counts + stage * b.n_scenes;
    scene_tile_prefix<kSortTile>(b, counts, [&](int s) { return sort_active(b, s, stage); }, s_pre);
    const int s = blockIdx.x;
    const int first = s_pre[s], last = s_pre[s + 1];
    if (first == last) return; // block-uniform: no tiles for this scene
    const int d = threadIdx.x;
    uint32_t tot = 0;
    for (int t = first; t < last; t++) tot += b.hist[(size_t)t * 256 + d];
    s_tot[d] = tot;
    __syncthreads();
    uint32_t run = 0;
    for (int k = 0; k < d; k++) run += s_tot[k];
    for (int t = first; t < last; t++) {
        const uint32_t v = b.hist[(size_t)t * 256 + d];
        b.hist[(size_t)t * 256 + d] = run;
        run += v;
    }
}

// Stable scatter of each tile (as k_radix_scatter: per round of 256 keys, each
// wave ranks the lanes sharing its digit with eight ballots, the four waves'
// counts are prefix-summed in LDS, rounds in order).
__global__ void __launch_bounds__(kBlock) k_b_sort_scatter(BatchDev b, int stage, int src, int shift, int work) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcount[4][256];
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile<kSortTile>(
        b, counts, b.work + work, [&](int s) { return sort_active(b, s, stage); },
        [&](int s, uint32_t base, int tile) {
            const BatchEntry e = entry(b, s);
            const int n = counts[s];
            const uint32_t *keys_in = b.keys[src] + e.q_offset, *vals_in = b.vals[src] + e.q_offset;
            uint32_t *keys_out = b.keys[src ^ 1] + e.q_offset, *vals_out = b.vals[src ^ 1] + e.q_offset;
            const int t = threadIdx.x, wave = t >> 6;
            const uint32_t lane = lane_id();
            run[t] = b.hist[(size_t)tile * 256 + t];
            const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
            for (int r = 0; r < kSortItems; r++) {
                const int i = (int)base + r * kBlock + t;
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
                for (int k = 0; k < 8; k++) {
                    const uint64_t m = __ballot((digit >> k) & 1u);
                    peers &= ((digit >> k) & 1u) ? m : ~m;
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
                run[t] += wcount[0][t] + wcount[1][t] + wcount[2][t] + wcount[3][t];
                if (active) {
                    keys_out[pos] = key;
                    vals_out[pos] = val;
                }
                __syncthreads();
            }
        });
}

// the rays of queue q in sorted order into queue q ^ 1
__global__ void __launch_bounds__(kBlock) k_b_sort_gather(BatchDev b, int stage, int q, int sorted, int work) {
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_tile(b, counts, b.work + work, [&](int s) { return sort_active(b, s, stage); },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      const int i = (int)(base + threadIdx.x);
                      if (i >= counts[s]) return;
                      const uint32_t j = b.vals[sorted][e.q_offset + i];
                      b.rays[q ^ 1][e.q_offset + i] = b.rays[q][e.q_offset + j];
                      b.aux[q ^ 1][e.q_offset + i] = b.aux[q][e.q_offset + j];
                  });
}

// the rays entering `stage` follow their paths to each scene's depth in one
// lane; Defer: scenes whose light samples are resolved after shading (1..8 lights);
// Solid: every scene of the batch has solid triangles only
template <bool Defer, int Tree, bool Solid>
__global__ void __launch_bounds__(kBlock) k_b_tail(BatchDev b, int stage, int q, int work) {
    extern __shared__ uint32_t lds_stack[];
    const int *counts = b.counts + stage * b.n_scenes;
    for_each_wave_tile(b, counts, b.work + work,
                  [&](int s) { return stage_active(b.table[s], stage) && ((b.table[s].nee_slots != 0) == Defer); },
                  [&](int s, uint32_t base) {
                      const BatchEntry e = entry(b, s);
                      const ShadeArgs a = batch_shade_args(b, e, s, stage, q, Defer);
                      path_tail_body<Defer, Tree, false, Solid>(a, stage - 1, e.depth, base, lds_stack);
                  });
}

__global__ void __launch_bounds__(kBlock) k_b_mix(BatchDev b, int work) {
    for_each_tile(b, b.counts, b.work + work, [](int) { return true; }, [&](int s, uint32_t base) {
        const BatchEntry e = entry(b, s);
        const int i = (int)(base + threadIdx.x);
        if (i >= (int)e.pixels || b.counts[s] == 0) return;
        mix_tonemap_pixel(b.temp + e.offset, b.clean + e.offset, b.final_ + e.offset, (size_t)i, e.pc.mix_k,
                          e.pc.cam.dtype, e.pc.cam.gamma, e.pc.cam.pass_settings.flags);
    });
}

__global__ void k_b_fill(float4 *buf, size_t n, float4 v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) buf[i] = v;
}

} // namespace

// The working set of one batch pass in flight: its stream, its copy of the
// scene table (the pass writes iteration and mixing weight into it), its
// pass counter, queue and tile counters, the queues, the pass framebuffer,
// the NEE buffers, spill slices, sort scratch and captured graph.  With
// several passes in flight (option pass_overlap: passes in flight besides the
// newest) rh_batch's fields of these names are the bound slot's and store[]
// holds the other slots' (bind_slot).
constexpr int kMaxBatchSlots = 4;
struct BatchSlot {
    hipStream_t stream = nullptr;
    BatchEntry *table = nullptr;
    int *pass = nullptr;
    int *counts = nullptr, *shadow_counts = nullptr, *work = nullptr;
    rh_ray *rays[2] = {nullptr, nullptr};
    rh_ray_aux *aux[2] = {nullptr, nullptr};
    rh_hit *hits = nullptr;
    float4 *temp = nullptr;
    float4 *nee_ray = nullptr, *nee_term = nullptr;
    int2 *nee_aux = nullptr;
    uint8_t *nee_vis = nullptr;
    float4 *pend_c = nullptr, *pend_alb = nullptr;
    int *pend_count = nullptr;
    uint32_t *shadow_queue = nullptr;
    uint32_t *spill = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_key[5] = {-1, -1, -1, -1, -1};
};
#define RH_BATCH_SLOT_FIELDS(X)                                                                                        \
    X(stream) X(table) X(pass) X(counts) X(shadow_counts) X(work) X(rays) X(aux) X(hits) X(temp) X(nee_ray)            \
        X(nee_term) X(nee_aux) X(nee_vis) X(pend_c) X(pend_alb) X(pend_count) X(shadow_queue) X(spill) X(keys)        \
            X(vals) X(hist) X(graph) X(graph_exec) X(graph_key)

struct rh_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    // passes in flight: the other slots (store[bound] unused: the bound slot
    // lives in the fields above and below), which slot is bound, the mix order
    BatchSlot store[kMaxBatchSlots];
    int bound = 0;
    int pass_overlap = kMaxBatchSlots - 1;
    hipEvent_t ev_mix = nullptr; // the last pass's mix (passes mix in pass order)
    bool mix_recorded = false;

    std::vector<std::vector<DevBuf>> scene_bufs;
    std::vector<BatchEntry> entries; // host copy of the table
    bool built = false;
    int max_depth = 0;
    int max_spp = 0;
    size_t total_pixels = 0, total_slots = 0;
    // scenes in groups whose rays fit the queue limit, rendered one group after
    // another within a pass: [first, end) per group, and the largest group's rays and NEE slots
    std::vector<std::pair<int, int>> groups;
    size_t queue_pixels = 0, queue_slots = 0;

    BatchEntry *table = nullptr;
    int *pass = nullptr;
    int *counts = nullptr, *shadow_counts = nullptr, *work = nullptr;
    rh_ray *rays[2] = {nullptr, nullptr};
    rh_ray_aux *aux[2] = {nullptr, nullptr};
    rh_hit *hits = nullptr;
    float4 *temp = nullptr, *clean = nullptr, *final_ = nullptr;
    float4 *nee_ray = nullptr, *nee_term = nullptr;
    int2 *nee_aux = nullptr;
    uint8_t *nee_vis = nullptr;
    float4 *pend_c = nullptr, *pend_alb = nullptr;
    int *pend_count = nullptr;
    uint32_t *shadow_queue = nullptr;
    uint32_t *spill = nullptr;
    size_t spill_cap = 0;
    float *halton = nullptr;
    unsigned long long *stats = nullptr;
    uint32_t *keys[2] = {nullptr, nullptr}, *vals[2] = {nullptr, nullptr};
    uint32_t *hist = nullptr;
    int passes_done = 0;
    int mix_work = 0; // tile counter of the mix (the last of a pass; the same for both slots' graphs)

    // options
    int lds_depth = 16;
    int path_tail = 2;
    int use_graph = 1;
    int blocks_per_cu = 8;
    int max_bounces = RH_MAX_BOUNCES; // depth clamp of scenes added from now on
    int sort_bits = 0;                 // staged bounces sorted by the top sort_bits of the ray key (0: no sort)
    long long queue_rays = 0;          // rays in flight at most (scenes grouped to fit; 0: every scene at once)

    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_key[5] = {-1, -1, -1, -1, -1};
};

namespace {

// Makes slot j the bound one: its working set into the batch's fields, the
// bound one's back to the store.
void bind_slot(rh_batch *b, int j) {
    if (j == b->bound) return;
#define RH_BIND_FIELD(f)                                                                                                \
    std::swap(b->f, b->store[j].f);                                                                                    \
    std::swap(b->store[j].f, b->store[b->bound].f);
    RH_BATCH_SLOT_FIELDS(RH_BIND_FIELD)
#undef RH_BIND_FIELD
    b->bound = j;
}

// Waits for every slot's stream.
int sync_all(rh_batch *b) {
    RH_CHECK(hipStreamSynchronize(b->stream));
    for (int j = 0; j < kMaxBatchSlots; j++)
        if (j != b->bound && b->store[j].stream) RH_CHECK(hipStreamSynchronize(b->store[j].stream));
    return 0;
}

void drop_graph(rh_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

// The bound slot's arrays and graph.
void free_slot_arrays(rh_batch *b) {
    drop_graph(b);
    void *ps[] = {b->table,    b->pass,    b->counts,  b->shadow_counts, b->work,       b->rays[0],   b->rays[1],
                  b->aux[0],   b->aux[1],  b->hits,    b->temp,          b->nee_ray,    b->nee_term,  b->nee_aux,
                  b->nee_vis,  b->pend_c,  b->pend_alb, b->pend_count,   b->shadow_queue, b->spill,   b->keys[0],
                  b->keys[1],  b->vals[0], b->vals[1], b->hist};
    for (void *p : ps) free_ptr(p);
    b->keys[0] = b->keys[1] = b->vals[0] = b->vals[1] = b->hist = nullptr;
    b->table = nullptr;
    b->pass = b->counts = b->shadow_counts = b->work = nullptr;
    b->rays[0] = b->rays[1] = nullptr;
    b->aux[0] = b->aux[1] = nullptr;
    b->hits = nullptr;
    b->temp = nullptr;
    b->nee_ray = b->nee_term = nullptr;
    b->nee_aux = nullptr;
    b->nee_vis = nullptr;
    b->pend_c = b->pend_alb = nullptr;
    b->pend_count = nullptr;
    b->shadow_queue = nullptr;
    b->spill = nullptr;
}

// Every slot's arrays and the shared framebuffers.
void free_arrays(rh_batch *b) {
    const int keep = b->bound;
    for (int j = 0; j < kMaxBatchSlots; j++) {
        bind_slot(b, j);
        free_slot_arrays(b);
    }
    bind_slot(b, keep);
    free_ptr(b->clean);
    free_ptr(b->final_);
    b->clean = b->final_ = nullptr;
    b->built = false;
}



template <typename T> int dev_alloc(T **p, size_t n) {
    RH_CHECK(hipMalloc(p, std::max<size_t>(n, 1) * sizeof(T)));
    return 0;
}

bool any_wide(const rh_batch *b) {
    for (const BatchEntry &e : b->entries)
        if (e.sc.wnodes) return true;
    return false;
}

bool batch_all_solid(const rh_batch *b) {
    for (const BatchEntry &e : b->entries)
        if (!e.sc.all_solid) return false;
    return true;
}

int batch_tree(const rh_batch *b) {
    int wide = 0;
    for (const BatchEntry &e : b->entries) wide += e.sc.wnodes ? 1 : 0;
    return wide == 0 ? kTreeBinary : (wide == (int)b->entries.size() ? kTreeWide : kTreeAny);
}

// The per-bounce launches of a pass: stage 0 is the primary hit; stage k >= 1
// traces bounce k - 1 (its queue first sorted per scene when sort_bits is
// set).  Bounces from path_tail on run in the tail kernel.
template <int Tree, bool Solid>
void launch_stages(rh_batch *b, const BatchDev &d, size_t lds, dim3 grid, dim3 blk, int tail, int &work) {
    int q = 0;
    for (int stage = 0; stage <= b->max_depth; stage++) {
        if (stage >= tail) {
            hipLaunchKernelGGL((k_b_tail<true, Tree, Solid>), grid, blk, lds, b->stream, d, stage, q, work++);
            hipLaunchKernelGGL((k_b_tail<false, Tree, false>), grid, blk, lds, b->stream, d, stage, q, work++);
            break;
        }
        if (stage >= 1 && b->sort_bits) {
            hipLaunchKernelGGL(k_b_sort_keys, grid, blk, 0, b->stream, d, stage, q, work++);
            int src = 0;
            for (int shift = 32 - b->sort_bits; shift < 32; shift += 8) {
                hipLaunchKernelGGL(k_b_sort_hist, grid, blk, 0, b->stream, d, stage, src, shift, work++);
                hipLaunchKernelGGL(k_b_sort_scan, dim3((unsigned)b->entries.size()), dim3(256), 0, b->stream, d, stage);
                hipLaunchKernelGGL(k_b_sort_scatter, grid, blk, 0, b->stream, d, stage, src, shift, work++);
                src ^= 1;
            }
            hipLaunchKernelGGL(k_b_sort_gather, grid, blk, 0, b->stream, d, stage, q, src, work++);
            q ^= 1; // the sorted rays
        }
        hipLaunchKernelGGL(k_b_trace<Tree>, grid, blk, lds, b->stream, d, stage, q, work++);
        hipLaunchKernelGGL(k_b_shade<true>, grid, blk, 0, b->stream, d, stage, q, work++);
        hipLaunchKernelGGL(k_b_shade<false>, grid, blk, lds, b->stream, d, stage, q, work++);
        hipLaunchKernelGGL((k_b_shadow<Tree, Solid>), grid, blk, lds, b->stream, d, stage, work++);
        hipLaunchKernelGGL(k_b_combine, grid, blk, 0, b->stream, d, stage, work++);
        q ^= 1;
    }
}

// Tile counters of the whole pass: kMaxStages per group, and the mix.
size_t work_counters(const rh_batch *b) { return b->groups.size() * (size_t)kMaxStages + 4; }

// Scenes in order into groups of at most queue_rays pixels (a larger scene
// alone), each scene's queue and NEE offsets within its group.
void group_scenes(rh_batch *b) {
    b->groups.clear();
    b->queue_pixels = b->queue_slots = 0;
    size_t qp = 0, qs = 0;
    int first = 0;
    for (int s = 0; s < (int)b->entries.size(); s++) {
        BatchEntry &e = b->entries[s];
        if (b->queue_rays > 0 && qp > 0 && qp + e.pixels > (size_t)b->queue_rays) {
            b->groups.push_back({first, s});
            first = s;
            qp = qs = 0;
        }
        e.q_offset = (uint32_t)qp;
        e.nee_offset = (uint32_t)qs;
        qp += e.pixels;
        qs += (size_t)e.pixels * e.nee_slots;
        b->queue_pixels = std::max(b->queue_pixels, qp);
        b->queue_slots = std::max(b->queue_slots, qs);
    }
    b->groups.push_back({first, (int)b->entries.size()});
}

// The bound slot's arrays: its table copy, counters, queues for the largest
// group and the pass framebuffer for every pixel of the batch.
int build_slot(rh_batch *b) {
    const size_t P = b->total_pixels, QP = b->queue_pixels, NS = b->queue_slots, S = b->entries.size();
    int rc = 0;
    rc |= dev_alloc(&b->table, S);
    rc |= dev_alloc(&b->pass, 1);
    rc |= dev_alloc(&b->counts, (RH_MAX_BOUNCES_EXT + 2) * S);
    rc |= dev_alloc(&b->shadow_counts, (RH_MAX_BOUNCES_EXT + 2) * S);
    rc |= dev_alloc(&b->work, work_counters(b));
    for (int k = 0; k < 2; k++) {
        rc |= dev_alloc(&b->rays[k], QP);
        rc |= dev_alloc(&b->aux[k], QP);
    }
    rc |= dev_alloc(&b->hits, QP);
    rc |= dev_alloc(&b->temp, P);
    rc |= dev_alloc(&b->nee_ray, NS);
    rc |= dev_alloc(&b->nee_term, NS);
    rc |= dev_alloc(&b->nee_aux, NS);
    rc |= dev_alloc(&b->nee_vis, NS);
    rc |= dev_alloc(&b->shadow_queue, NS);
    rc |= dev_alloc(&b->pend_c, QP);
    rc |= dev_alloc(&b->pend_alb, QP);
    rc |= dev_alloc(&b->pend_count, QP);
    // entries, and their distances when a scene has 8-wide trees
    rc |= dev_alloc(&b->spill, (size_t)kSpillDepth * (any_wide(b) ? 2 : 1) * b->spill_cap);
    return rc;
}

// Every scene back to pass 0: each slot's table copy and pass counter (with
// S slots, slot k takes passes k + 1, k + 1 + S, ...: its counter starts at
// k + 1 - S and steps by S), the accumulated images cleared.
int reset_passes(rh_batch *b) {
    int rc = sync_all(b);
    if (rc) return rc;
    const int slots = b->pass_overlap + 1;
    int init[kMaxBatchSlots];
    for (int k = 0; k < slots; k++) init[k] = k + 1 - slots;
    for (int k = 0; k < slots; k++) {
        bind_slot(b, k);
        RH_CHECK(hipMemcpyAsync(b->table, b->entries.data(), b->entries.size() * sizeof(BatchEntry),
                                hipMemcpyHostToDevice, b->stream));
        RH_CHECK(hipMemcpyAsync(b->pass, &init[k], sizeof(int), hipMemcpyHostToDevice, b->stream));
    }
    bind_slot(b, 0);
    const size_t n4 = b->total_pixels;
    hipLaunchKernelGGL(k_b_fill, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, b->stream, b->clean, n4,
                       make_float4(0, 0, 0, 0));
    hipLaunchKernelGGL(k_b_fill, dim3((unsigned)((n4 + 255) / 256)), dim3(256), 0, b->stream, b->final_, n4,
                       make_float4(0, 0, 0, 0));
    RH_CHECK(hipGetLastError());
    rc = sync_all(b); // init[] is read by the copies above
    if (rc) return rc;
    b->passes_done = 0;
    b->mix_recorded = false;
    return 0;
}

// Lays out the queues for the scenes added so far and allocates the shared
// framebuffers and each slot's arrays.
int build(rh_batch *b) {
    if (b->built) return 0;
    if (b->entries.empty()) return arg_error("batch has no scenes");
    int rc = sync_all(b);
    if (rc) return rc;
    free_arrays(b);
    group_scenes(b);
    const size_t P = b->total_pixels;
    b->spill_cap = std::max(b->queue_pixels, b->queue_slots);
    rc |= dev_alloc(&b->clean, P);
    rc |= dev_alloc(&b->final_, P);
    for (int k = 0; k <= b->pass_overlap; k++) {
        bind_slot(b, k);
        rc |= build_slot(b);
    }
    bind_slot(b, 0);
    if (rc) return rc;
    rc = reset_passes(b);
    if (rc) return rc;
    b->built = true;
    return 0;
}

BatchDev dev_view(rh_batch *b) {
    BatchDev d;
    memset(&d, 0, sizeof(d));
    d.table = b->table;
    d.n_scenes = (int)b->entries.size();
    d.s_begin = 0;
    d.s_end = d.n_scenes;
    d.pass = b->pass;
    d.counts = b->counts;
    d.shadow_counts = b->shadow_counts;
    d.work = b->work;
    d.rays[0] = b->rays[0];
    d.rays[1] = b->rays[1];
    d.aux[0] = b->aux[0];
    d.aux[1] = b->aux[1];
    d.hits = b->hits;
    d.temp = b->temp;
    d.clean = b->clean;
    d.final_ = b->final_;
    d.nee_ray = b->nee_ray;
    d.nee_term = b->nee_term;
    d.nee_aux = b->nee_aux;
    d.nee_vis = b->nee_vis;
    d.pend_c = b->pend_c;
    d.pend_alb = b->pend_alb;
    d.pend_count = b->pend_count;
    d.shadow_queue = b->shadow_queue;
    d.spill = b->spill;
    d.spill_stride = (uint32_t)b->spill_cap;
    d.lds_depth = (uint32_t)b->lds_depth;
    d.halton = b->halton;
    d.stats = b->stats;
    for (int k = 0; k < 2; k++) {
        d.keys[k] = b->keys[k];
        d.vals[k] = b->vals[k];
    }
    d.hist = b->hist;
    return d;
}

// The sort's buffers (once a pass sorts): keys and indices for every ray slot,
// digit counts for the most tiles the scenes' queues can make.
int ensure_sort_buffers(rh_batch *b) {
    if (!b->sort_bits || b->hist) return 0;
    const size_t P = b->queue_pixels;
    const size_t tiles = P / kSortTile + b->entries.size() + 1;
    int rc = 0;
    for (int k = 0; k < 2; k++) {
        rc |= dev_alloc(&b->keys[k], P);
        rc |= dev_alloc(&b->vals[k], P);
    }
    rc |= dev_alloc(&b->hist, tiles * 256);
    if (rc) {
        for (int k = 0; k < 2; k++) {
            rh_rt::release(b->keys[k]);
            rh_rt::release(b->vals[k]);
        }
        rh_rt::release(b->hist);
    }
    return rc;
}

// The launch sequence of one pass (fixed for given options): begin, generate,
// then per stage trace / shade / shadow / combine, the path tail.  The mix
// follows outside (rh_batch_render_pass) at tile counter *mix_work.
int enqueue_pass(rh_batch *b, int stride, int *mix_work) {
    // first stage run by the tail kernel (bounce path_tail)
    const int tail = b->path_tail < 0 ? b->max_depth + 1 : std::min(b->path_tail + 1, b->max_depth + 1);
    const int sort_launches = b->sort_bits ? 2 + 2 * (b->sort_bits / 8) : 0;
    // tile counters of one group: generation, the staged bounces (with their sorts), the tails
    const int launches = 1 + 5 * tail + sort_launches * std::max(0, tail - 1) + (tail <= b->max_depth ? 2 : 0);
    if (launches > kMaxStages) return arg_error("batch pass needs more stage counters than reserved");
    const BatchDev all = dev_view(b);
    const size_t lds = (size_t)b->lds_depth * kBlock * sizeof(uint32_t) * (any_wide(b) ? 2 : 1);
    const dim3 grid((unsigned)(b->cus * b->blocks_per_cu)), blk(kBlock);
    int work = 0;
    hipLaunchKernelGGL(k_b_begin_pass, dim3(1), dim3(1024), 0, b->stream, all, b->table, (int)work_counters(b), stride);
    const int tree = batch_tree(b);
    const bool solid = batch_all_solid(b);
    // the groups one after another: each fills the queues with its scenes' rays
    for (const auto &g : b->groups) {
        BatchDev d = all;
        d.s_begin = g.first;
        d.s_end = g.second;
        hipLaunchKernelGGL(k_b_gen, grid, blk, 0, b->stream, d, work++);
        if (tree == kTreeWide && solid) launch_stages<kTreeWide, true>(b, d, lds, grid, blk, tail, work);
        else if (tree == kTreeBinary) launch_stages<kTreeBinary, false>(b, d, lds, grid, blk, tail, work);
        else if (tree == kTreeWide) launch_stages<kTreeWide, false>(b, d, lds, grid, blk, tail, work);
        else launch_stages<kTreeAny, false>(b, d, lds, grid, blk, tail, work);
    }
    *mix_work = work++;
    RH_CHECK(hipGetLastError());
    return 0;
}

} // namespace

extern "C" {

int rh_batch_create(int device, rh_batch **out) {
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        rh_rt::err() = "no HIP device available";
        return -1002;
    }
    if (device < 0) RH_CHECK(hipGetDevice(&device));
    if (device >= ndev) return arg_error("device index out of range");
    RH_CHECK(hipSetDevice(device));
    auto *b = new rh_batch;
    b->device = device;
    // one stream per pass slot, from the device's shared pool
    static_assert(kMaxBatchSlots <= rh_rt::kPassStreams, "a stream per pass slot");
    hipStream_t streams[rh_rt::kPassStreams];
    const int src = rh_rt::pass_streams(device, streams);
    if (src) {
        delete b;
        return src;
    }
    b->stream = streams[0];
    for (int j = 1; j < kMaxBatchSlots; j++) b->store[j].stream = streams[j];
    RH_CHECK(hipEventCreateWithFlags(&b->ev_mix, hipEventDisableTiming));
    hipDeviceProp_t prop;
    RH_CHECK(hipGetDeviceProperties(&prop, device));
    b->cus = prop.multiProcessorCount;
    RH_CHECK(hipMalloc(&b->halton, RH_HALTON_TABLE_FLOATS * sizeof(float)));
    RH_CHECK(hipMemsetAsync(b->halton, 0, RH_HALTON_TABLE_FLOATS * sizeof(float), b->stream));
    RH_CHECK(hipMalloc(&b->stats, 8 * sizeof(unsigned long long)));
    RH_CHECK(hipMemsetAsync(b->stats, 0, 8 * sizeof(unsigned long long), b->stream));
    RH_CHECK(hipStreamSynchronize(b->stream));
    *out = b;
    return 0;
}

void rh_batch_destroy(rh_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)sync_all(b);
    free_arrays(b);
    for (auto &v : b->scene_bufs)
        for (auto &x : v) free_ptr(x.p);
    free_ptr(b->halton);
    free_ptr(b->stats);
    if (b->ev_mix) (void)hipEventDestroy(b->ev_mix);
    delete b; // the streams stay in the device's pool
}

int rh_batch_add_scene(rh_batch *b, const rh_scene_view *v, const rh_pass_desc *desc, int w, int h, int spp) {
    if ((int)b->entries.size() >= kMaxBatchScenes) return arg_error("batch is full (256 scenes)");
    if (w <= 0 || h <= 0 || w > 65535 || h > 65535) return arg_error("framebuffer size must be 1..65535 per side");
    if (spp < 1) return arg_error("spp must be >= 1");
    if (desc->rect[0] != 0 || desc->rect[1] != 0 || desc->rect[2] != w || desc->rect[3] != h)
        return arg_error("batched scenes render their whole frame");
    if (b->total_pixels + (size_t)w * h > 0x7fffffffu) return arg_error("batch exceeds 2^31 pixels");
    RH_CHECK(hipSetDevice(b->device));
    int rc0 = sync_all(b);
    if (rc0) return rc0;
    std::vector<DevBuf> bufs;
    BatchEntry e;
    memset(&e, 0, sizeof(e));
    const int rc = build_device_scene(v, b->stream, bufs, e.sc);
    if (rc) {
        for (auto &x : bufs) free_ptr(x.p);
        return rc;
    }
    RH_CHECK(hipStreamSynchronize(b->stream));
    rh_pass_desc p = *desc;
    p.iteration = 1;
    e.pc = make_pass_const(w, h, &p);
    e.pi.iteration = 1;
    e.pi.bounce = 2;
    e.pi.settings = desc->cam.pass_settings;
    e.depth = std::min<int>(desc->cam.pass_settings.max_total_depth, b->max_bounces);
    e.pi.settings.max_total_depth = (uint8_t)e.depth;
    e.pc.cam.pass_settings.max_total_depth = (uint8_t)e.depth;
    e.offset = (uint32_t)b->total_pixels;
    e.pixels = (uint32_t)((size_t)w * h);
    e.nee_slots = (v->lights_count >= 1 && v->lights_count <= kMaxDeferredLights) ? v->lights_count : 0;
    e.nee_offset = (uint32_t)b->total_slots;
    e.spp = spp;
    b->total_pixels += e.pixels;
    b->total_slots += (size_t)e.pixels * e.nee_slots;
    b->max_depth = std::max(b->max_depth, e.depth);
    b->max_spp = std::max(b->max_spp, spp);
    b->entries.push_back(e);
    b->scene_bufs.push_back(std::move(bufs));
    b->built = false;
    return (int)b->entries.size() - 1;
}

int rh_batch_set_halton(rh_batch *b, const float *seq, size_t count) {
    if (count < (size_t)RH_HALTON_COUNT * RH_HALTON_SEQ_LEN) return arg_error("Halton table needs 256 x 24 floats");
    const int rc = sync_all(b); // passes in flight read the table
    if (rc) return rc;
    RH_CHECK(upload_halton(b->halton, seq, b->stream));
    RH_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int rh_batch_passes(const rh_batch *b) { return b->max_spp; }

int rh_batch_passes_done(const rh_batch *b) { return b->built ? b->passes_done : 0; }

int rh_batch_render_pass(rh_batch *b) {
    RH_CHECK(hipSetDevice(b->device));
    int rc = build(b);
    if (rc) return rc;
    if (b->passes_done >= b->max_spp) return arg_error("every scene of the batch has rendered its passes");
    // with S slots, pass p (from 1) on slot (p - 1) % S; its stages run while
    // the other slots' passes finish
    const int stride = b->pass_overlap + 1;
    bind_slot(b, b->passes_done % stride);
    rc = ensure_sort_buffers(b);
    if (rc) return rc;
    const int key[5] = {b->lds_depth, b->path_tail, b->blocks_per_cu, b->sort_bits, stride};
    if (b->use_graph) {
        if (!b->graph_exec || memcmp(key, b->graph_key, sizeof(key)) != 0) {
            drop_graph(b);
            RH_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            rc = enqueue_pass(b, stride, &b->mix_work);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return fail(e, "hipStreamEndCapture");
            b->graph = g;
            RH_CHECK(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
            memcpy(b->graph_key, key, sizeof(key));
        }
        RH_CHECK(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
        rc = enqueue_pass(b, stride, &b->mix_work);
        if (rc) return rc;
    }
    // passes mix into the accumulated images in pass order
    if (b->mix_recorded) RH_CHECK(hipStreamWaitEvent(b->stream, b->ev_mix, 0));
    hipLaunchKernelGGL(k_b_mix, dim3((unsigned)(b->cus * b->blocks_per_cu)), dim3(kBlock), 0, b->stream, dev_view(b),
                       b->mix_work);
    RH_CHECK(hipGetLastError());
    RH_CHECK(hipEventRecord(b->ev_mix, b->stream));
    b->mix_recorded = true;
    b->passes_done++;
    return 0;
}

int rh_batch_restart(rh_batch *b) {
    RH_CHECK(hipSetDevice(b->device));
    if (!b->built) return 0;
    return reset_passes(b);
}

int rh_batch_synchronize(rh_batch *b) { return sync_all(b); }

int rh_batch_read_pixels(rh_batch *b, int scene, float *rgba, size_t count) {
    if (scene < 0 || scene >= (int)b->entries.size()) return arg_error("scene index out of range");
    if (!b->built) return arg_error("batch has not rendered");
    const BatchEntry &e = b->entries[scene];
    if (count < (size_t)e.pixels * 4) return arg_error("output buffer too small");
    const int rc = sync_all(b);
    if (rc) return rc;
    RH_CHECK(hipMemcpyAsync(rgba, b->final_ + e.offset, (size_t)e.pixels * sizeof(float4), hipMemcpyDefault,
                            b->stream));
    RH_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int rh_batch_read_queue(rh_batch *b, int scene, int stage, int buffer, rh_ray *out, size_t cap, int *count) {
    if (scene < 0 || scene >= (int)b->entries.size()) return arg_error("scene index out of range");
    if (stage < 0 || stage > RH_MAX_BOUNCES_EXT + 1 || buffer < 0 || buffer > 1)
        return arg_error("stage or buffer out of range");
    if (!b->built) return arg_error("batch has not rendered");
    // the queues of the last pass (the bound slot's)
    const int rc = sync_all(b);
    if (rc) return rc;
    int n = 0;
    RH_CHECK(hipMemcpy(&n, b->counts + (size_t)stage * b->entries.size() + scene, sizeof(int), hipMemcpyDeviceToHost));
    *count = n;
    const size_t k = std::min<size_t>((size_t)std::max(n, 0), cap);
    if (k) RH_CHECK(hipMemcpy(out, b->rays[buffer] + b->entries[scene].q_offset, k * sizeof(rh_ray), hipMemcpyDeviceToHost));
    return 0;
}

int rh_batch_reset_stats(rh_batch *b) {
    const int rc = sync_all(b);
    if (rc) return rc;
    RH_CHECK(hipMemsetAsync(b->stats, 0, 8 * sizeof(unsigned long long), b->stream));
    RH_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int rh_batch_get_ray_counts(rh_batch *b, rh_ray_counts *out) {
    const int rc = sync_all(b);
    if (rc) return rc;
    unsigned long long v[8];
    RH_CHECK(hipMemcpyAsync(v, b->stats, sizeof(v), hipMemcpyDeviceToHost, b->stream));
    RH_CHECK(hipStreamSynchronize(b->stream));
    memset(out, 0, sizeof(*out));
    out->primary = v[0];
    out->secondary = v[1];
    out->shadow = v[2];
    out->overflow = v[4];
    return 0;
}

int rh_batch_set_option(rh_batch *b, const char *name, int value) {
    const std::string n = name;
    if (n == "lds_stack_depth") {
        if (value < 4 || value > 64) return arg_error("lds_stack_depth must be 4..64");
        b->lds_depth = value;
    } else if (n == "path_tail") {
        if (value < -1 || value > 255) return arg_error("path_tail must be -1..255");
        b->path_tail = value;
    } else if (n == "use_graph") {
        b->use_graph = value != 0;
    } else if (n == "blocks_per_cu") {
        if (value < 1 || value > 64) return arg_error("blocks_per_cu must be 1..64");
        b->blocks_per_cu = value;
    } else if (n == "max_bounces") {
        if (value < 1 || value > RH_MAX_BOUNCES_EXT) return arg_error("max_bounces must be 1..16");
        b->max_bounces = value;
    } else if (n == "queue_rays") {
        if (value < 0) return arg_error("queue_rays must be >= 0 (0: no limit)");
        if (value != b->queue_rays) {
            b->queue_rays = value;
            b->built = false; // queues and groups are laid out again at the next pass (which restarts the batch)
        }
    } else if (n == "sort_bits") {
        if (value != 0 && value != 8 && value != 16 && value != 24 && value != 32)
            return arg_error("sort_bits must be 0, 8, 16, 24 or 32");
        b->sort_bits = value;
    } else if (n == "pass_overlap") {
        if (value < 0 || value >= kMaxBatchSlots) return arg_error("pass_overlap must be 0..3");
        if (value != b->pass_overlap) {
            const int rc = sync_all(b);
            if (rc) return rc;
            b->pass_overlap = value;
            b->built = false; // the slots are laid out again at the next pass (which restarts the batch)
        }
    } else {
        return arg_error("unknown option");
    }
    return 0;
}

} // extern "C"
