// Device-side building blocks of the HIP backend (gfx950).
//
// Every routine here reproduces the arithmetic of the Ray library's CPU Ref
// backend (src/Ray/internal/CoreRef.cpp) operation for operation: the file is
// compiled with -ffp-contract=off, so each float multiply and add rounds
// separately in source order, divisions and square roots are IEEE correctly
// rounded, and sums over vector components run left to right as the
// reference's generic simd_vec does (simd_vec.h:319-327, 559-563).  Cosine,
// sine, power and arc-cosine are evaluated in double precision and rounded to
// float, which agrees with the host C library's float functions except in
// rare last-bit cases.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "rh_types.h"

namespace rh {

// ---------------------------------------------------------------- vectors
struct f3 {
    float x, y, z;
};
struct f2 {
    float x, y;
};
struct f4 {
    float x, y, z, w;
};

__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 ld3(const float *p) { return f3{p[0], p[1], p[2]}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(f3 a, f3 b) {
    float r = 0.0f;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
__device__ __forceinline__ float length(f3 a) { return __builtin_sqrtf(dot(a, a)); }
__device__ __forceinline__ f3 normalize(f3 a) { return a / length(a); }
__device__ __forceinline__ f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

// std::min / std::max argument order
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
// std::modf fractional part for the finite values used here
__device__ __forceinline__ float fract_modf(float x) { return x - __builtin_truncf(x); }
// simd_vec floor (simd_vec.h:408-412): int truncation minus one for negatives
__device__ __forceinline__ float simd_floor(float v) { return (float)((int)v - (v < 0.0f)); }

__device__ __forceinline__ float rcos(float x) { return (float)cos((double)x); }
__device__ __forceinline__ float rsin(float x) { return (float)sin((double)x); }
__device__ __forceinline__ float racos(float x) { return (float)acos((double)x); }
__device__ __forceinline__ float rpow(float x, float y) { return (float)pow((double)x, (double)y); }

// hash (CoreRef.cpp:264-270)
__device__ __forceinline__ int hash(int x) {
    unsigned ret = (unsigned)x;
    ret = ((ret >> 16) ^ ret) * 0x45d9f3bu;
    ret = ((ret >> 16) ^ ret) * 0x45d9f3bu;
    ret = (ret >> 16) ^ ret;
    return (int)ret;
}
// construct_float (CoreRef.cpp:359-368)
__device__ __forceinline__ float construct_float(uint32_t m) {
    m &= 0x007FFFFFu;
    m |= 0x3F800000u;
    return __uint_as_float(m) - 1.0f;
}
// to_norm_float (Core.h:249-252)
__device__ __forceinline__ float to_norm_float(uint32_t v) {
    const uint32_t val = 0x3f800000u + v * 0x8080u + (v + 1u) / 2u;
    return __uint_as_float(val) - 1.0f;
}

// ---------------------------------------------------------------- scene
struct Scene {
    const rh_bvh_node *nodes;
    const rh_tri_accel *tris;
    const uint32_t *tri_indices;
    // tris in tri_indices order (leaf_tris[k] = tris[tri_indices[k]], with
    // the material field mi holding the triangle index tri_indices[k] instead:
    // the leaf test needs no material): one load round per triangle test
    const rh_tri_accel *leaf_tris;
    const rh_mesh *meshes;
    const rh_mesh_instance *mesh_instances;
    // one record of 5 float4 per mi_indices entry, the instance that entry
    // names with its tree root and inverse transform resolved (load_instance)
    const float4 *inst_recs;
    const uint32_t *mi_indices;
    const rh_transform *transforms;
    const rh_vertex *vertices;
    const uint32_t *vtx_indices;
    const rh_material *materials;
    const rh_texture *textures;
    const rh_tex_mip *tex_mips;
    const uint32_t *texels;
    const rh_light *lights;
    const uint32_t *li_indices;
    const rh_bvh_node *light_leaves; // light-tree leaves in depth-first visit order
    uint32_t light_leaf_count;
    rh_environment env;
    uint32_t macro_root;
    // 8-wide scenes: the macro tree's root as wide_root reads it (a leaf
    // root's range without the load)
    uint32_t macro_word, macro_count;
    // 8-wide scenes whose macro root is a leaf of one instance (a scene of one
    // mesh instance): that instance's record (inst_recs), kept here so that
    // every ray tests it without reading memory
    uint32_t macro_single;
    float4 macro_inst[5];
    // and when that instance's tree root is an interior node: the root's
    // device node (its first step of every ray reads no memory)
    uint32_t root_cached;
    float4 root_node[16];
    uint32_t light_root;
    // 8-wide trees (see the wide traversal below); null for binary scenes
    const float4 *wnodes;
    // leaf children's words carry small prim counts (kLeafCountBits) and are
    // stacked as they are (a device tree of fewer than 2^24 nodes)
    uint32_t wide_leaf_counts;
    // the wide light tree is a single leaf: its lights are taken for every
    // point, with no bounds test (the wide walk tests only child slots)
    uint32_t light_untested;
    // every triangle is solid (no transparent material reachable): a
    // visibility query then only asks whether any triangle lies in between
    uint32_t all_solid;
};

struct Hit {
    int mask, obj, prim;
    float t, u, v;
};

// ---------------------------------------------------------------- traversal
// bbox_test (CoreRef.cpp:64-86): slab test, tmax widened by 1.00000024
__device__ __forceinline__ bool bbox_test(f3 o, f3 inv_d, float t, const float *bmin, const float *bmax) {
    float lo_x = inv_d.x * (bmin[0] - o.x);
    float hi_x = inv_d.x * (bmax[0] - o.x);
    if (lo_x > hi_x) {
        const float tmp = lo_x;
        lo_x = hi_x;
        hi_x = tmp;
    }
    float lo_y = inv_d.y * (bmin[1] - o.y);
    float hi_y = inv_d.y * (bmax[1] - o.y);
    if (lo_y > hi_y) {
        const float tmp = lo_y;
        lo_y = hi_y;
        hi_y = tmp;
    }
    float lo_z = inv_d.z * (bmin[2] - o.z);
    float hi_z = inv_d.z * (bmax[2] - o.z);
    if (lo_z > hi_z) {
        const float tmp = lo_z;
        lo_z = hi_z;
        hi_z = tmp;
    }
    float tmin = lo_x > lo_y ? lo_x : lo_y;
    if (lo_z > tmin) tmin = lo_z;
    float tmax = hi_x < hi_y ? hi_x : hi_y;
    if (hi_z < tmax) tmax = hi_z;
    tmax *= 1.00000024f;
    return tmin <= tmax && tmin <= t && tmax > 0;
}

// The same slab test for rays whose inv_d has no infinite component: no slab
// value can then be NaN (finite * finite), and for non-NaN operands the
// reference's swaps and selects are exactly min and max (only the sign of a
// zero may differ, which no comparison sees), so hardware min/max (min3/max3)
// give the same verdict with fewer instructions.
__device__ __forceinline__ bool bbox_test_finite(f3 o, f3 inv_d, float t, const float *bmin, const float *bmax) {
    const float lx = inv_d.x * (bmin[0] - o.x), hx = inv_d.x * (bmax[0] - o.x);
    const float ly = inv_d.y * (bmin[1] - o.y), hy = inv_d.y * (bmax[1] - o.y);
    const float lz = inv_d.z * (bmin[2] - o.z), hz = inv_d.z * (bmax[2] - o.z);
    const float tmin = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(lx, hx), __builtin_fminf(ly, hy)),
                                       __builtin_fminf(lz, hz));
    const float tmax = __builtin_fminf(__builtin_fminf(__builtin_fmaxf(lx, hx), __builtin_fmaxf(ly, hy)),
                                       __builtin_fmaxf(lz, hz)) *
                       1.00000024f;
    return tmin <= tmax && tmin <= t && tmax > 0;
}

template <bool Finite>
__device__ __forceinline__ bool box_hit(f3 o, f3 inv_d, float t, const float *bmin, const float *bmax) {
    return Finite ? bbox_test_finite(o, inv_d, t, bmin, bmax) : bbox_test(o, inv_d, t, bmin, bmax);
}

// safe_invert (CoreRef.cpp:272-294)
__device__ __forceinline__ f3 safe_invert(f3 v) {
    f3 o = f3{1.0f / v.x, 1.0f / v.y, 1.0f / v.z};
    const float big = 3.402823466e+38f;
    if (v.x <= RH_FLT_EPS && v.x >= 0) o.x = big;
    else if (v.x >= -RH_FLT_EPS && v.x < 0) o.x = -big;
    if (v.y <= RH_FLT_EPS && v.y >= 0) o.y = big;
    else if (v.y >= -RH_FLT_EPS && v.y < 0) o.y = -big;
    if (v.z <= RH_FLT_EPS && v.z >= 0) o.z = big;
    else if (v.z >= -RH_FLT_EPS && v.z < 0) o.z = -big;
    return o;
}

// _IntersectTri (CoreRef.cpp:11-42): projected-plane test with HIT_EPS edge tolerance
__device__ __forceinline__ bool intersect_tri(f3 o, f3 d, const rh_tri_accel &tri, float cur_t, float &t_out,
                                              float &u_out, float &v_out) {
    const int iw = tri.ci & RH_TRI_W_BITS;
    const int iu = iw == 0 ? 1 : 0;
    const int iv = iw == 2 ? 1 : 2;
    const float dw = comp(d, iw), du = comp(d, iu), dv = comp(d, iv);
    const float ow = comp(o, iw), ou = comp(o, iu), ov = comp(o, iv);

    const float det = du * tri.nu + dv * tri.nv + dw;
    const float dett = tri.np - (ou * tri.nu + ov * tri.nv + ow);
    const float Du = du * dett - (tri.pu - ou) * det;
    const float Dv = dv * dett - (tri.pv - ov) * det;
    const float detu = tri.e1v * Du - tri.e1u * Dv;
    const float detv = tri.e0u * Dv - tri.e0v * Du;

    const float tmpdet0 = det - detu - detv;
    if ((tmpdet0 > -RH_HIT_EPS && detu > -RH_HIT_EPS && detv > -RH_HIT_EPS) ||
        (tmpdet0 < RH_HIT_EPS && detu < RH_HIT_EPS && detv < RH_HIT_EPS)) {
        const float rdet = 1.0f / det;
        const float t = dett * rdet;
        if (t > 0 && t < cur_t) {
            t_out = t;
            u_out = detu * rdet;
            v_out = detv * rdet;
            return true;
        }
    }
    return false;
}

__device__ __forceinline__ rh_tri_accel load_tri(const rh_tri_accel *tris, uint32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(tris + i);
    const float4 a = p[0], b = p[1], c = p[2];
    rh_tri_accel t;
    t.nu = a.x;
    t.nv = a.y;
    t.np = a.z;
    t.pu = a.w;
    t.pv = b.x;
    t.ci = __float_as_int(b.y);
    t.e0u = b.z;
    t.e0v = b.w;
    t.e1u = c.x;
    t.e1v = c.y;
    t.mi = __float_as_uint(c.z);
    t.back_mi = __float_as_uint(c.w);
    return t;
}

struct NodeBox {
    float bmin[3], bmax[3];
    uint32_t a, b; // prim_index/left_child, prim_count/right_child
};
__device__ __forceinline__ NodeBox load_node(const rh_bvh_node *nodes, uint32_t i) {
    const float4 *p = reinterpret_cast<const float4 *>(nodes + i);
    const float4 lo = p[0], hi = p[1];
    NodeBox n;
    n.bmin[0] = lo.x;
    n.bmin[1] = lo.y;
    n.bmin[2] = lo.z;
    n.a = __float_as_uint(lo.w);
    n.bmax[0] = hi.x;
    n.bmax[1] = hi.y;
    n.bmax[2] = hi.z;
    n.b = __float_as_uint(hi.w);
    return n;
}

// Instance record k (the instance mi_indices[k] names): r0 = (bbox_min,
// tree root), r1 = (bbox_max, instance index), r2..r4 = the rows of its
// inverse transform's 3x4 part, so that the box test and TransformRay need
// one round of independent loads instead of three dependent ones.
struct InstanceRec {
    float bmin[3], bmax[3];
    uint32_t root, mi_index;
    float4 x0, x1, x2;
};
__device__ __forceinline__ InstanceRec instance_rec(const float4 *p) {
    const float4 a = p[0], b = p[1];
    InstanceRec r;
    r.bmin[0] = a.x;
    r.bmin[1] = a.y;
    r.bmin[2] = a.z;
    r.root = __float_as_uint(a.w);
    r.bmax[0] = b.x;
    r.bmax[1] = b.y;
    r.bmax[2] = b.z;
    r.mi_index = __float_as_uint(b.w);
    r.x0 = p[2];
    r.x1 = p[3];
    r.x2 = p[4];
    return r;
}
__device__ __forceinline__ InstanceRec load_instance(const Scene &sc, uint32_t k) {
    return instance_rec(sc.inst_recs + (size_t)k * 5);
}
// TransformRay (CoreRef.cpp:1605-1617), the same products and sums in the same order
__device__ __forceinline__ void transform_ray(const InstanceRec &r, f3 o, f3 d, f3 &mo, f3 &md) {
    mo = f3{r.x0.x * o.x + r.x0.y * o.y + r.x0.z * o.z + r.x0.w, r.x1.x * o.x + r.x1.y * o.y + r.x1.z * o.z + r.x1.w,
            r.x2.x * o.x + r.x2.y * o.y + r.x2.z * o.z + r.x2.w};
    md = f3{r.x0.x * d.x + r.x0.y * d.y + r.x0.z * d.z, r.x1.x * d.x + r.x1.y * d.y + r.x1.z * d.z,
            r.x2.x * d.x + r.x2.y * d.y + r.x2.z * d.z};
}

// Per-lane traversal stack: the first LDS_DEPTH entries live in LDS (lane-minor,
// so the 64 lanes of a wave hit distinct banks), deeper entries spill to a
// per-thread slice of global memory; pushes past LDS_DEPTH + SPILL_DEPTH are
// dropped and counted (the reference's fixed 32-entry stacks, CoreRef.cpp:1099-1111,
// have no overflow check at all).
struct Stack {
    uint32_t *lds;      // base of this block's LDS stack, [depth][blockDim]
    uint32_t *spill;    // base of this thread's spill slice, stride = spill_stride
    uint32_t spill_stride;
    uint32_t lds_depth;
    uint32_t spill_depth;
    uint32_t lane;
    uint32_t block;
    uint32_t sp;
    uint32_t overflow;
    // the 8-wide walk keeps an entry distance beside each entry: the LDS
    // stack's second half and the spill slice's second half
    float *lds_d;
    float *spill_d;

    __device__ __forceinline__ void init(uint32_t *lds_base, uint32_t *spill_base, uint32_t stride, uint32_t depth,
                                         uint32_t spill_entries) {
        lds = lds_base;
        spill = spill_base;
        spill_stride = stride;
        lds_depth = depth;
        spill_depth = spill_entries;
        lane = threadIdx.x;
        block = blockDim.x;
        sp = 0;
        overflow = 0;
        lds_d = reinterpret_cast<float *>(lds_base + (size_t)depth * blockDim.x);
        spill_d = reinterpret_cast<float *>(spill_base + (size_t)spill_entries * stride);
    }
    // the stack of ray `ray` of the `rays` a block walks with lane groups
    // (every lane of a group uses the same entries)
    __device__ __forceinline__ void init_ray(uint32_t *lds_base, uint32_t *spill_base, uint32_t stride, uint32_t depth,
                                             uint32_t spill_entries, uint32_t ray, uint32_t rays) {
        init(lds_base, spill_base, stride, depth, spill_entries);
        lane = ray;
        block = rays;
        lds_d = reinterpret_cast<float *>(lds_base + (size_t)depth * rays);
    }
    // entry k (any position below capacity) with its distance
    __device__ __forceinline__ void put(uint32_t k, uint32_t v, float d) {
        if (k < lds_depth) {
            lds[k * block + lane] = v;
            lds_d[k * block + lane] = d;
        } else if (k < lds_depth + spill_depth) {
            spill[(size_t)(k - lds_depth) * spill_stride] = v;
            spill_d[(size_t)(k - lds_depth) * spill_stride] = d;
        } else {
            overflow++;
        }
    }
    __device__ __forceinline__ void push2(uint32_t v, float d) {
        if (sp < lds_depth + spill_depth) {
            put(sp, v, d);
            sp++;
        } else {
            overflow++;
        }
    }
    __device__ __forceinline__ void pop2(uint32_t &v, float &d) {
        --sp;
        if (sp < lds_depth) {
            v = lds[sp * block + lane];
            d = lds_d[sp * block + lane];
        } else {
            v = spill[(size_t)(sp - lds_depth) * spill_stride];
            d = spill_d[(size_t)(sp - lds_depth) * spill_stride];
        }
    }

    __device__ __forceinline__ void push(uint32_t v) {
        if (sp < lds_depth) {
            lds[sp * block + lane] = v;
        } else if (sp < lds_depth + spill_depth) {
            spill[(size_t)(sp - lds_depth) * spill_stride] = v;
        } else {
            overflow++;
            return;
        }
        sp++;
    }
    __device__ __forceinline__ uint32_t pop() {
        --sp;
        if (sp < lds_depth) return lds[sp * block + lane];
        return spill[(size_t)(sp - lds_depth) * spill_stride];
    }
};

struct TraceCounters {
    uint32_t nodes, leaf_tris, instances;
};

// Leaf triangle loop (IntersectTris_ClosestHit / _AnyHit, CoreRef.cpp:760-825).
// Returns true when an any-hit search may stop (a solid triangle was hit).
template <bool AnyHit>
__device__ __forceinline__ bool intersect_leaf(const Scene &sc, const NodeBox &n, f3 o, f3 d, int obj, Hit &hit,
                                               TraceCounters *cnt) {
    const uint32_t first = n.a & RH_PRIM_INDEX_BITS, count = n.b;
    bool leaf_hit = false;
    int hit_ci = 0; // flags of the leaf's latest hit triangle
    for (uint32_t k = first; k < first + count; k++) {
        const rh_tri_accel tri = load_tri(sc.leaf_tris, k);
        const uint32_t ti = tri.mi; // the triangle's index (leaf_tris)
        if (cnt) cnt->leaf_tris++;
        float t, u, v;
        if (intersect_tri(o, d, tri, hit.t, t, u, v)) {
            hit.mask = -1;
            hit.prim = (int)ti;
            hit.obj = obj;
            hit.t = t;
            hit.u = u;
            hit.v = v;
            leaf_hit = true;
            hit_ci = tri.ci;
        }
        if (AnyHit && leaf_hit && (tri.ci & RH_TRI_SOLID_BIT)) break;
    }
    return AnyHit && leaf_hit && (hit_ci & RH_TRI_SOLID_BIT);
}

// Interior-node step: both children's boxes are tested here (siblings are
// adjacent in the node array, so this is one 64-byte pair load), the lane
// descends into the near child and pushes the far one only when both boxes
// pass.  This visits exactly the nodes the reference visits, in its order:
// the reference pushes far then near and box-tests each node when popped; the
// near child is popped at once (same t), and a pushed far child is re-tested
// against the current t when popped here too (boxes only get culled as t
// shrinks).  Returns false when neither child is entered.
template <bool Finite>
__device__ __forceinline__ bool enter_children(const Scene &sc, const NodeBox &n, f3 o, f3 d, f3 inv_d, float t,
                                               Stack &st, NodeBox &next, TraceCounters *cnt) {
    const uint32_t axis = n.b >> 30;
    const uint32_t right = n.b & RH_RIGHT_CHILD_BITS, left = n.a;
    const bool neg = comp(d, axis) < 0;
    const uint32_t near_i = neg ? right : left, far_i = neg ? left : right;
    const NodeBox nb = load_node(sc.nodes, near_i), fb = load_node(sc.nodes, far_i);
    if (cnt) cnt->nodes += 2;
    const bool hn = box_hit<Finite>(o, inv_d, t, nb.bmin, nb.bmax);
    const bool hf = box_hit<Finite>(o, inv_d, t, fb.bmin, fb.bmax);
    if (hn) {
        if (hf) st.push(far_i);
        next = nb;
        return true;
    }
    if (hf) {
        next = fb;
        return true;
    }
    return false;
}

// Pops until a node passes its box test against the current t.  A popped node
// was counted when its parent's children were tested: the reference visits
// (box-tests) each child once, at its pop, and so does this count.
template <bool Finite>
__device__ __forceinline__ bool pop_entered(const Scene &sc, uint32_t base, f3 o, f3 inv_d, float t, Stack &st,
                                            NodeBox &next, TraceCounters *) {
    while (st.sp > base) {
        const NodeBox n = load_node(sc.nodes, st.pop());
        if (box_hit<Finite>(o, inv_d, t, n.bmin, n.bmax)) {
            next = n;
            return true;
        }
    }
    return false;
}

// One instance's triangle tree, from its root node mc.  Returns true when an
// any-hit query found a solid occluder.
template <bool AnyHit, bool Finite>
__device__ __forceinline__ bool walk_micro(const Scene &sc, NodeBox mc, f3 mo, f3 md, f3 minv, int mi_index, Hit &hit,
                                           Stack &st, TraceCounters *cnt) {
    const uint32_t base = st.sp;
    bool mhave = box_hit<Finite>(mo, minv, hit.t, mc.bmin, mc.bmax);
    while (true) {
        if (mhave) {
            if (!(mc.a & RH_LEAF_NODE_BIT)) {
                mhave = enter_children<Finite>(sc, mc, mo, md, minv, hit.t, st, mc, cnt);
                continue;
            }
            if (intersect_leaf<AnyHit>(sc, mc, mo, md, mi_index, hit, cnt)) return true;
        }
        mhave = pop_entered<Finite>(sc, base, mo, minv, hit.t, st, mc, cnt);
        if (!mhave) return false;
    }
}

// Two-level traversal, entered-node form (see enter_children): the lane holds
// the node it is in; the stack holds far siblings still to be re-tested.  The
// macro tree uses safe_invert's finite inverse; an instance's object-space
// ray takes the exact NaN-preserving box test only when 1/d overflowed.
template <bool AnyHit>
__device__ __forceinline__ void traverse_fast(const Scene &sc, f3 o, f3 d, Hit &hit, Stack &st, TraceCounters *cnt) {
    const f3 inv_d = safe_invert(d);
    st.sp = 0;
    NodeBox cur = load_node(sc.nodes, sc.macro_root);
    if (cnt) cnt->nodes++;
    if (!box_hit<true>(o, inv_d, hit.t, cur.bmin, cur.bmax)) return;

    uint32_t leaf_cur = 0, leaf_end = 0;
    bool have = true; // macro phase: cur is an entered node
    while (true) {
        if (have) {
            if (!(cur.a & RH_LEAF_NODE_BIT)) {
                have = enter_children<true>(sc, cur, o, d, inv_d, hit.t, st, cur, cnt);
                continue;
            }
            leaf_cur = cur.a & RH_PRIM_INDEX_BITS;
            leaf_end = leaf_cur + cur.b;
            have = false;
        }
        // instances of the current macro leaf, in index order
        while (leaf_cur < leaf_end) {
            const InstanceRec mi = load_instance(sc, leaf_cur++);
            const uint32_t mi_index = mi.mi_index;
            if (cnt) cnt->instances++;
            if (!box_hit<true>(o, inv_d, hit.t, mi.bmin, mi.bmax)) continue;
            f3 mo, md;
            transform_ray(mi, o, d, mo, md);
            const f3 minv = f3{1.0f / md.x, 1.0f / md.y, 1.0f / md.z};
            const NodeBox mc = load_node(sc.nodes, mi.root);
            if (cnt) cnt->nodes++;
            const bool finite = __builtin_isfinite(minv.x) && __builtin_isfinite(minv.y) && __builtin_isfinite(minv.z);
            const bool solid = finite ? walk_micro<AnyHit, true>(sc, mc, mo, md, minv, (int)mi_index, hit, st, cnt)
                                      : walk_micro<AnyHit, false>(sc, mc, mo, md, minv, (int)mi_index, hit, st, cnt);
            if (solid) {
                st.sp = 0;
                return; // solid occluder found
            }
        }
        have = pop_entered<true>(sc, 0, o, inv_d, hit.t, st, cur, cnt);
        if (!have) break;
    }
}

// ---------------------------------------------------------------- 8-wide trees
// The reference's default trees (use_wide_bvh): FlattenBVH's 8-wide nodes,
// walked as Traverse_{Macro,Micro}Tree_WithStack_ClosestHit (CoreRef.cpp:
// 1133-1222, 1395-1468) walk them.  Device form of node i: 8 slots of two
// float4 at wnodes[16 i + 2 s]: a = (bbox_min, word), b = (bbox_max, count).
// The used slots come first, in the reference's slot order (which is all the
// walk depends on: unused slots never hit), then RH_MBVH_EMPTY slots.  word =
// an interior child's node index, or prim index | RH_LEAF_NODE_BIT for a leaf
// child (count = its prim count: leaves have no device node of their own), or
// RH_MBVH_EMPTY.  A word that names an interior node carries kWideHalf when
// that node uses at most 4 slots: its second 128-byte line holds only empty
// slots and is not read (most nodes of a typical tree: 71 % of a 100k-triangle
// soup's).  Otherwise all 8 slots are loaded at once whatever their use: loads
// issued under conditions that wait on a slot's own word measured 16 % slower;
// kWideHalf is known before the node's first load.  A tree root that is a leaf
// holds itself in slot 0 (see wide_root).  Stacked leaf children are
// RH_LEAF_NODE_BIT | (8 node + s), a reference to the slot that describes them.
constexpr uint32_t kWideSlotRef = RH_LEAF_NODE_BIT;
constexpr uint32_t kWideLeafRoot = 0x40000000u;
constexpr uint32_t kWideHalf = 0x20000000u;
constexpr uint32_t kWideIndex = kWideHalf - 1u; // node indices are below 2^28
// A leaf child's slot word may carry its triangle count (1..15) in bits 27-30
// when its first index is below 2^27 (0 there: the count is only in the
// slot's count word), so entering it needs no further load, and such a word
// is stacked as it is (a stacked slot reference, node x 8 + slot below 2^27,
// has those bits clear).  Only trees of fewer than 2^24 device nodes do this
// (Scene::wide_leaf_counts).
constexpr uint32_t kLeafCountShift = 27, kLeafCountBits = 15u << kLeafCountShift;
// the walk's form of a leaf child's slot word: prim index | leaf bit
__device__ __forceinline__ uint32_t leaf_word(uint32_t w) {
    return (w & kLeafCountBits) ? (w & ~kLeafCountBits) : w;
}

// slab test of one slot with its entry distance (bbox_test_oct, CoreRef.cpp:133-161)
template <bool Finite>
__device__ __forceinline__ bool slot_hit(f3 o, f3 inv_d, float t, float4 a, float4 b, float &tmin) {
    if (Finite) {
        const float lx = inv_d.x * (a.x - o.x), hx = inv_d.x * (b.x - o.x);
        const float ly = inv_d.y * (a.y - o.y), hy = inv_d.y * (b.y - o.y);
        const float lz = inv_d.z * (a.z - o.z), hz = inv_d.z * (b.z - o.z);
        tmin = __builtin_fmaxf(__builtin_fmaxf(__builtin_fminf(lx, hx), __builtin_fminf(ly, hy)),
                               __builtin_fminf(lz, hz));
        const float tmax =
            __builtin_fminf(__builtin_fminf(__builtin_fmaxf(lx, hx), __builtin_fmaxf(ly, hy)), __builtin_fmaxf(lz, hz)) *
            1.00000024f;
        return tmin <= tmax && tmin <= t && tmax > 0;
    }
    float lo_x = inv_d.x * (a.x - o.x), hi_x = inv_d.x * (b.x - o.x);
    if (lo_x > hi_x) {
        const float tmp = lo_x;
        lo_x = hi_x;
        hi_x = tmp;
    }
    float lo_y = inv_d.y * (a.y - o.y), hi_y = inv_d.y * (b.y - o.y);
    if (lo_y > hi_y) {
        const float tmp = lo_y;
        lo_y = hi_y;
        hi_y = tmp;
    }
    float lo_z = inv_d.z * (a.z - o.z), hi_z = inv_d.z * (b.z - o.z);
    if (lo_z > hi_z) {
        const float tmp = lo_z;
        lo_z = hi_z;
        hi_z = tmp;
    }
    tmin = lo_x > lo_y ? lo_x : lo_y;
    if (lo_z > tmin) tmin = lo_z;
    float tmax = hi_x < hi_y ? hi_x : hi_y;
    if (hi_z < tmax) tmax = hi_z;
    tmax *= 1.00000024f;
    return tmin <= tmax && tmin <= t && tmax > 0;
}

// f(integral_constant<int, i>) for i = 0..N-1, unrolled by instantiation: the
// array indices below are constants from the start, so the arrays live in
// registers (a loop unrolled later leaves indexed loads that put them in
// scratch memory)
template <int N, int I = 0, class F> __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, I + 1>(f);
    }
}

// v[i] for a per-lane i as a tree of selects on the bits of i
template <class T> __device__ __forceinline__ T pick8(const T (&v)[8], uint32_t i) {
    const bool b0 = i & 1u, b1 = i & 2u, b2 = i & 4u;
    const T a0 = b0 ? v[1] : v[0], a1 = b0 ? v[3] : v[2], a2 = b0 ? v[5] : v[4], a3 = b0 ? v[7] : v[6];
    const T c0 = b1 ? a1 : a0, c1 = b1 ? a3 : a2;
    return b2 ? c1 : c0;
}

// The wide node the walk is in: an interior node's index, or a leaf's range.
struct WideCur {
    uint32_t word, count;
};

// Tests the slots of interior node `node` and picks the next node as the
// reference does: one hit slot is entered; of two, the nearer is entered and
// the other stacked; of three or four, all are stacked in slot order, the top
// ordered by TraversalStack::sort_top3/sort_top4 (CoreRef.cpp:192-237, their
// handling of equal distances included) and the nearest popped; of five or
// more, the same with the stable insertion sort of sort_topN, done here by
// writing each entry straight to its sorted place.  false: no slot hit.
// (wide_order: the ordering part, given the hit slots' distances and words;
// returns the slot entered)
__device__ __forceinline__ uint32_t wide_order(const float (&dist)[8], const uint32_t (&word)[8], uint32_t mask,
                                               uint32_t node, Stack &st) {
    const uint32_t k = __builtin_popcount(mask);
    uint32_t top; // slot entered next
    if (k == 1) {
        top = __builtin_ctz(mask);
    } else if (k <= 4) {
        // slots in push (slot) order
        uint32_t e[4];
        float d[4];
        uint32_t m = mask;
        static_for<4>([&](auto C) {
            constexpr int c = C;
            e[c] = __builtin_ctz(m | 0x100u); // 8 once the hit slots are used up
            m &= m - 1;
            d[c] = pick8(dist, e[c]);
        });
        auto swap = [&](int x, int y) {
            const uint32_t te = e[x];
            e[x] = e[y];
            e[y] = te;
            const float td = d[x];
            d[x] = d[y];
            d[y] = td;
        };
        if (k == 2) {
            // nearer entered, the other stacked (dist[i] < dist[i2] picks i)
            if (!(d[0] < d[1])) swap(0, 1);
            swap(0, 1); // e[1] = entered, e[0] = stacked
        } else if (k == 3) {
            if (d[0] > d[1]) {
                if (!(d[1] > d[2])) {
                    if (d[0] > d[2]) {
                        swap(1, 2);
                    } else { // [2, 0, 1]
                        swap(0, 2);
                        swap(1, 2);
                    }
                }
            } else {
                if (d[0] > d[2]) {
                    swap(0, 1);
                } else if (d[2] > d[1]) {
                    swap(0, 2);
                } else { // [1, 2, 0]
                    swap(0, 1);
                    swap(1, 2);
                }
            }
        } else {
            if (d[0] < d[1]) swap(0, 1);
            if (d[2] < d[3]) swap(2, 3);
            if (d[0] < d[2]) swap(0, 2);
            if (d[1] < d[3]) swap(1, 3);
            if (d[1] < d[2]) swap(1, 2);
        }
        // stack e[0..k-2] bottom to top, enter e[k-1]
        static_for<3>([&](auto C) {
            constexpr uint32_t c = C;
            if (c + 1 < k) {
                const uint32_t w = pick8(word, e[c]);
                st.push2(((w & RH_LEAF_NODE_BIT) && !(w & kLeafCountBits)) ? (kWideSlotRef | (node * 8 + e[c])) : w,
                         d[c]);
            }
        });
        top = k == 2 ? e[1] : (k == 3 ? e[2] : e[3]);
    } else {
        // stable order by decreasing distance (ties keep slot order): slot s
        // goes to sp + #{hit t: d_t > d_s} + #{hit t < s: d_t == d_s}
        top = 0;
        static_for<8>([&](auto S) {
            constexpr uint32_t s = S;
            if (!((mask >> s) & 1u)) return;
            uint32_t pos = 0;
            static_for<8>([&](auto U) {
                constexpr uint32_t u = U;
                if (u != s && ((mask >> u) & 1u))
                    pos += (dist[u] > dist[s] || (dist[u] == dist[s] && u < s)) ? 1u : 0u;
            });
            if (pos == k - 1) {
                top = s;
            } else {
                st.put(st.sp + pos,
                       ((word[s] & RH_LEAF_NODE_BIT) && !(word[s] & kLeafCountBits)) ? (kWideSlotRef | (node * 8 + s))
                                                                                   : word[s],
                       dist[s]);
            }
        });
        if (st.sp + k - 1 <= st.lds_depth + st.spill_depth) st.sp += k - 1;
        else st.sp = st.lds_depth + st.spill_depth;
    }
    return top;
}

// wide_step on the node at p in the node array (node: the word naming it,
// for its flags and the slot references it stacks), its slots read by
// slot(S, a, bm) for a constant slot index S: from p, or from a copy the
// caller holds
template <bool Finite, class SlotRead>
__device__ __forceinline__ bool wide_step_at(const float4 *p, SlotRead slot, uint32_t node, f3 o, f3 inv_d, float t,
                                             Stack &st, WideCur &next) {
    const bool full = !(node & kWideHalf);
    node &= kWideIndex;
    float dist[8];
    uint32_t word[8];
    uint32_t mask = 0;
    static_for<8>([&](auto S) {
        constexpr int s = S;
        // bbox_max without the count word: the count is read below for the
        // entered slot only, when it is a leaf (8 registers fewer in flight)
        float4 a = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(RH_MBVH_EMPTY));
        float3 bm = make_float3(0.0f, 0.0f, 0.0f);
        if (s < 4 || full) slot(S, a, bm);
        const float4 b = make_float4(bm.x, bm.y, bm.z, 0.0f);
        word[s] = __float_as_uint(a.w);
        const bool h = slot_hit<Finite>(o, inv_d, t, a, b, dist[s]);
        if (h && word[s] != RH_MBVH_EMPTY) mask |= 1u << s;
    });
    if (!mask) return false;
    const uint32_t top = wide_order(dist, word, mask, node, st);
    const uint32_t w = pick8(word, top);
    next.word = w;
    next.count = 0;
    if (w & RH_LEAF_NODE_BIT) {
        const uint32_t c = (w & kLeafCountBits) >> kLeafCountShift;
        next.word = leaf_word(w);
        next.count = c ? c : __float_as_uint(reinterpret_cast<const float *>(p + 2 * top + 1)[3]);
    }
    return true;
}

template <bool Finite>
__device__ __forceinline__ bool wide_step(const float4 *wn, uint32_t node, f3 o, f3 inv_d, float t, Stack &st,
                                          WideCur &next) {
    const float4 *p = wn + (size_t)(node & kWideIndex) * 16;
    auto slot = [&](auto S, float4 &a, float3 &bm) {
        constexpr int s = S;
        a = p[2 * s];
        bm = *reinterpret_cast<const float3 *>(p + 2 * s + 1);
    };
    return wide_step_at<Finite>(p, slot, node, o, inv_d, t, st, next);
}

// wide_step for a ray walked by G lanes (G = 2, 4 or 8, aligned lane groups of
// one wave) that hold identical copies of the ray's state: lane j of the
// group loads and tests slots j, j + G, ... (the group's loads of a node's
// first 128-byte line are one cache line), the hit mask comes from ballots
// and the distances and words from the group's lanes, and every lane then
// orders and stacks the same entries as wide_step, so that the group's
// copies stay identical (its lanes write the same stack entries).  All G
// lanes must be active.
template <bool Finite, int G>
__device__ __forceinline__ bool wide_step_group(const float4 *wn, uint32_t node, f3 o, f3 inv_d, float t, Stack &st,
                                                WideCur &next) {
    static_assert(G == 2 || G == 4 || G == 8, "groups of 2, 4 or 8 lanes");
    constexpr int S = 8 / G; // slots per lane
    const uint32_t lane = __lane_id(), gb = lane & ~(uint32_t)(G - 1), j = lane & (G - 1);
    const bool full = !(node & kWideHalf);
    node &= kWideIndex;
    const float4 *p = wn + (size_t)node * 16;
    float dl[S];
    uint32_t wl[S], cl[S];
    uint32_t mask = 0;
    static_for<S>([&](auto I) {
        constexpr int i = I;
        const uint32_t s = j + i * G;
        float4 a = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(RH_MBVH_EMPTY));
        float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (s < 4 || full) {
            a = p[2 * s];
            b = p[2 * s + 1];
        }
        wl[i] = __float_as_uint(a.w);
        cl[i] = __float_as_uint(b.w);
        const bool h = slot_hit<Finite>(o, inv_d, t, a, b, dl[i]) && wl[i] != RH_MBVH_EMPTY;
        mask |= (uint32_t)((__ballot(h) >> gb) & ((1ull << G) - 1)) << (i * G);
    });
    if (!mask) return false;
    float dist[8];
    uint32_t word[8];
    static_for<8>([&](auto SS) {
        constexpr int s = SS;
        dist[s] = __shfl(dl[s / G], (int)(gb + s % G));
        word[s] = __shfl(wl[s / G], (int)(gb + s % G));
    });
    const uint32_t top = wide_order(dist, word, mask, node, st);
    next.word = pick8(word, top);
    if (next.word & RH_LEAF_NODE_BIT) next.word = leaf_word(next.word);
    uint32_t c = cl[0];
    static_for<S>([&](auto I) {
        if (top / G == (uint32_t)I) c = cl[I];
    });
    c = __shfl(c, (int)(gb + (top & (G - 1))));
    next.count = (next.word & RH_LEAF_NODE_BIT) ? c : 0u;
    return true;
}

// wide_step of one lane (G = 1) or of a lane group
template <bool Finite, int G>
__device__ __forceinline__ bool wide_step_g(const float4 *wn, uint32_t node, f3 o, f3 inv_d, float t, Stack &st,
                                            WideCur &next) {
    if constexpr (G == 1) return wide_step<Finite>(wn, node, o, inv_d, t, st, next);
    else return wide_step_group<Finite, G>(wn, node, o, inv_d, t, st, next);
}

// Pops until an entry nearer than t: false when the stack is back at base.
__device__ __forceinline__ bool wide_pop(const float4 *wn, uint32_t base, float t, Stack &st, WideCur &next,
                                         uint32_t leaf_counts) {
    while (st.sp > base) {
        uint32_t w;
        float d;
        st.pop2(w, d);
        if (d > t) continue;
        if (leaf_counts && (w & RH_LEAF_NODE_BIT) && (w & kLeafCountBits)) {
            // a leaf word with its count: no load
            next.word = w & ~kLeafCountBits;
            next.count = (w & kLeafCountBits) >> kLeafCountShift;
        } else if (w & kWideSlotRef) {
            const float *slot = reinterpret_cast<const float *>(wn + (size_t)(w & RH_PRIM_INDEX_BITS) * 2);
            next.word = leaf_word(__float_as_uint(slot[3]));
            next.count = __float_as_uint(slot[7]);
        } else {
            next.word = w;
            next.count = 0;
        }
        return true;
    }
    return false;
}

// A tree's root word: its node's index, with kWideLeafRoot set when the root
// is a leaf, whose node then describes it in slot 0.
__device__ __forceinline__ WideCur wide_root(const float4 *wn, uint32_t root_word) {
    if (!(root_word & kWideLeafRoot)) return WideCur{root_word, 0};
    const float *slot = reinterpret_cast<const float *>(wn + (size_t)(root_word & ~kWideLeafRoot) * 16);
    return WideCur{__float_as_uint(slot[3]), __float_as_uint(slot[7])};
}

// Traverse_MicroTree_WithStack_ClosestHit, 8-wide (CoreRef.cpp:1395-1468).
// The reference's wide any-hit walk takes the closest hit inside each
// instance too (CoreRef.cpp:1357), so this is the only micro walk.  With
// first_hit the walk ends at the first solid triangle it hits instead: for a
// scene whose triangles are all solid, whether a segment is blocked does not
// depend on which blocking triangle is found (visibility queries only).
template <bool Finite, int G = 1>
__device__ __forceinline__ bool walk_micro_wide(const Scene &sc, uint32_t root, f3 mo, f3 md, f3 minv, int mi_index,
                                                Hit &hit, Stack &st, TraceCounters *cnt, bool first_hit,
                                                const float4 *root_node = nullptr) {
    const uint32_t base = st.sp;
    WideCur cur = wide_root(sc.wnodes, root);
    bool res = false;
    if (root_node && !(cur.word & RH_LEAF_NODE_BIT)) {
        // the root's step from the copy of its node the caller holds (the
        // same slots as the node array's: its index is read by constants)
        if (cnt) cnt->nodes++;
        auto slot = [&](auto S, float4 &a, float3 &bm) {
            constexpr int s = S;
            a = root_node[2 * s];
            const float4 b = root_node[2 * s + 1];
            bm = make_float3(b.x, b.y, b.z);
        };
        if (!wide_step_at<Finite>(sc.wnodes + (size_t)(cur.word & kWideIndex) * 16, slot, cur.word, mo, minv, hit.t,
                                  st, cur) &&
            !wide_pop(sc.wnodes, base, hit.t, st, cur, sc.wide_leaf_counts))
            return false;
    }
    while (true) {
        if (cur.word & RH_LEAF_NODE_BIT) {
            NodeBox leaf;
            leaf.a = cur.word;
            leaf.b = cur.count;
            const int before = hit.prim;
            const float t_before = hit.t;
            if (first_hit) {
                if (intersect_leaf<true>(sc, leaf, mo, md, mi_index, hit, cnt)) {
                    st.sp = base;
                    return true;
                }
            } else {
                intersect_leaf<false>(sc, leaf, mo, md, mi_index, hit, cnt);
            }
            res |= hit.t != t_before || hit.prim != before;
        } else {
            if (cnt) cnt->nodes++;
            if (wide_step_g<Finite, G>(sc.wnodes, cur.word, mo, minv, hit.t, st, cur)) continue;
        }
        if (!wide_pop(sc.wnodes, base, hit.t, st, cur, sc.wide_leaf_counts)) return res;
    }
}

// Traverse_MacroTree_WithStack_{ClosestHit,AnyHit}, 8-wide (CoreRef.cpp:
// 1133-1222, 1270-1368): an any-hit walk ends at the first instance whose
// closest hit is on a solid triangle.  G > 1: the ray is walked by a lane
// group (wide_step_group), every other step done alike by all its lanes.
template <bool AnyHit, int G = 1>
__device__ __forceinline__ void traverse_wide(const Scene &sc, f3 o, f3 d, Hit &hit, Stack &st, TraceCounters *cnt,
                                              bool first_hit = false) {
    const f3 inv_d = safe_invert(d);
    st.sp = 0;
    // an instance of a macro leaf: true when the walk ends (an any-hit walk
    // found a solid triangle)
    auto instance = [&](const InstanceRec &mi, const float4 *root_node) -> bool {
        const uint32_t mi_index = mi.mi_index;
        if (cnt) cnt->instances++;
        if (!box_hit<true>(o, inv_d, hit.t, mi.bmin, mi.bmax)) return false;
        f3 mo, md;
        transform_ray(mi, o, d, mo, md);
        const f3 minv = f3{1.0f / md.x, 1.0f / md.y, 1.0f / md.z};
        const uint32_t root = mi.root;
        const bool finite = __builtin_isfinite(minv.x) && __builtin_isfinite(minv.y) && __builtin_isfinite(minv.z);
        const bool fh = AnyHit && first_hit;
        const bool found =
            finite ? walk_micro_wide<true, G>(sc, root, mo, md, minv, (int)mi_index, hit, st, cnt, fh, root_node)
                   : walk_micro_wide<false, G>(sc, root, mo, md, minv, (int)mi_index, hit, st, cnt, fh);
        if (AnyHit && found) {
            // first_hit (every triangle solid): the walk found a solid one
            if (fh) {
                st.sp = 0;
                return true;
            }
            const rh_tri_accel ht = load_tri(sc.tris, (uint32_t)hit.prim);
            if (ht.ci & RH_TRI_SOLID_BIT) {
                st.sp = 0;
                return true;
            }
        }
        return false;
    };
    if (!AnyHit && sc.macro_single) {
        // the leaf root's one instance, then nothing is stacked (closest-hit
        // walks only: in the any-hit walk the second copy of the instance
        // walk costs the shadow kernel more registers than it saves)
        instance(instance_rec(sc.macro_inst), (G == 1 && sc.root_cached) ? sc.root_node : nullptr);
        return;
    }
    WideCur cur = WideCur{sc.macro_word, sc.macro_count};
    while (true) {
        if (cur.word & RH_LEAF_NODE_BIT) {
            const uint32_t first = cur.word & RH_PRIM_INDEX_BITS, end = first + cur.count;
            for (uint32_t i = first; i < end; i++)
                if (instance(load_instance(sc, i), nullptr)) return;
        } else {
            if (cnt) cnt->nodes++;
            if (wide_step_g<true, G>(sc.wnodes, cur.word, o, inv_d, hit.t, st, cur)) continue;
        }
        if (!wide_pop(sc.wnodes, 0, hit.t, st, cur, sc.wide_leaf_counts)) return;
    }
}

// The scene's trees: Tree = kTreeBinary or kTreeWide compiles one walk (the
// kernel is launched only for scenes of that kind), kTreeAny both, chosen per
// scene at run time (at the register cost of the larger).
constexpr int kTreeBinary = 0, kTreeWide = 1, kTreeAny = 2;
// first_hit: see walk_micro_wide (the binary any-hit walk already stops at
// the first solid triangle, as the reference's does).
template <bool AnyHit, int Tree = kTreeAny>
__device__ __forceinline__ void traverse(const Scene &sc, f3 o, f3 d, Hit &hit, Stack &st, TraceCounters *cnt,
                                         bool first_hit = false) {
    if (Tree == kTreeWide || (Tree == kTreeAny && sc.wnodes)) traverse_wide<AnyHit>(sc, o, d, hit, st, cnt, first_hit);
    else traverse_fast<AnyHit>(sc, o, d, hit, st, cnt);
}

// ---------------------------------------------------------------- textures
// Texel fetch at reference atlas coordinates (x, y) on the atlas page where the
// texture's mip block sits: the block's own compact storage holds it.
// (px, py) is the block's atlas position, the texture's pos[lod].
__device__ __forceinline__ uint32_t atlas_index(const Scene &sc, uint32_t tex, int lod, int px, int py, int x, int y) {
    const rh_tex_mip m = sc.tex_mips[tex * RH_NUM_MIP_LEVELS + lod];
    int lx = x - px, ly = y - py;
    lx = clampi(lx, 0, (int)m.bw - 1);
    ly = clampi(ly, 0, (int)m.bh - 1);
    return m.offset + (uint32_t)ly * m.bw + (uint32_t)lx;
}
__device__ __forceinline__ uint32_t atlas_get(const Scene &sc, uint32_t tex, int lod, int px, int py, int x, int y) {
    return sc.texels[atlas_index(sc, tex, lod, px, py, x, y)];
}

// The texels a bilinear sample read (indices into the texel array: 00, 01,
// 10, 11) and its weights: value = sum of texel * weight, w00 = (1-kx)(1-ky),
// w01 = kx(1-ky), w10 = (1-kx)ky, w11 = kx ky.
struct TexFoot {
    uint32_t idx[4];
    float kx, ky;
    int lod;
};

__device__ __forceinline__ float texel_ch(uint32_t p, int c) { return (float)((p >> (8 * c)) & 0xffu); }

// SampleBilinear (CoreRef.cpp:1652-1678) with TransformUV (CoreRef.cpp:1625-1632)
__device__ __forceinline__ f4 sample_bilinear(const Scene &sc, uint32_t tex, f2 uvs, int lod, TexFoot *foot = nullptr) {
    // fields read in place: copying the record and indexing its arrays by lod
    // would put it in scratch memory
    const rh_texture *t = sc.textures + tex;
    const int ipx = (int)t->pos[lod][0], ipy = (int)t->pos[lod][1];
    const float atlas = (float)RH_TEXTURE_ATLAS_SIZE;
    const float posx = (float)ipx, posy = (float)ipy;
    const float sizex = (float)(((uint32_t)t->width & RH_TEXTURE_WIDTH_BITS) >> lod);
    const float sizey = (float)((uint32_t)t->height >> lod);
    const float ux = uvs.x - simd_floor(uvs.x), uy = uvs.y - simd_floor(uvs.y);
    float rx = posx + ux * sizex + 1.0f, ry = posy + uy * sizey + 1.0f;
    rx /= atlas;
    ry /= atlas;
    rx = rx * atlas - 0.5f;
    ry = ry * atlas - 0.5f;
    const int ix = (int)rx, iy = (int)ry;
    const uint32_t i00 = atlas_index(sc, tex, lod, ipx, ipy, ix + 0, iy + 0);
    const uint32_t i01 = atlas_index(sc, tex, lod, ipx, ipy, ix + 1, iy + 0);
    const uint32_t i10 = atlas_index(sc, tex, lod, ipx, ipy, ix + 0, iy + 1);
    const uint32_t i11 = atlas_index(sc, tex, lod, ipx, ipy, ix + 1, iy + 1);
    const uint32_t p00 = sc.texels[i00], p01 = sc.texels[i01], p10 = sc.texels[i10], p11 = sc.texels[i11];
    const float kx = rx - __builtin_floorf(rx), ky = ry - __builtin_floorf(ry);
    if (foot) *foot = TexFoot{{i00, i01, i10, i11}, kx, ky, lod};
    float r[4];
    const float k = 1.0f / 255.0f;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float p0 = texel_ch(p01, c) * kx + texel_ch(p00, c) * (1 - kx);
        const float p1 = texel_ch(p11, c) * kx + texel_ch(p10, c) * (1 - kx);
        r[c] = (p1 * ky + p0 * (1.0f - ky)) * k;
    }
    return f4{r[0], r[1], r[2], r[3]};
}

// fast_log2 (CoreRef.cpp:391-400)
__device__ __forceinline__ float fast_log2(float val) {
    int32_t x = __float_as_int(val);
    float log_2 = (float)(((x >> 23) & 255) - 128);
    x &= ~(255 << 23);
    x += 127 << 23;
    const float v = __int_as_float(x);
    log_2 += ((-0.34484843f) * v + 2.02466578f) * v - 0.67487759f;
    return log_2;
}

// get_texture_lod (CoreRef.cpp:402-415)
__device__ __forceinline__ float get_texture_lod(const rh_texture &t, f2 duv_dx, f2 duv_dy) {
    const float szx = (float)((uint32_t)t.width & RH_TEXTURE_WIDTH_BITS), szy = (float)t.height;
    const f2 dx = f2{duv_dx.x * szx, duv_dx.y * szy}, dy = f2{duv_dy.x * szx, duv_dy.y * szy};
    const f2 dg = f2{dx.x + dy.x, dx.y + dy.y};
    float l_dx = 0.0f, l_dy = 0.0f, l_dg = 0.0f;
    l_dx += dx.x * dx.x;
    l_dx += dx.y * dx.y;
    l_dy += dy.x * dy.x;
    l_dy += dy.y * dy.y;
    l_dg += dg.x * dg.x;
    l_dg += dg.y * dg.y;
    const float min_length2 = smin(smin(l_dx, l_dy), l_dg);
    float lod = fast_log2(min_length2);
    lod = clampf(0.5f * lod - 1.0f, 0.0f, (float)RH_MAX_MIP_LEVEL);
    return lod;
}

__device__ __forceinline__ f4 srgb_to_rgb(f4 c) {
    float v[3] = {c.x, c.y, c.z};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (v[i] > 0.04045f) {
            v[i] = rpow((v[i] + 0.055f) / 1.055f, 2.4f);
        } else {
            v[i] = v[i] / 12.92f;
        }
    }
    return f4{v[0], v[1], v[2], c.w};
}

// SampleLatlong_RGBE (CoreRef.cpp:1772-1799)
__device__ __forceinline__ f4 sample_latlong_rgbe(const Scene &sc, uint32_t tex, f3 dir) {
    const rh_texture t = sc.textures[tex];
    const float theta = racos(clampf(dir.y, -1.0f, 1.0f)) / RH_PI;
    const float r = __builtin_sqrtf(dir.x * dir.x + dir.z * dir.z);
    float u = 0.5f * racos(r > RH_FLT_EPS ? clampf(dir.x / r, -1.0f, 1.0f) : 0.0f) / RH_PI;
    if (dir.z < 0.0f) u = 1.0f - u;
    const float posx = (float)t.pos[0][0], posy = (float)t.pos[0][1];
    const float sizex = (float)((uint32_t)t.width & RH_TEXTURE_WIDTH_BITS), sizey = (float)t.height;
    const float ux = posx + u * sizex + 1.0f, uy = posy + theta * sizey + 1.0f;
    const int ipx = (int)t.pos[0][0], ipy = (int)t.pos[0][1];
    const uint32_t p00 = atlas_get(sc, tex, 0, ipx, ipy, (int)(ux + 0), (int)(uy + 0));
    const uint32_t p01 = atlas_get(sc, tex, 0, ipx, ipy, (int)(ux + 1), (int)(uy + 0));
    const uint32_t p10 = atlas_get(sc, tex, 0, ipx, ipy, (int)(ux + 0), (int)(uy + 1));
    const uint32_t p11 = atlas_get(sc, tex, 0, ipx, ipy, (int)(ux + 1), (int)(uy + 1));
    const float kx = ux - simd_floor(ux), ky = uy - simd_floor(uy);
    auto rgbe = [](uint32_t p, float *o) {
        const float f = exp2f((float)((p >> 24) & 0xffu) - 128.0f);
        o[0] = to_norm_float(p & 0xffu) * f;
        o[1] = to_norm_float((p >> 8) & 0xffu) * f;
        o[2] = to_norm_float((p >> 16) & 0xffu) * f;
        o[3] = 1.0f;
    };
    float a00[4], a01[4], a10[4], a11[4], res[4];
    rgbe(p00, a00);
    rgbe(p01, a01);
    rgbe(p10, a10);
    rgbe(p11, a11);
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const float p0 = a01[c] * kx + a00[c] * (1 - kx);
        const float p1 = a11[c] * kx + a10[c] * (1 - kx);
        res[c] = p1 * ky + p0 * (1 - ky);
    }
    return f4{res[0], res[1], res[2], res[3]};
}

// ---------------------------------------------------------------- shading
// ExtractPlaneNormal (Core.cpp:216-228); note the reference scales by the
// length rather than dividing by it
__device__ __forceinline__ f3 extract_plane_normal(const rh_tri_accel &tri) {
    // component w is 1, (u, v) = the next two axes in (1,2), (0,2), (0,1) order
    const int iw = tri.ci & RH_TRI_W_BITS;
    const float n0 = iw == 0 ? 1.0f : tri.nu;
    const float n1 = iw == 1 ? 1.0f : (iw == 0 ? tri.nu : tri.nv);
    const float n2 = iw == 2 ? 1.0f : tri.nv;
    float inv_l = __builtin_sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
    if (tri.ci & RH_TRI_INV_NORMAL_BIT) inv_l = -inv_l;
    return f3{n0 * inv_l, n1 * inv_l, n2 * inv_l};
}

// TransformNormal (CoreRef.cpp:1625-1629)
__device__ __forceinline__ f3 transform_normal(f3 n, const float *inv) {
    return f3{inv[0] * n.x + inv[1] * n.y + inv[2] * n.z, inv[4] * n.x + inv[5] * n.y + inv[6] * n.z,
              inv[8] * n.x + inv[9] * n.y + inv[10] * n.z};
}

// BRDF_OrenNayar (CoreRef.cpp:1579-1603)
__device__ __forceinline__ float brdf_oren_nayar(f3 L, f3 I, f3 N, f3 T, float sigma) {
    const float sigma_sqr = sigma * sigma;
    const float A = 1.0f - (sigma_sqr / (2.0f * (sigma_sqr + 0.33f)));
    const float B = (0.45f * sigma_sqr) / (sigma_sqr + 0.09f);
    const float cos_theta_I = dot(L, N);
    const float cos_theta_O = -dot(I, N);
    const float sin_theta_I = __builtin_sqrtf(smax(0.0f, 1.0f - (cos_theta_I * cos_theta_I)));
    const float sin_theta_O = __builtin_sqrtf(smax(0.0f, 1.0f - (cos_theta_O * cos_theta_O)));
    const float cos_phi_I = dot(L, T);
    const float cos_phi_O = -dot(I, T);
    const float sin_phi_I = __builtin_sqrtf(smax(0.0f, 1.0f - (cos_phi_I * cos_phi_I)));
    const float sin_phi_O = __builtin_sqrtf(smax(0.0f, 1.0f - (cos_phi_O * cos_phi_O)));
    const float temp = smax(0.0f, (cos_phi_I * cos_phi_O) + (sin_phi_I * sin_phi_O));
    float sin_alpha, tan_beta;
    if (cos_theta_I > cos_theta_O) {
        sin_alpha = sin_theta_O;
        tan_beta = (sin_theta_I / cos_theta_I);
    } else {
        sin_alpha = sin_theta_I;
        tan_beta = (sin_theta_O / cos_theta_O);
    }
    const float result = (A + (B * temp * sin_alpha * tan_beta));
    return clampf(result, 0.0f, 1.0f);
}

struct Derivatives {
    f3 do_dx, do_dy, dd_dx, dd_dy;
    f2 duv_dx, duv_dy;
    f3 dndx, dndy;
    float ddn_dx, ddn_dy;
};

// ComputeDerivatives (CoreRef.cpp:1993-2052)
__device__ __forceinline__ void compute_derivatives(f3 I, float t, f3 do_dx, f3 do_dy, f3 dd_dx, f3 dd_dy,
                                                    const rh_vertex &v1, const rh_vertex &v2, const rh_vertex &v3,
                                                    f3 plane_N, Derivatives &out) {
    const float dot_I_N = -dot(I, plane_N);
    const float inv_dot = fabsf(dot_I_N) < RH_FLT_EPS ? 0.0f : 1.0f / dot_I_N;
    const float dt_dx = dot(do_dx + t * dd_dx, plane_N) * inv_dot;
    const float dt_dy = dot(do_dy + t * dd_dy, plane_N) * inv_dot;

    out.do_dx = (do_dx + t * dd_dx) + dt_dx * I;
    out.do_dy = (do_dy + t * dd_dy) + dt_dy * I;
    out.dd_dx = dd_dx;
    out.dd_dy = dd_dy;

    const f2 duv13 = f2{v1.t[0][0] - v3.t[0][0], v1.t[0][1] - v3.t[0][1]};
    const f2 duv23 = f2{v2.t[0][0] - v3.t[0][0], v2.t[0][1] - v3.t[0][1]};
    const f3 dp13 = ld3(v1.p) - ld3(v3.p), dp23 = ld3(v2.p) - ld3(v3.p);

    const float det_uv = duv13.x * duv23.y - duv13.y * duv23.x;
    const float inv_det_uv = fabsf(det_uv) < RH_FLT_EPS ? 0 : 1.0f / det_uv;
    const f3 dpdu = (duv23.y * dp13 - duv13.y * dp23) * inv_det_uv;
    const f3 dpdv = ((-duv23.x) * dp13 + duv13.x * dp23) * inv_det_uv;

    f2 A0 = f2{dpdu.x, dpdv.x}, A1 = f2{dpdu.y, dpdv.y};
    f2 Bx = f2{out.do_dx.x, out.do_dx.y};
    f2 By = f2{out.do_dy.x, out.do_dy.y};
    if (fabsf(plane_N.x) > fabsf(plane_N.y) && fabsf(plane_N.x) > fabsf(plane_N.z)) {
        A0 = f2{dpdu.y, dpdv.y};
        A1 = f2{dpdu.z, dpdv.z};
        Bx = f2{out.do_dx.y, out.do_dx.z};
        By = f2{out.do_dy.y, out.do_dy.z};
    } else if (fabsf(plane_N.y) > fabsf(plane_N.z)) {
        A1 = f2{dpdu.z, dpdv.z};
        Bx = f2{out.do_dx.x, out.do_dx.z};
        By = f2{out.do_dy.x, out.do_dy.z};
    }
    const float det = A0.x * A1.y - A0.y * A1.x;
    const float inv_det = fabsf(det) < RH_FLT_EPS ? 0 : 1.0f / det;
    out.duv_dx = f2{(A1.y * Bx.x - A0.y * Bx.y) * inv_det, (A0.x * Bx.y - A1.x * Bx.x) * inv_det};
    out.duv_dy = f2{(A1.y * By.x - A0.y * By.y) * inv_det, (A0.x * By.y - A1.x * By.x) * inv_det};

    const f3 dn1 = ld3(v1.n) - ld3(v3.n), dn2 = ld3(v2.n) - ld3(v3.n);
    const f3 dndu = (duv23.y * dn1 - duv13.y * dn2) * inv_det_uv;
    const f3 dndv = ((-duv23.x) * dn1 + duv13.x * dn2) * inv_det_uv;
    out.dndx = dndu * out.duv_dx.x + dndv * out.duv_dx.y;
    out.dndy = dndu * out.duv_dy.x + dndv * out.duv_dy.y;
    out.ddn_dx = dot(dd_dx, plane_N) + dot(I, out.dndx);
    out.ddn_dy = dot(dd_dy, plane_N) + dot(I, out.dndy);
}

} // namespace rh
