// TEST INFRASTRUCTURE — the CPU oracle (see ray_oracle.h).
//
// Scalar restatement of the Ray Ref backend.  Each function names the
// reference routine it follows (paths relative to src/Ray/internal/).  Float
// arithmetic is written in the reference's operation order and the file is
// built with -ffp-contract=off, so results are bit-identical to the reference
// built the same way (checked in tests/test_oracle.py).
#include "ray_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

namespace {

struct V3 {
    float x, y, z;
};
inline V3 v3(const float *p) { return {p[0], p[1], p[2]}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 mul(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 mul(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 mulv(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 divs(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
// generic simd_vec<float,3>: sums left to right starting from zero (simd_vec.h:559-563)
inline float dot(V3 a, V3 b) {
    float r = 0;
    r += a.x * b.x;
    r += a.y * b.y;
    r += a.z * b.z;
    return r;
}
inline float len(V3 a) { return std::sqrt(dot(a, a)); }
inline V3 norm(V3 a) { return divs(a, len(a)); }
inline V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float at(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline float clampf(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
inline float fracpart(float x) {
    float ip;
    return std::modf(x, &ip);
}

// hash, construct_float (CoreRef.cpp:264-270, 359-368)
inline int hash(int x) {
    unsigned r = (unsigned)x;
    r = ((r >> 16) ^ r) * 0x45d9f3b;
    r = ((r >> 16) ^ r) * 0x45d9f3b;
    r = (r >> 16) ^ r;
    return (int)r;
}
inline float construct_float(uint32_t m) {
    m = (m & 0x007FFFFFu) | 0x3F800000u;
    float f;
    memcpy(&f, &m, 4);
    return f - 1.0f;
}
inline float bits_f(uint32_t u) {
    float f;
    memcpy(&f, &u, 4);
    return f;
}
// to_norm_float (Core.h:249-252)
inline float to_norm_float(uint32_t v) { return bits_f(0x3f800000u + v * 0x8080u + (v + 1u) / 2u) - 1.0f; }

const int kPrimes[RH_HALTON_COUNT] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89};

// ComputeRadicalInversePermutations + Shuffle (Halton.h:6-43), mt19937(0) (RendererRef.cpp:11-12)
std::vector<uint16_t> perms() {
    auto rnd = std::bind(std::uniform_int_distribution<int>(), std::mt19937(0));
    std::vector<uint16_t> out;
    for (int p : kPrimes) {
        std::vector<uint16_t> a(p);
        std::iota(a.begin(), a.end(), 0);
        for (int i = 0; i < p; i++) {
            const uint32_t thres = (~((uint32_t)p) + 1u) % p;
            int other;
            for (;;) {
                const uint32_t r = rnd();
                if (r >= thres) {
                    other = r % p;
                    break;
                }
            }
            std::swap(a[i], a[other]);
        }
        out.insert(out.end(), a.begin(), a.end());
    }
    return out;
}

// ScrambledRadicalInverse (Halton.h:60-74)
float scrambled_ri(int base, const uint16_t *perm, uint64_t a) {
    const float inv_base = 1.0f / base;
    uint64_t rev = 0;
    float inv_base_n = 1;
    while (a) {
        const uint64_t next = a / base;
        const uint64_t digit = a - next * base;
        rev = rev * base + perm[digit];
        inv_base_n *= inv_base;
        a = next;
    }
    return std::min(inv_base_n * (rev + inv_base * perm[0] / (1 - inv_base)), 1.0f - std::numeric_limits<float>::epsilon());
}

void halton_table(const std::vector<uint16_t> &pm, int iteration, float *out) {
    for (int i = 0; i < RH_HALTON_SEQ_LEN; i++) {
        int off = 0;
        for (int d = 0; d < RH_HALTON_COUNT; d++) {
            out[i * RH_HALTON_COUNT + d] = scrambled_ri(kPrimes[d], pm.data() + off, (uint64_t)(iteration + i));
            off += kPrimes[d];
        }
    }
}

// the table with its extra row (a copy of row 0), as the device holds it
void halton_table_padded(const std::vector<uint16_t> &pm, int iteration, float *out) {
    halton_table(pm, iteration, out);
    std::copy(out, out + RH_HALTON_COUNT, out + RH_HALTON_COUNT * RH_HALTON_SEQ_LEN);
}

// ------------------------------------------------------------------ traversal
struct Hit {
    int mask = 0, obj = -1, prim = -1;
    float t = RH_MAX_DIST, u = 0, v = 0;
};

// bbox_test (CoreRef.cpp:64-86)
bool bbox_test(const float o[3], const float inv_d[3], float t, const float *bmin, const float *bmax) {
    float lo_x = inv_d[0] * (bmin[0] - o[0]), hi_x = inv_d[0] * (bmax[0] - o[0]);
    if (lo_x > hi_x) std::swap(lo_x, hi_x);
    float lo_y = inv_d[1] * (bmin[1] - o[1]), hi_y = inv_d[1] * (bmax[1] - o[1]);
    if (lo_y > hi_y) std::swap(lo_y, hi_y);
    float lo_z = inv_d[2] * (bmin[2] - o[2]), hi_z = inv_d[2] * (bmax[2] - o[2]);
    if (lo_z > hi_z) std::swap(lo_z, hi_z);
    float tmin = lo_x > lo_y ? lo_x : lo_y;
    if (lo_z > tmin) tmin = lo_z;
    float tmax = hi_x < hi_y ? hi_x : hi_y;
    if (hi_z < tmax) tmax = hi_z;
    tmax *= 1.00000024f;
    return tmin <= tmax && tmin <= t && tmax > 0;
}

// _IntersectTri (CoreRef.cpp:11-42)
bool intersect(const float o[3], const float d[3], const rh_tri_accel &tri, uint32_t idx, Hit &h) {
    static const int nu_[] = {1, 0, 0}, nv_[] = {2, 2, 1};
    const int iw = tri.ci & RH_TRI_W_BITS, iu = nu_[iw], iv = nv_[iw];
    const float det = d[iu] * tri.nu + d[iv] * tri.nv + d[iw];
    const float dett = tri.np - (o[iu] * tri.nu + o[iv] * tri.nv + o[iw]);
    const float Du = d[iu] * dett - (tri.pu - o[iu]) * det;
    const float Dv = d[iv] * dett - (tri.pv - o[iv]) * det;
    const float detu = tri.e1v * Du - tri.e1u * Dv;
    const float detv = tri.e0u * Dv - tri.e0v * Du;
    const float tmpdet0 = det - detu - detv;
    if ((tmpdet0 > -RH_HIT_EPS && detu > -RH_HIT_EPS && detv > -RH_HIT_EPS) ||
        (tmpdet0 < RH_HIT_EPS && detu < RH_HIT_EPS && detv < RH_HIT_EPS)) {
        const float rdet = 1.0f / det;
        const float t = dett * rdet;
        if (t > 0 && t < h.t) {
            h.mask = -1;
            h.prim = (int)idx;
            h.t = t;
            h.u = detu * rdet;
            h.v = detv * rdet;
            return true;
        }
    }
    return false;
}

// Traverse_MicroTree_WithStack_{ClosestHit,AnyHit} (CoreRef.cpp:1370-1393, 1470-1498)
bool micro(const rh_scene_view &sc, const float o[3], const float d[3], const float inv_d[3], uint32_t root, int obj,
           bool any, Hit &h, orc_counters *cnt) {
    bool res = false;
    std::vector<uint32_t> st;
    st.reserve(64);
    st.push_back(root);
    while (!st.empty()) {
        const uint32_t cur = st.back();
        st.pop_back();
        const rh_bvh_node &n = sc.nodes[cur];
        if (cnt) cnt->nodes++;
        if (!bbox_test(o, inv_d, h.t, n.bbox_min, n.bbox_max)) continue;
        if (!(n.prim_index & RH_LEAF_NODE_BIT)) {
            const uint32_t ax = n.prim_count >> 30, right = n.prim_count & RH_RIGHT_CHILD_BITS;
            const bool neg = d[ax] < 0;
            st.push_back(neg ? n.prim_index : right);
            st.push_back(neg ? right : n.prim_index);
        } else {
            // IntersectTris_ClosestHit / _AnyHit (CoreRef.cpp:760-825)
            bool leaf_hit = false;
            const uint32_t first = n.prim_index & RH_PRIM_INDEX_BITS;
            for (uint32_t k = first; k < first + n.prim_count; k++) {
                const uint32_t ti = sc.tri_indices[k];
                if (cnt) cnt->tris++;
                if (intersect(o, d, sc.tris[ti], ti, h)) {
                    h.obj = obj;
                    leaf_hit = true;
                }
                if (any && leaf_hit && (sc.tris[ti].ci & RH_TRI_SOLID_BIT)) break;
            }
            res |= leaf_hit;
            if (any && leaf_hit && (sc.tris[h.prim].ci & RH_TRI_SOLID_BIT)) break;
        }
    }
    return res;
}

// safe_invert (CoreRef.cpp:272-294)
void safe_invert(const float v[3], float o[3]) {
    for (int i = 0; i < 3; i++) {
        o[i] = 1.0f / v[i];
        if (v[i] <= RH_FLT_EPS && v[i] >= 0) o[i] = std::numeric_limits<float>::max();
        else if (v[i] >= -RH_FLT_EPS && v[i] < 0) o[i] = -std::numeric_limits<float>::max();
    }
}

// Traverse_MacroTree_WithStack_{ClosestHit,AnyHit} (CoreRef.cpp:1091-1131, 1224-1268)
void macro_binary(const rh_scene_view &sc, const float o[3], const float d[3], bool any, Hit &h, orc_counters *cnt) {
    if (sc.macro_nodes_root == RH_INVALID) return;
    float inv_d[3];
    safe_invert(d, inv_d);
    std::vector<uint32_t> st;
    st.push_back(sc.macro_nodes_root);
    while (!st.empty()) {
        const uint32_t cur = st.back();
        st.pop_back();
        const rh_bvh_node &n = sc.nodes[cur];
        if (cnt) cnt->nodes++;
        if (!bbox_test(o, inv_d, h.t, n.bbox_min, n.bbox_max)) continue;
        if (!(n.prim_index & RH_LEAF_NODE_BIT)) {
            const uint32_t ax = n.prim_count >> 30, right = n.prim_count & RH_RIGHT_CHILD_BITS;
            const bool negd = d[ax] < 0;
            st.push_back(negd ? n.prim_index : right);
            st.push_back(negd ? right : n.prim_index);
            continue;
        }
        const uint32_t first = n.prim_index & RH_PRIM_INDEX_BITS;
        for (uint32_t i = first; i < first + n.prim_count; i++) {
            const uint32_t mii = sc.mi_indices[i];
            const rh_mesh_instance &mi = sc.mesh_instances[mii];
            if (cnt) cnt->instances++;
            if (!bbox_test(o, inv_d, h.t, mi.bbox_min, mi.bbox_max)) continue;
            // TransformRay (CoreRef.cpp:1605-1617)
            const float *x = sc.transforms[mi.tr_index].inv_xform;
            const float oo[3] = {x[0] * o[0] + x[4] * o[1] + x[8] * o[2] + x[12],
                                 x[1] * o[0] + x[5] * o[1] + x[9] * o[2] + x[13],
                                 x[2] * o[0] + x[6] * o[1] + x[10] * o[2] + x[14]};
            const float dd[3] = {x[0] * d[0] + x[4] * d[1] + x[8] * d[2], x[1] * d[0] + x[5] * d[1] + x[9] * d[2],
                                 x[2] * d[0] + x[6] * d[1] + x[10] * d[2]};
            const float inv2[3] = {1.0f / dd[0], 1.0f / dd[1], 1.0f / dd[2]};
            const bool found = micro(sc, oo, dd, inv2, sc.meshes[mi.mesh_index].node_index, (int)mii, any, h, cnt);
            if (any && found && (sc.tris[h.prim].ci & RH_TRI_SOLID_BIT)) return;
        }
    }
}

// ------------------------------------------------------------------ 8-wide trees
// bbox_test_oct (CoreRef.cpp:133-161): the slab test of the 8 slots, each
// slot's entry distance in dist; bit i of the result = slot i hit (an unused
// slot has NaN bounds and never passes)
int bbox_test_oct(const float o[3], const float inv_d[3], float t, const rh_mbvh_node &n, float dist[8]) {
    int mask = 0;
    for (int i = 0; i < 8; i++) {
        float lo_x = inv_d[0] * (n.bbox_min[0][i] - o[0]), hi_x = inv_d[0] * (n.bbox_max[0][i] - o[0]);
        if (lo_x > hi_x) std::swap(lo_x, hi_x);
        float lo_y = inv_d[1] * (n.bbox_min[1][i] - o[1]), hi_y = inv_d[1] * (n.bbox_max[1][i] - o[1]);
        if (lo_y > hi_y) std::swap(lo_y, hi_y);
        float lo_z = inv_d[2] * (n.bbox_min[2][i] - o[2]), hi_z = inv_d[2] * (n.bbox_max[2][i] - o[2]);
        if (lo_z > hi_z) std::swap(lo_z, hi_z);
        float tmin = lo_x > lo_y ? lo_x : lo_y;
        if (lo_z > tmin) tmin = lo_z;
        float tmax = hi_x < hi_y ? hi_x : hi_y;
        if (hi_z < tmax) tmax = hi_z;
        tmax *= 1.00000024f;
        dist[i] = tmin;
        if (tmin <= tmax && tmin <= t && tmax > 0) mask |= 1 << i;
    }
    return mask;
}

// TraversalStack (CoreRef.cpp:164-262): (node, entry distance) entries; after
// pushing the hit children of a node in slot order, its top entries are
// ordered by decreasing distance (nearest on top) by the reference's own
// exchange sequences, whose handling of equal distances this keeps.
struct WEntry {
    uint32_t index;
    float dist;
};
struct WStack {
    std::vector<WEntry> e;
    void push(uint32_t i, float d) { e.push_back({i, d}); }
    WEntry pop() {
        const WEntry x = e.back();
        e.pop_back();
        return x;
    }
    void sort_top3() {
        WEntry *s = &e[e.size() - 3];
        if (s[0].dist > s[1].dist) {
            if (s[1].dist > s[2].dist) return;
            if (s[0].dist > s[2].dist) {
                std::swap(s[1], s[2]);
            } else { // s2 >= s0 > s1
                const WEntry a = s[0];
                s[0] = s[2];
                s[2] = s[1];
                s[1] = a;
            }
        } else {
            if (s[0].dist > s[2].dist) {
                std::swap(s[0], s[1]);
            } else if (s[2].dist > s[1].dist) {
                std::swap(s[0], s[2]);
            } else { // s1 >= s2 >= s0
                const WEntry a = s[0];
                s[0] = s[1];
                s[1] = s[2];
                s[2] = a;
            }
        }
    }
    void sort_top4() {
        WEntry *s = &e[e.size() - 4];
        const int pairs[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
        for (const auto &p : pairs)
            if (s[p[0]].dist < s[p[1]].dist) std::swap(s[p[0]], s[p[1]]);
    }
    void sort_top(int n) { // insertion sort, stable: equal distances keep push order
        const size_t start = e.size() - n;
        for (size_t i = start + 1; i < e.size(); i++) {
            const WEntry key = e[i];
            size_t j = i;
            while (j > start && e[j - 1].dist < key.dist) {
                e[j] = e[j - 1];
                j--;
            }
            e[j] = key;
        }
    }
};

// The common walk of the four wide traversals (CoreRef.cpp:1133-1222,
// 1270-1368, 1395-1468, 1500-1568): a node is entered without a test of its
// own bounds; of its hit slots the nearest is entered next and the others
// are stacked nearest-on-top; a stacked entry farther than the current t is
// skipped when popped.  leaf(first, count) handles a leaf and returns true to
// end the walk.  Counts interior node visits.
template <class Leaf>
void wide_walk(const rh_scene_view &sc, uint32_t root, const float o[3], const float inv_d[3], Hit &h, orc_counters *cnt,
               Leaf leaf) {
    WStack st;
    st.push(root, 0.0f);
    while (!st.e.empty()) {
        WEntry cur = st.pop();
        if (cur.dist > h.t) continue;
        while (true) {
            const rh_mbvh_node &n = sc.mnodes[cur.index];
            if (n.child[0] & RH_LEAF_NODE_BIT) {
                if (leaf(n.child[0] & RH_PRIM_INDEX_BITS, n.child[1])) return;
                break;
            }
            if (cnt) cnt->nodes++;
            float dist[8];
            int mask = bbox_test_oct(o, inv_d, h.t, n, dist);
            if (!mask) break;
            int hit[8], k = 0;
            while (mask) {
                const int i = __builtin_ctz(mask);
                mask &= mask - 1;
                hit[k++] = i;
            }
            if (k == 1) {
                cur.index = n.child[hit[0]];
                continue;
            }
            if (k == 2) {
                const int a = hit[0], b = hit[1];
                if (dist[a] < dist[b]) {
                    st.push(n.child[b], dist[b]);
                    cur.index = n.child[a];
                } else {
                    st.push(n.child[a], dist[a]);
                    cur.index = n.child[b];
                }
                continue;
            }
            for (int j = 0; j < k; j++) st.push(n.child[hit[j]], dist[hit[j]]);
            if (k == 3) st.sort_top3();
            else if (k == 4) st.sort_top4();
            else st.sort_top(k);
            cur.index = st.pop().index;
        }
    }
}

// Traverse_MicroTree_WithStack_{ClosestHit,AnyHit}, 8-wide (CoreRef.cpp:1395-1568);
// the reference's wide walks call only the closest-hit form
bool micro_wide(const rh_scene_view &sc, const float o[3], const float d[3], const float inv_d[3], uint32_t root,
                int obj, bool any, Hit &h, orc_counters *cnt) {
    bool res = false;
    wide_walk(sc, root, o, inv_d, h, cnt, [&](uint32_t first, uint32_t count) {
        // IntersectTris_ClosestHit / _AnyHit (CoreRef.cpp:760-825)
        bool leaf_hit = false;
        for (uint32_t k = first; k < first + count; k++) {
            const uint32_t ti = sc.tri_indices[k];
            if (cnt) cnt->tris++;
            if (intersect(o, d, sc.tris[ti], ti, h)) {
                h.obj = obj;
                leaf_hit = true;
            }
            if (any && leaf_hit && (sc.tris[ti].ci & RH_TRI_SOLID_BIT)) break;
        }
        res |= leaf_hit;
        return any && leaf_hit && (sc.tris[h.prim].ci & RH_TRI_SOLID_BIT);
    });
    return res;
}

// Traverse_MacroTree_WithStack_{ClosestHit,AnyHit}, 8-wide (CoreRef.cpp:1133-1222, 1270-1368)
void macro_wide(const rh_scene_view &sc, const float o[3], const float d[3], bool any, Hit &h, orc_counters *cnt) {
    if (sc.macro_nodes_root == RH_INVALID) return;
    float inv_d[3];
    safe_invert(d, inv_d);
    wide_walk(sc, sc.macro_nodes_root, o, inv_d, h, cnt, [&](uint32_t first, uint32_t count) {
        for (uint32_t i = first; i < first + count; i++) {
            const uint32_t mii = sc.mi_indices[i];
            const rh_mesh_instance &mi = sc.mesh_instances[mii];
            if (cnt) cnt->instances++;
            if (!bbox_test(o, inv_d, h.t, mi.bbox_min, mi.bbox_max)) continue;
            const float *x = sc.transforms[mi.tr_index].inv_xform;
            const float oo[3] = {x[0] * o[0] + x[4] * o[1] + x[8] * o[2] + x[12],
                                 x[1] * o[0] + x[5] * o[1] + x[9] * o[2] + x[13],
                                 x[2] * o[0] + x[6] * o[1] + x[10] * o[2] + x[14]};
            const float dd[3] = {x[0] * d[0] + x[4] * d[1] + x[8] * d[2], x[1] * d[0] + x[5] * d[1] + x[9] * d[2],
                                 x[2] * d[0] + x[6] * d[1] + x[10] * d[2]};
            const float inv2[3] = {1.0f / dd[0], 1.0f / dd[1], 1.0f / dd[2]};
            // the wide any-hit walk also takes the closest hit inside each
            // instance (CoreRef.cpp:1357), then stops at a solid one
            const bool found =
                micro_wide(sc, oo, dd, inv2, sc.meshes[mi.mesh_index].node_index, (int)mii, false, h, cnt);
            if (any && found && (sc.tris[h.prim].ci & RH_TRI_SOLID_BIT)) return true;
        }
        return false;
    });
}

// the scene's trees: 8-wide when it has them
void macro(const rh_scene_view &sc, const float o[3], const float d[3], bool any, Hit &h, orc_counters *cnt) {
    if (sc.mnodes) macro_wide(sc, o, d, any, h, cnt);
    else macro_binary(sc, o, d, any, h, cnt);
}

// The light-tree leaves (first, count) whose bounds hold P strictly, in the
// order ComputeDirectLighting visits them (CoreRef.cpp:1935-1991): the binary
// tree tests every popped node and pushes left then right; the wide tree tests
// the slots of an interior node (not the root's own bounds) and pushes the
// passing ones in slot order.
void light_leaves(const rh_scene_view &sc, V3 P, std::vector<std::pair<uint32_t, uint32_t>> &out) {
    out.clear();
    if (sc.light_nodes_root == RH_INVALID) return;
    auto inside = [&](const float *mn, const float *mx) {
        return P.x > mn[0] && P.x < mx[0] && P.y > mn[1] && P.y < mx[1] && P.z > mn[2] && P.z < mx[2];
    };
    std::vector<uint32_t> st{sc.light_nodes_root};
    while (!st.empty()) {
        const uint32_t cur = st.back();
        st.pop_back();
        if (sc.mnodes) {
            const rh_mbvh_node &n = sc.mnodes[cur];
            if (n.child[0] & RH_LEAF_NODE_BIT) {
                out.push_back({n.child[0] & RH_PRIM_INDEX_BITS, n.child[1]});
                continue;
            }
            for (int i = 0; i < 8; i++) {
                const float mn[3] = {n.bbox_min[0][i], n.bbox_min[1][i], n.bbox_min[2][i]};
                const float mx[3] = {n.bbox_max[0][i], n.bbox_max[1][i], n.bbox_max[2][i]};
                if (inside(mn, mx)) st.push_back(n.child[i]);
            }
            continue;
        }
        const rh_bvh_node &n = sc.nodes[cur];
        if (!inside(n.bbox_min, n.bbox_max)) continue;
        if (!(n.prim_index & RH_LEAF_NODE_BIT)) {
            st.push_back(n.prim_index);
            st.push_back(n.prim_count & RH_RIGHT_CHILD_BITS);
            continue;
        }
        out.push_back({n.prim_index & RH_PRIM_INDEX_BITS, n.prim_count});
    }
}

// ------------------------------------------------------------------ textures
uint32_t texel_index(const rh_scene_view &sc, uint32_t tex, int lod, int x, int y) {
    const rh_texture &t = sc.textures[tex];
    const rh_tex_mip &m = sc.tex_mips[tex * RH_NUM_MIP_LEVELS + lod];
    const int lx = clampi(x - (int)t.pos[lod][0], 0, m.bw - 1), ly = clampi(y - (int)t.pos[lod][1], 0, m.bh - 1);
    return m.offset + (uint32_t)ly * m.bw + (uint32_t)lx;
}
uint32_t texel(const rh_scene_view &sc, uint32_t tex, int lod, int x, int y) {
    return sc.texels[texel_index(sc, tex, lod, x, y)];
}

// the stored texels a bilinear sample read (00, 01, 10, 11) and its weights
struct Foot {
    uint32_t idx[4];
    float kx, ky;
};
inline float ch(uint32_t p, int c) { return (float)((p >> (8 * c)) & 0xff); }
// generic simd_vec floor (simd_vec.h:408-412)
inline float vfloor(float v) { return (float)((int)v - (v < 0.0f)); }

// SampleBilinear + TransformUV (CoreRef.cpp:1625-1678)
void bilinear(const rh_scene_view &sc, uint32_t tex, float u, float v, int lod, float out[4], Foot *foot = nullptr) {
    const rh_texture &t = sc.textures[tex];
    const float A = (float)RH_TEXTURE_ATLAS_SIZE;
    const float sx = (float)((t.width & RH_TEXTURE_WIDTH_BITS) >> lod), sy = (float)(t.height >> lod);
    float x = (float)t.pos[lod][0] + (u - vfloor(u)) * sx + 1.0f, y = (float)t.pos[lod][1] + (v - vfloor(v)) * sy + 1.0f;
    x /= A;
    y /= A;
    x = x * A - 0.5f;
    y = y * A - 0.5f;
    const uint32_t p00 = texel(sc, tex, lod, (int)x, (int)y), p01 = texel(sc, tex, lod, (int)x + 1, (int)y);
    const uint32_t p10 = texel(sc, tex, lod, (int)x, (int)y + 1), p11 = texel(sc, tex, lod, (int)x + 1, (int)y + 1);
    const float kx = x - std::floor(x), ky = y - std::floor(y);
    if (foot) {
        *foot = Foot{{texel_index(sc, tex, lod, (int)x, (int)y), texel_index(sc, tex, lod, (int)x + 1, (int)y),
                      texel_index(sc, tex, lod, (int)x, (int)y + 1), texel_index(sc, tex, lod, (int)x + 1, (int)y + 1)},
                     kx, ky};
    }
    for (int c = 0; c < 4; c++) {
        const float a = ch(p01, c) * kx + ch(p00, c) * (1 - kx), b = ch(p11, c) * kx + ch(p10, c) * (1 - kx);
        out[c] = (b * ky + a * (1.0f - ky)) * (1.0f / 255.0f);
    }
}

// SampleLatlong_RGBE (CoreRef.cpp:1772-1799)
void latlong(const rh_scene_view &sc, uint32_t tex, V3 dir, float out[4]) {
    const rh_texture &t = sc.textures[tex];
    const float theta = std::acos(clampf(dir.y, -1.0f, 1.0f)) / RH_PI;
    const float r = std::sqrt(dir.x * dir.x + dir.z * dir.z);
    float u = 0.5f * std::acos(r > RH_FLT_EPS ? clampf(dir.x / r, -1.0f, 1.0f) : 0.0f) / RH_PI;
    if (dir.z < 0.0f) u = 1.0f - u;
    const float x = (float)t.pos[0][0] + u * (float)(t.width & RH_TEXTURE_WIDTH_BITS) + 1.0f;
    const float y = (float)t.pos[0][1] + theta * (float)t.height + 1.0f;
    const uint32_t p[4] = {texel(sc, tex, 0, (int)(x + 0), (int)(y + 0)), texel(sc, tex, 0, (int)(x + 1), (int)(y + 0)),
                           texel(sc, tex, 0, (int)(x + 0), (int)(y + 1)), texel(sc, tex, 0, (int)(x + 1), (int)(y + 1))};
    float c[4][4];
    for (int k = 0; k < 4; k++) {
        const float f = std::exp2(float((p[k] >> 24) & 0xff) - 128.0f);
        c[k][0] = to_norm_float(p[k] & 0xff) * f;
        c[k][1] = to_norm_float((p[k] >> 8) & 0xff) * f;
        c[k][2] = to_norm_float((p[k] >> 16) & 0xff) * f;
        c[k][3] = 1.0f;
    }
    const float kx = x - vfloor(x), ky = y - vfloor(y);
    for (int i = 0; i < 4; i++) {
        const float a = c[1][i] * kx + c[0][i] * (1 - kx), b = c[3][i] * kx + c[2][i] * (1 - kx);
        out[i] = b * ky + a * (1 - ky);
    }
}

// fast_log2, get_texture_lod (CoreRef.cpp:391-415)
float fast_log2(float val) {
    int32_t x;
    memcpy(&x, &val, 4);
    float l = (float)(((x >> 23) & 255) - 128);
    x &= ~(255 << 23);
    x += 127 << 23;
    float v;
    memcpy(&v, &x, 4);
    l += ((-0.34484843f) * v + 2.02466578f) * v - 0.67487759f;
    return l;
}
float texture_lod(const rh_texture &t, const float dx_[2], const float dy_[2]) {
    const float sx = (float)(t.width & RH_TEXTURE_WIDTH_BITS), sy = (float)t.height;
    const float dx[2] = {dx_[0] * sx, dx_[1] * sy}, dy[2] = {dy_[0] * sx, dy_[1] * sy};
    const float dg[2] = {dx[0] + dy[0], dx[1] + dy[1]};
    float a = 0, b = 0, c = 0;
    a += dx[0] * dx[0];
    a += dx[1] * dx[1];
    b += dy[0] * dy[0];
    b += dy[1] * dy[1];
    c += dg[0] * dg[0];
    c += dg[1] * dg[1];
    const float m = std::min(std::min(a, b), c);
    return clampf(0.5f * fast_log2(m) - 1.0f, 0.0f, (float)RH_MAX_MIP_LEVEL);
}

// ------------------------------------------------------------------ shading helpers
V3 plane_normal(const rh_tri_accel &tri) { // ExtractPlaneNormal (Core.cpp:216-228)
    static const int nu_[] = {1, 0, 0}, nv_[] = {2, 2, 1};
    float n[3];
    const int iw = tri.ci & RH_TRI_W_BITS;
    n[iw] = 1.0f;
    n[nu_[iw]] = tri.nu;
    n[nv_[iw]] = tri.nv;
    float l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (tri.ci & RH_TRI_INV_NORMAL_BIT) l = -l;
    return {n[0] * l, n[1] * l, n[2] * l};
}
V3 xform_normal(V3 n, const float *x) { // TransformNormal (CoreRef.cpp:1625-1629)
    return {x[0] * n.x + x[1] * n.y + x[2] * n.z, x[4] * n.x + x[5] * n.y + x[6] * n.z, x[8] * n.x + x[9] * n.y + x[10] * n.z};
}
float oren_nayar(V3 L, V3 I, V3 N, V3 T, float sigma) { // BRDF_OrenNayar (CoreRef.cpp:1579-1603)
    const float s2 = sigma * sigma;
    const float A = 1.0f - (s2 / (2.0f * (s2 + 0.33f)));
    const float B = (0.45f * s2) / (s2 + 0.09f);
    const float cti = dot(L, N), cto = -dot(I, N);
    const float sti = std::sqrt(std::max(0.0f, 1.0f - (cti * cti))), sto = std::sqrt(std::max(0.0f, 1.0f - (cto * cto)));
    const float cpi = dot(L, T), cpo = -dot(I, T);
    const float spi = std::sqrt(std::max(0.0f, 1.0f - (cpi * cpi))), spo = std::sqrt(std::max(0.0f, 1.0f - (cpo * cpo)));
    const float temp = std::max(0.0f, (cpi * cpo) + (spi * spo));
    float sa, tb;
    if (cti > cto) {
        sa = sto;
        tb = (sti / cti);
    } else {
        sa = sti;
        tb = (sto / cto);
    }
    return clampf((A + (B * temp * sa * tb)), 0.0f, 1.0f);
}

struct Der {
    V3 do_dx, do_dy, dd_dx, dd_dy;
    float duv_dx[2], duv_dy[2];
    V3 dndx, dndy;
    float ddn_dx, ddn_dy;
};
// ComputeDerivatives (CoreRef.cpp:1993-2052)
void derivatives(V3 I, float t, V3 do_dx, V3 do_dy, V3 dd_dx, V3 dd_dy, const rh_vertex &w1, const rh_vertex &w2,
                 const rh_vertex &w3, V3 pN, Der &o) {
    const float dIN = -dot(I, pN);
    const float inv_dot = std::abs(dIN) < RH_FLT_EPS ? 0.0f : 1.0f / dIN;
    const float dt_dx = dot(add(do_dx, mul(t, dd_dx)), pN) * inv_dot, dt_dy = dot(add(do_dy, mul(t, dd_dy)), pN) * inv_dot;
    o.do_dx = add(add(do_dx, mul(t, dd_dx)), mul(dt_dx, I));
    o.do_dy = add(add(do_dy, mul(t, dd_dy)), mul(dt_dy, I));
    o.dd_dx = dd_dx;
    o.dd_dy = dd_dy;
    const float duv13[2] = {w1.t[0][0] - w3.t[0][0], w1.t[0][1] - w3.t[0][1]};
    const float duv23[2] = {w2.t[0][0] - w3.t[0][0], w2.t[0][1] - w3.t[0][1]};
    const V3 dp13 = sub(v3(w1.p), v3(w3.p)), dp23 = sub(v3(w2.p), v3(w3.p));
    const float det_uv = duv13[0] * duv23[1] - duv13[1] * duv23[0];
    const float idu = std::abs(det_uv) < RH_FLT_EPS ? 0 : 1.0f / det_uv;
    const V3 dpdu = mul(sub(mul(duv23[1], dp13), mul(duv13[1], dp23)), idu);
    const V3 dpdv = mul(add(mul(-duv23[0], dp13), mul(duv13[0], dp23)), idu);
    float A[2][2] = {{dpdu.x, dpdv.x}, {dpdu.y, dpdv.y}};
    float Bx[2] = {o.do_dx.x, o.do_dx.y}, By[2] = {o.do_dy.x, o.do_dy.y};
    if (std::abs(pN.x) > std::abs(pN.y) && std::abs(pN.x) > std::abs(pN.z)) {
        A[0][0] = dpdu.y;
        A[0][1] = dpdv.y;
        A[1][0] = dpdu.z;
        A[1][1] = dpdv.z;
        Bx[0] = o.do_dx.y;
        Bx[1] = o.do_dx.z;
        By[0] = o.do_dy.y;
        By[1] = o.do_dy.z;
    } else if (std::abs(pN.y) > std::abs(pN.z)) {
        A[1][0] = dpdu.z;
        A[1][1] = dpdv.z;
        Bx[1] = o.do_dx.z;
        By[1] = o.do_dy.z;
    }
    const float det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
    const float id = std::abs(det) < RH_FLT_EPS ? 0 : 1.0f / det;
    o.duv_dx[0] = (A[1][1] * Bx[0] - A[0][1] * Bx[1]) * id;
    o.duv_dx[1] = (A[0][0] * Bx[1] - A[1][0] * Bx[0]) * id;
    o.duv_dy[0] = (A[1][1] * By[0] - A[0][1] * By[1]) * id;
    o.duv_dy[1] = (A[0][0] * By[1] - A[1][0] * By[0]) * id;
    const V3 dn1 = sub(v3(w1.n), v3(w3.n)), dn2 = sub(v3(w2.n), v3(w3.n));
    const V3 dndu = mul(sub(mul(duv23[1], dn1), mul(duv13[1], dn2)), idu);
    const V3 dndv = mul(add(mul(-duv23[0], dn1), mul(duv13[0], dn2)), idu);
    o.dndx = add(mul(dndu, o.duv_dx[0]), mul(dndv, o.duv_dx[1]));
    o.dndy = add(mul(dndu, o.duv_dy[0]), mul(dndv, o.duv_dy[1]));
    o.ddn_dx = dot(dd_dx, pN) + dot(I, o.dndx);
    o.ddn_dy = dot(dd_dy, pN) + dot(I, o.dndy);
}

struct Ctx {
    const rh_scene_view &sc;
    const float *halton;
    unsigned long long shadow = 0;
};

// ComputeVisibility (CoreRef.cpp:1801-1896)
float visibility(Ctx &c, V3 p1, V3 p2, int hi, int rand_hash2) {
    const rh_scene_view &sc = c.sc;
    V3 dir = sub(p2, p1);
    float dist = len(dir);
    dir = divs(dir, dist);
    float o[3] = {p1.x, p1.y, p1.z};
    const float d[3] = {dir.x, dir.y, dir.z};
    float vis = 1.0f;
    while (dist > RH_HIT_EPS) {
        Hit h;
        h.t = dist;
        c.shadow++;
        macro(sc, o, d, true, h, nullptr);
        if (!h.mask) break;
        const rh_tri_accel &tri = sc.tris[h.prim];
        if (tri.ci & RH_TRI_SOLID_BIT) {
            vis = 0;
            break;
        }
        const rh_material *mat = &sc.materials[tri.mi];
        const float *ix = sc.transforms[sc.mesh_instances[h.obj].tr_index].inv_xform;
        const rh_vertex &a = sc.vertices[sc.vtx_indices[h.prim * 3 + 0]], &b = sc.vertices[sc.vtx_indices[h.prim * 3 + 1]],
                        &e = sc.vertices[sc.vtx_indices[h.prim * 3 + 2]];
        const float w = 1.0f - h.u - h.v;
        const float su = a.t[0][0] * w + b.t[0][0] * h.u + e.t[0][0] * h.v, sv = a.t[0][1] * w + b.t[0][1] * h.u + e.t[0][1] * h.v;
        V3 pN = xform_normal(plane_normal(tri), ix);
        bool skip = false;
        if (dot(pN, dir) < 0.0f) {
            if (tri.back_mi == RH_INVALID) skip = true;
            else {
                mat = &sc.materials[tri.back_mi];
                pN = neg(pN);
            }
        }
        if (!skip) {
            float r = fracpart(c.halton[hi] + construct_float((uint32_t)hash(rand_hash2)));
            while (mat->type == RH_MAT_MIX) {
                float tx[4];
                bilinear(sc, mat->textures[RH_MAIN_TEXTURE], su, sv, 0, tx);
                const float mv = tx[0] * mat->roughness;
                if (r > mv) {
                    mat = &sc.materials[mat->textures[RH_MIX_MAT1]];
                    r = (r - mv) / (1.0f - mv);
                } else {
                    mat = &sc.materials[mat->textures[RH_MIX_MAT2]];
                    r = r / mv;
                }
            }
            if (mat->type != RH_MAT_TRANSPARENT) {
                vis = 0;
                break;
            }
        }
        const float t = h.t + RH_HIT_BIAS;
        o[0] += d[0] * t;
        o[1] += d[1] * t;
        o[2] += d[2] * t;
        dist -= t;
    }
    return vis;
}

// ComputeDirectLighting + AcumulateLightContribution (CoreRef.cpp:1898-1991)
V3 direct_light(Ctx &c, V3 I, V3 P, V3 N, V3 B, V3 pN, float sigma, int hi, int rh2, float ro1, float ro2) {
    const rh_scene_view &sc = c.sc;
    V3 col = {0, 0, 0};
    std::vector<std::pair<uint32_t, uint32_t>> leaves;
    light_leaves(sc, P, leaves);
    for (const auto &leaf : leaves) {
        const uint32_t first = leaf.first;
        for (uint32_t i = first; i < first + leaf.second; i++) {
            const rh_light &l = sc.lights[sc.li_indices[i]];
            V3 L = sub(P, v3(l.pos));
            const float distance = len(L);
            const float dd = std::max(distance - l.radius, 0.0f);
            L = divs(L, distance);
            const float z = fracpart(c.halton[hi + 0] + ro1);
            const float dir = std::sqrt(z);
            const float phi = 2 * RH_PI * fracpart(c.halton[hi + 1] + ro2);
            const float cp = std::cos(phi), sp = std::sin(phi);
            const V3 TT = cross(L, B), BB = cross(L, TT);
            const V3 V = add(add(mul(dir * sp, BB), mul(std::sqrt(1.0f - dir), L)), mul(dir * cp, TT));
            L = norm(sub(add(v3(l.pos), mul(V, l.radius)), P));
            const float denom = dd / l.radius + 1.0f;
            float atten = 1.0f / (denom * denom);
            atten = (atten - RH_LIGHT_ATTEN_CUTOFF / l.brightness) / (1.0f - RH_LIGHT_ATTEN_CUTOFF);
            atten = std::max(atten, 0.0f);
            const float d1 = std::max(dot(L, N), 0.0f), d2 = dot(L, v3(l.dir));
            if (d1 > RH_FLT_EPS && d2 > l.spot && (l.brightness * atten) > RH_FLT_EPS) {
                const float vis = visibility(c, add(P, mul(RH_HIT_BIAS, pN)), v3(l.pos), hi, rh2);
                col = add(col, mul(mul(mul(mul(v3(l.col), d1), vis), atten), oren_nayar(L, I, N, B, sigma)));
            }
        }
    }
    return col;
}

inline float pow5(float v) { return (v * v) * (v * v) * v; }

// ShadeSurface (CoreRef.cpp:2055-2391)
// alb_mat (optional): the material whose main_color scales this vertex's
// colour and the rest of the path once (emissive; diffuse and glossy when the
// albedo is considered), or -1
struct AlbedoSample {
    Foot foot;
    float raw[3]; // before the sRGB conversion
    bool srgb;
};

void shade(Ctx &c, const rh_pass_info &pi, const rh_hit &inter, const rh_ray_packet &ray, float out[4],
           rh_ray_packet *sec, bool &emitted, int *alb_mat = nullptr, AlbedoSample *alb_sample = nullptr) {
    const rh_scene_view &sc = c.sc;
    const float *halton = c.halton;
    emitted = false;
    if (alb_mat) *alb_mat = -1;
    const uint32_t fl = pi.settings.flags;
    if (!inter.mask) {
        float env[4] = {0, 0, 0, 0};
        if (!(fl & RH_NO_BACKGROUND) || pi.bounce > 2) {
            latlong(sc, sc.env->env_map, v3(ray.d), env);
            if (sc.env->env_clamp > RH_FLT_EPS)
                for (float &e : env) e = e < sc.env->env_clamp ? e : sc.env->env_clamp;
            env[3] = 1.0f;
        }
        for (int i = 0; i < 3; i++) out[i] = ray.c[i] * env[i] * sc.env->env_col[i];
        out[3] = env[3];
        return;
    }
    const V3 I = v3(ray.d), P = add(v3(ray.o), mul(inter.t, I));
    const rh_tri_accel &tri = sc.tris[inter.prim];
    const rh_material *mat = &sc.materials[tri.mi];
    const float *ix = sc.transforms[sc.mesh_instances[inter.obj].tr_index].inv_xform;
    const rh_vertex &v1 = sc.vertices[sc.vtx_indices[inter.prim * 3 + 0]], &v2 = sc.vertices[sc.vtx_indices[inter.prim * 3 + 1]],
                    &v3_ = sc.vertices[sc.vtx_indices[inter.prim * 3 + 2]];
    const float w = 1.0f - inter.u - inter.v;
    V3 N = add(add(mul(v3(v1.n), w), mul(v3(v2.n), inter.u)), mul(v3(v3_.n), inter.v));
    const float uvs[2] = {v1.t[0][0] * w + v2.t[0][0] * inter.u + v3_.t[0][0] * inter.v,
                          v1.t[0][1] * w + v2.t[0][1] * inter.u + v3_.t[0][1] * inter.v};
    V3 pN = xform_normal(plane_normal(tri), ix);
    const bool back = dot(pN, I) > 0.0f;
    if (back) {
        if (tri.back_mi == RH_INVALID) {
            out[0] = out[1] = out[2] = out[3] = 0.0f;
            return;
        }
        mat = &sc.materials[tri.back_mi];
        pN = neg(pN);
        N = neg(N);
    }
    Der der;
    derivatives(I, inter.t, v3(ray.do_dx), v3(ray.do_dy), v3(ray.dd_dx), v3(ray.dd_dy), v1, v2, v3_, pN, der);

    const int h1 = hash(pi.rand_index), h2 = hash(h1), h3 = hash(h2);
    const float ro1 = construct_float((uint32_t)h1), ro2 = construct_float((uint32_t)h2), ro3 = construct_float((uint32_t)h3);
    const int hi = (pi.iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT + pi.bounce * 2;
    float mix_rand = fracpart(halton[hi] + ro1);
    while (mat->type == RH_MAT_MIX) {
        float tx[4];
        bilinear(sc, mat->textures[RH_MAIN_TEXTURE], uvs[0], uvs[1], 0, tx);
        float mv = tx[0] * mat->roughness;
        const float ior = back ? mat->ext_ior : mat->int_ior;
        float R0 = (ray.ior - ior) / (ray.ior + ior);
        R0 *= R0;
        float RR;
        if (ray.ior > ior) {
            const float eta = ray.ior / ior, cosi = -dot(I, N), cost2 = 1.0f - eta * eta * (1.0f - cosi * cosi);
            if (cost2 >= 0.0f) {
                const float m = eta * cosi - std::sqrt(cost2);
                RR = R0 + (1.0f - R0) * pow5(1.0f + dot(add(mul(eta, I), mul(m, N)), N));
            } else {
                RR = 1.0f;
            }
        } else {
            RR = R0 + (1.0f - R0) * pow5(1.0f + dot(I, N));
        }
        mv *= clampf(RR, 0.0f, 1.0f);
        if (mix_rand > mv) {
            mat = &sc.materials[mat->textures[RH_MIX_MAT1]];
            mix_rand = (mix_rand - mv) / (1.0f - mv);
        } else {
            mat = &sc.materials[mat->textures[RH_MIX_MAT2]];
            mix_rand = mix_rand / mv;
        }
    }
    V3 B = add(add(mul(v3(v1.b), w), mul(v3(v2.b), inter.u)), mul(v3(v3_.b), inter.v));
    V3 T = cross(B, N);
    float nm[4];
    bilinear(sc, mat->textures[RH_NORMALS_TEXTURE], uvs[0], uvs[1], 0, nm);
    for (float &x : nm) x = x * 2.0f - 1.0f;
    N = add(add(mul(nm[0], B), mul(nm[2], N)), mul(nm[1], T));
    N = xform_normal(N, ix);
    B = xform_normal(B, ix);
    T = xform_normal(T, ix);

    const uint32_t mt = mat->textures[RH_MAIN_TEXTURE];
    const float lod = texture_lod(sc.textures[mt], der.duv_dx, der.duv_dy);
    float alb[4];
    bilinear(sc, mt, uvs[0], uvs[1], (int)lod, alb, alb_sample ? &alb_sample->foot : nullptr);
    if (alb_sample) {
        for (int i = 0; i < 3; i++) alb_sample->raw[i] = alb[i];
        alb_sample->srgb = (sc.textures[mt].width & RH_TEXTURE_SRGB_BIT) != 0;
    }
    if (sc.textures[mt].width & RH_TEXTURE_SRGB_BIT) {
        for (int i = 0; i < 3; i++) alb[i] = alb[i] > 0.04045f ? std::pow((alb[i] + 0.055f) / 1.055f, 2.4f) : alb[i] / 12.92f;
    }
    for (int i = 0; i < 3; i++) alb[i] *= mat->main_color[i];
    const V3 A = {alb[0], alb[1], alb[2]};

    V3 col = {0, 0, 0};
    const int dd = ray.ray_depth & 0xff, gd = (ray.ray_depth >> 8) & 0xff, rd = (ray.ray_depth >> 16) & 0xff,
              td = (ray.ray_depth >> 24) & 0xff, tot = dd + gd + rd + td;
    const bool cant_terminate = tot < pi.settings.termination_start_depth;
    const bool consider_albedo = !(fl & RH_LIGHTING_ONLY) || pi.bounce > 2;
    if (alb_mat && (mat->type == RH_MAT_EMISSIVE ||
                    (consider_albedo && (mat->type == RH_MAT_DIFFUSE || mat->type == RH_MAT_GLOSSY))))
        *alb_mat = (int)(mat - sc.materials);
    rh_ray_packet r;
    memset(&r, 0, sizeof(r));
    bool emit = false;
    auto set_common = [&](V3 o, V3 dir, V3 cc, float ior, int depth) {
        r.xy = ray.xy;
        r.ior = ior;
        r.ray_depth = depth;
        memcpy(r.o, &o, 12);
        memcpy(r.d, &dir, 12);
        memcpy(r.c, &cc, 12);
    };
    auto reflect_diffs = [&]() {
        const float dIN = dot(I, pN);
        const V3 a = sub(der.dd_dx, mul(2, add(mul(dIN, der.dndx), mul(der.ddn_dx, pN))));
        const V3 b = sub(der.dd_dy, mul(2, add(mul(dIN, der.dndy), mul(der.ddn_dy, pN))));
        memcpy(r.do_dx, &der.do_dx, 12);
        memcpy(r.do_dy, &der.do_dy, 12);
        memcpy(r.dd_dx, &a, 12);
        memcpy(r.dd_dy, &b, 12);
    };
    if (mat->type == RH_MAT_DIFFUSE) {
        if (!(fl & RH_SKIP_DIRECT_LIGHT) || pi.bounce > 2) {
            col = direct_light(c, I, P, N, B, pN, mat->roughness, hi, h2, ro1, ro2);
            if (consider_albedo) col = mulv(col, A);
        }
        if (dd < pi.settings.max_diff_depth && tot < pi.settings.max_total_depth) {
            const float u1 = fracpart(halton[hi + 0] + ro1), u2 = fracpart(halton[hi + 1] + ro2);
            const float phi = 2 * RH_PI * u2, cp = std::cos(phi), sp = std::sin(phi);
            V3 V;
            float weight = 1.0f;
            if ((fl & RH_OUTPUT_SH) && pi.bounce <= 2) {
                const float dir = std::sqrt(1.0f - u1 * u1);
                V = add(add(mul(dir * sp, B), mul(u1, N)), mul(dir * cp, T));
                weight = 2 * u1;
            } else {
                const float dir = std::sqrt(u1);
                V = add(add(mul(dir * sp, B), mul(std::sqrt(1.0f - u1), N)), mul(dir * cp, T));
            }
            weight *= oren_nayar(V, I, N, B, mat->roughness);
            V3 cc = {ray.c[0] * weight, ray.c[1] * weight, ray.c[2] * weight};
            if (consider_albedo) cc = mulv(cc, A);
            set_common(add(P, mul(RH_HIT_BIAS, pN)), V, cc, ray.ior, ray.ray_depth + 0x00000001);
            reflect_diffs();
            emit = true;
        }
    } else if (mat->type == RH_MAT_GLOSSY) {
        if (gd < pi.settings.max_glossy_depth && tot < pi.settings.max_total_depth) {
            V3 V = sub(I, mul(2 * dot(N, I), N));
            const float h = 1.0f - std::cos(0.5f * RH_PI * mat->roughness * mat->roughness);
            const float z = h * fracpart(halton[hi + 0] + ro1);
            const float dir = std::sqrt(z), phi = 2 * RH_PI * fracpart(halton[hi + 1] + ro2);
            const float cp = std::cos(phi), sp = std::sin(phi);
            const V3 TT = cross(V, B), BB = cross(V, TT);
            if (dot(V, pN) > 0) V = add(add(mul(dir * sp, BB), mul(std::sqrt(1.0f - dir), V)), mul(dir * cp, TT));
            else V = sub(add(mul(-dir * sp, BB), mul(std::sqrt(1.0f - dir), V)), mul(dir * cp, TT));
            V3 cc = v3(ray.c);
            if (consider_albedo) cc = mulv(cc, A);
            set_common(add(P, mul(RH_HIT_BIAS, pN)), V, cc, ray.ior, ray.ray_depth + 0x00000100);
            reflect_diffs();
            emit = true;
        }
    } else if (mat->type == RH_MAT_REFRACTIVE) {
        if (rd < pi.settings.max_refr_depth && tot < pi.settings.max_total_depth) {
            const float ior = back ? mat->ext_ior : mat->int_ior;
            const float eta = ray.ior / ior, cosi = -dot(I, N), cost2 = 1.0f - eta * eta * (1.0f - cosi * cosi);
            if (cost2 < 0) {
                out[0] = out[1] = out[2] = 0.0f;
                out[3] = 1.0f;
                return;
            }
            const float m = eta * cosi - std::sqrt(cost2);
            V3 V = add(mul(eta, I), mul(m, N));
            const float z = 1.0f - halton[hi + 0] * mat->roughness, temp = std::sqrt(1.0f - z * z);
            const float phi = halton[(((hash(hi) + pi.iteration) & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT + pi.bounce * 2) + 0] * 2 * RH_PI;
            const float cp = std::cos(phi), sp = std::sin(phi);
            const V3 TT = norm(cross(V, B)), BB = norm(cross(V, TT));
            V = add(add(mul(temp * sp, BB), mul(z, V)), mul(temp * cp, TT));
            const float k = (eta - eta * eta * dot(I, pN) / dot(V, pN));
            const float dmdx = k * der.ddn_dx, dmdy = k * der.ddn_dy;
            set_common(add(P, mul(RH_HIT_BIAS, I)), V, mul(v3(ray.c), z), ior, ray.ray_depth + 0x00010000);
            const V3 a = sub(mul(eta, der.dd_dx), add(mul(m, der.dndx), mul(dmdx, pN)));
            const V3 b = sub(mul(eta, der.dd_dy), add(mul(m, der.dndy), mul(dmdy, pN)));
            memcpy(r.do_dx, &der.do_dx, 12);
            memcpy(r.do_dy, &der.do_dy, 12);
            memcpy(r.dd_dx, &a, 12);
            memcpy(r.dd_dy, &b, 12);
            emit = true;
        }
    } else if (mat->type == RH_MAT_EMISSIVE) {
        col = mul(mat->roughness, A);
    } else if (mat->type == RH_MAT_TRANSPARENT) {
        if (td < pi.settings.max_transp_depth && tot < pi.settings.max_total_depth) {
            set_common(add(P, mul(RH_HIT_BIAS, I)), I, v3(ray.c), ray.ior, ray.ray_depth + 0x01000000);
            memcpy(r.do_dx, ray.do_dx, 12);
            memcpy(r.do_dy, ray.do_dy, 12);
            memcpy(r.dd_dx, ray.dd_dx, 12);
            memcpy(r.dd_dy, ray.dd_dy, 12);
            emit = true;
        }
    }
    if (emit) {
        const float lum = std::max(r.c[0], std::max(r.c[1], r.c[2]));
        const float p = fracpart(halton[hi + 0] + ro3);
        const float q = cant_terminate ? 0.0f : std::max(0.05f, 1.0f - lum);
        if (p >= q) {
            for (float &x : r.c) x /= 1.0f - q;
            *sec = r;
            emitted = true;
        }
    }
    out[0] = ray.c[0] * col.x;
    out[1] = ray.c[1] * col.y;
    out[2] = ray.c[2] * col.z;
    out[3] = 1.0f;
}

// Ray-hash tables (Core.cpp:129-150)
const int8_t kOmega[33] = {15, 14, 13, 12, 12, 11, 11, 11, 10, 10, 9, 9, 9, 8, 8, 8, 8, 7, 7, 7, 6, 6, 6, 5, 5, 4, 4, 4, 3, 3, 2, 1, 0};
const int8_t kPhi[17][17] = {{2, 2, 2, 2, 2, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 6},
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
uint32_t morton(uint32_t v, int stride, int bits) {
    uint32_t r = 0;
    for (int k = 0; k < bits; k++) r |= ((v >> k) & 1u) << (stride * k);
    return r;
}

// get_ray_hash (CoreRef.cpp:316-333)
uint32_t ray_hash(const rh_ray_packet &r, const float *rmin, const float *cell) {
    const int x = (int)morton(clampi((int)((r.o[0] - rmin[0]) / cell[0]), 0, 255), 3, 8);
    const int y = (int)morton(clampi((int)((r.o[1] - rmin[1]) / cell[1]), 0, 255), 3, 8);
    const int z = (int)morton(clampi((int)((r.o[2] - rmin[2]) / cell[2]), 0, 255), 3, 8);
    const int o = (int)morton(kOmega[clampi(int((1.0f + r.d[2]) / 0.0625f), 0, 32)], 2, 4);
    const int p = (int)morton(kPhi[clampi(int((1.0f + r.d[1]) / 0.125f), 0, 16)][clampi(int((1.0f + r.d[0]) / 0.125f), 0, 16)], 2, 4);
    return (uint32_t)((o << 25) | (p << 24) | (y << 2) | (z << 1) | (x << 0));
}

} // namespace

extern "C" {

int orc_halton_perms(uint16_t *out, size_t count) {
    const std::vector<uint16_t> p = perms();
    if (count < p.size()) return -1;
    memcpy(out, p.data(), p.size() * 2);
    return (int)p.size();
}

int orc_halton_table(int iteration, float *out) {
    static const std::vector<uint16_t> p = perms();
    halton_table(p, iteration, out);
    return 0;
}

// GeneratePrimaryRays (CoreRef.cpp:427-505)
int orc_gen_primary(const rh_camera *cam, const int32_t rect[4], int w, int h, int iteration, const float *halton,
                    rh_ray_packet *out) {
    const V3 co = v3(cam->origin), fwd = v3(cam->fwd), side = v3(cam->side), up = v3(cam->up);
    const float fd = cam->focus_distance;
    const float k = float(w) / h;
    const float fov_k = std::tan(0.5f * cam->fov * RH_PI / 180.0f) * fd;
    auto pix_dir = [&](float x, float y, V3 origin) {
        const V3 p = {2 * fov_k * x / w - fov_k, 2 * fov_k * (-y) / h + fov_k, fd};
        const V3 q = add(add(add(co, mul(k * p.x, side)), mul(p.y, up)), mul(p.z, fwd));
        return norm(sub(q, origin));
    };
    size_t i = 0;
    for (int y = rect[1]; y < rect[1] + rect[3]; y++) {
        for (int x = rect[0]; x < rect[0] + rect[2]; x++) {
            rh_ray_packet &r = out[i++];
            const int index = y * w + x;
            const int hi = (iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT;
            float fx = (float)x, fy = (float)y;
            const int hv = hash(index);
            if (cam->filter == RH_FILTER_TENT) {
                float rx = fracpart(halton[hi + 0] + construct_float((uint32_t)hv));
                rx = rx < 0.5f ? std::sqrt(2.0f * rx) - 1.0f : 1.0f - std::sqrt(2.0f - 2 * rx);
                float ry = fracpart(halton[hi + 1] + construct_float((uint32_t)hash(hv)));
                ry = ry < 0.5f ? std::sqrt(2.0f * ry) - 1.0f : 1.0f - std::sqrt(2.0f - 2.0f * ry);
                fx += 0.5f + rx;
                fy += 0.5f + ry;
            } else {
                fx += fracpart(halton[hi + 0] + construct_float((uint32_t)hv));
                fy += fracpart(halton[hi + 1] + construct_float((uint32_t)hash(hv)));
            }
            const float ff1 = cam->focus_factor * (-0.5f + fracpart(halton[hi + 2] + construct_float((uint32_t)hv)));
            const float ff2 = cam->focus_factor * (-0.5f + fracpart(halton[hi + 3] + construct_float((uint32_t)hash(hv))));
            const V3 o = add(add(co, mul(side, ff1)), mul(up, ff2));
            const V3 d = pix_dir(fx, fy, o), dx = pix_dir(fx + 1, fy, o), dy = pix_dir(fx, fy + 1, o);
            const V3 ddx = sub(dx, d), ddy = sub(dy, d);
            memcpy(r.o, &o, 12);
            memcpy(r.d, &d, 12);
            r.c[0] = r.c[1] = r.c[2] = 1.0f;
            r.ior = 1.0f;
            memset(r.do_dx, 0, 12);
            memset(r.do_dy, 0, 12);
            memcpy(r.dd_dx, &ddx, 12);
            memcpy(r.dd_dy, &ddy, 12);
            r.xy = (x << 16) | y;
            r.ray_depth = 0;
        }
    }
    return (int)i;
}

int orc_trace(const rh_scene_view *sc, const rh_ray_packet *rays, size_t n, int any_hit, const float *tmax,
              rh_hit *out, orc_counters *counters) {
    for (size_t i = 0; i < n; i++) {
        Hit h;
        if (tmax) h.t = tmax[i];
        macro(*sc, rays[i].o, rays[i].d, any_hit != 0, h, counters);
        out[i] = rh_hit{h.mask, h.obj, h.prim, h.t, h.u, h.v, rays[i].xy, 0};
    }
    return 0;
}

int orc_ray_hashes(const rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3], uint32_t *out) {
    for (size_t i = 0; i < n; i++) out[i] = ray_hash(rays[i], root_min, cell_size);
    return 0;
}

// SortRays_CPU (CoreRef.cpp:615-657): the run-length/radix/decompress steps
// amount to a stable sort by hash key
int orc_sort_rays(rh_ray_packet *rays, size_t n, const float root_min[3], const float cell_size[3]) {
    std::vector<uint32_t> key(n);
    std::vector<size_t> idx(n);
    for (size_t i = 0; i < n; i++) {
        key[i] = ray_hash(rays[i], root_min, cell_size);
        idx[i] = i;
    }
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
    std::vector<rh_ray_packet> tmp(rays, rays + n);
    for (size_t i = 0; i < n; i++) rays[i] = tmp[idx[i]];
    return 0;
}

int orc_shade(const rh_scene_view *sc, const rh_pass_info *pi, const rh_ray_packet *rays, const rh_hit *hits, size_t n,
              const float *halton, float *out_rgba, rh_ray_packet *out_secondary, int *out_count, int *out_src) {
    Ctx c{*sc, halton};
    int cnt = 0;
    for (size_t i = 0; i < n; i++) {
        bool e;
        shade(c, pi[i], hits[i], rays[i], out_rgba + 4 * i, out_secondary + cnt, e);
        if (e) {
            if (out_src) out_src[cnt] = (int)i;
            cnt++;
        }
    }
    *out_count = cnt;
    return 0;
}

// the max_total_depth clamp: MAX_BOUNCES of the reference, or a larger
// value for the HIP backend's deeper-path extension ("max_bounces")
static int g_max_bounces = RH_MAX_BOUNCES;
int orc_set_max_bounces(int n) {
    if (n < 1 || n > RH_MAX_BOUNCES_EXT) return -1;
    g_max_bounces = n;
    return 0;
}

// UseCoherentSampling: the random index of a pixel is its entry of the 8x8
// sampling pattern (RendererRef.cpp:165-168, 262-265; pattern Core.cpp) instead
// of the pixel index
static int rand_index_of(const rh_pass_info &pi, int x, int y) {
    static const int pattern[64] = {0, 1, 2, 3, 1, 2, 3, 0, 2, 3, 0, 1, 3, 0, 1, 2, 1, 2, 3, 0, 2, 1,
                                    0, 3, 3, 0, 1, 2, 0, 3, 2, 1, 2, 3, 0, 1, 1, 0, 3, 2, 0, 1, 2, 3,
                                    0, 3, 2, 1, 3, 2, 1, 0, 3, 2, 1, 0, 1, 0, 3, 2, 2, 1, 0, 3};
    return (pi.settings.flags & RH_USE_COHERENT_SAMPLING) ? pattern[(y % 8) * 8 + (x % 8)] : pi.index;
}

// float -> int as x86 cvttss2si: NaN and out-of-range values give INT_MIN
static int f2i_x86(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT32_MIN;
    return (int)f;
}

// SampleMeshInTextureSpace (CoreRef.cpp:507-613): the region's rays and hits
// for a geometry camera, triangle by triangle in order, the first triangle
// whose texture-space footprint covers a pixel's jittered sample taking it
static void sample_mesh_in_texture_space(const rh_scene_view *sc, const rh_camera *cam, int w, int h,
                                         const int32_t r[4], int iteration, const float *halton,
                                         rh_ray_packet *rays, rh_hit *hits) {
    for (int y = r[1]; y < r[1] + r[3]; y++) {
        for (int x = r[0]; x < r[0] + r[2]; x++) {
            const int i = (y - r[1]) * r[2] + (x - r[0]);
            memset(&rays[i], 0, sizeof(rays[i]));
            rays[i].xy = (x << 16) | y;
            rays[i].c[0] = rays[i].c[1] = rays[i].c[2] = 1.0f;
            rays[i].ior = 1.0f;
            memset(&hits[i], 0, sizeof(hits[i]));
            hits[i].obj = hits[i].prim = -1;
            hits[i].t = RH_MAX_DIST;
            hits[i].xy = rays[i].xy;
        }
    }
    const rh_mesh_instance &mi = sc->mesh_instances[cam->mi_index];
    const rh_mesh &mesh = sc->meshes[mi.mesh_index];
    const rh_transform &tr = sc->transforms[mi.tr_index];
    const int uv = (int)cam->uv_index;
    const float W = (float)w, H = (float)h;
    const int rmin[2] = {r[0], r[1]}, rmax[2] = {r[0] + r[2] - 1, r[1] + r[3] - 1};
    for (uint32_t tri = mesh.tris_index; tri < mesh.tris_index + mesh.tris_count; tri++) {
        const rh_vertex &v0 = sc->vertices[sc->vtx_indices[tri * 3 + 0]];
        const rh_vertex &v1 = sc->vertices[sc->vtx_indices[tri * 3 + 1]];
        const rh_vertex &v2 = sc->vertices[sc->vtx_indices[tri * 3 + 2]];
        const float t0[2] = {v0.t[uv][0] * W, (1.0f - v0.t[uv][1]) * H};
        const float t1[2] = {v1.t[uv][0] * W, (1.0f - v1.t[uv][1]) * H};
        const float t2[2] = {v2.t[uv][0] * W, (1.0f - v2.t[uv][1]) * H};
        float bmin[2], bmax[2];
        int ib0[2], ib1[2];
        for (int k = 0; k < 2; k++) {
            bmin[k] = std::min(std::min(t0[k], t1[k]), t2[k]);
            bmax[k] = std::max(std::max(t0[k], t1[k]), t2[k]);
            ib0[k] = f2i_x86(bmin[k]);
            ib1[k] = f2i_x86(std::round(bmax[k]));
        }
        if (ib1[0] < rmin[0] || ib1[1] < rmin[1] || ib0[0] > rmax[0] || ib0[1] > rmax[1]) continue;
        for (int k = 0; k < 2; k++) {
            ib0[k] = std::max(ib0[k], rmin[k]);
            ib1[k] = std::min(ib1[k], rmax[k]);
        }
        const float d01[2] = {t0[0] - t1[0], t0[1] - t1[1]}, d12[2] = {t1[0] - t2[0], t1[1] - t2[1]},
                    d20[2] = {t2[0] - t0[0], t2[1] - t0[1]};
        const float area = d01[0] * d20[1] - d20[0] * d01[1];
        if (area < RH_FLT_EPS) continue;
        const float inv_area = 1.0f / area;
        for (int y = ib0[1]; y <= ib1[1]; y++) {
            for (int x = ib0[0]; x <= ib1[0]; x++) {
                const int i = (y - r[1]) * r[2] + (x - r[0]);
                if (hits[i].mask) continue;
                const int index = y * w + x;
                const int hi = (iteration & (RH_HALTON_SEQ_LEN - 1)) * RH_HALTON_COUNT;
                const int hv = hash(index);
                const float sx = float(x) + fracpart(halton[hi + 0] + construct_float((uint32_t)hv));
                const float sy = float(y) + fracpart(halton[hi + 1] + construct_float((uint32_t)hash(hv)));
                float u = d01[0] * (sy - t0[1]) - d01[1] * (sx - t0[0]);
                float v = d12[0] * (sy - t1[1]) - d12[1] * (sx - t1[0]);
                float ww = d20[0] * (sy - t2[1]) - d20[1] * (sx - t2[0]);
                if (u >= -RH_FLT_EPS && v >= -RH_FLT_EPS && ww >= -RH_FLT_EPS) {
                    u *= inv_area;
                    v *= inv_area;
                    ww *= inv_area;
                    float pl[3], nl[3];
                    for (int k = 0; k < 3; k++) {
                        pl[k] = v0.p[k] * v + v1.p[k] * ww + v2.p[k] * u;
                        nl[k] = v0.n[k] * v + v1.n[k] * ww + v2.n[k] * u;
                    }
                    const float *X = tr.xform, *Y = tr.inv_xform;
                    float p[3], n[3];
                    for (int k = 0; k < 3; k++) {
                        p[k] = X[k] * pl[0] + X[4 + k] * pl[1] + X[8 + k] * pl[2] + X[12 + k];
                        n[k] = Y[4 * k] * nl[0] + Y[4 * k + 1] * nl[1] + Y[4 * k + 2] * nl[2];
                    }
                    for (int k = 0; k < 3; k++) {
                        rays[i].o[k] = p[k] + n[k];
                        rays[i].d[k] = -n[k];
                    }
                    hits[i].mask = -1;
                    hits[i].prim = (int)tri;
                    hits[i].obj = (int)cam->mi_index;
                    hits[i].t = 1.0f;
                    hits[i].u = ww;
                    hits[i].v = u;
                }
            }
        }
    }
}

// RendererRef::RenderScene (RendererRef.cpp:19-325) over one region
int orc_render(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
               int *iteration, float *clean, float *final_rgba, unsigned long long ray_counts[3]) {
    return orc_render_sh(sc, cam, w, h, rect, passes, iteration, clean, final_rgba, ray_counts, nullptr, nullptr);
}

// ... with the OutputSH data (FramebufferRef.h:58-90, FramebufferRef.cpp:48-94)
// in sh_temp / sh_clean (12 floats per pixel, kept by the caller between calls)
int orc_render_sh(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
                  int *iteration, float *clean, float *final_rgba, unsigned long long ray_counts[3], float *sh_temp,
                  float *sh_clean) {
    static const std::vector<uint16_t> pm = perms();
    std::vector<float> halton(RH_HALTON_TABLE_FLOATS);
    int have_halton = -1;
    float rmin[3] = {0, 0, 0}, cell[3] = {1, 1, 1};
    if (sc->macro_nodes_root != RH_INVALID && sc->mnodes) {
        // the wide root as RendererRef.cpp:52-66 reads it: slot 0 of a leaf,
        // else the last used slot
        const rh_mbvh_node &root = sc->mnodes[sc->macro_nodes_root];
        const bool leaf = (root.child[0] & RH_LEAF_NODE_BIT) != 0;
        float rmax[3] = {0, 0, 0};
        for (int j = 0; j < (leaf ? 1 : 8); j++) {
            if (!leaf && root.child[j] == RH_MBVH_EMPTY) continue;
            for (int i = 0; i < 3; i++) {
                rmin[i] = root.bbox_min[i][j];
                rmax[i] = root.bbox_max[i][j];
            }
        }
        for (int i = 0; i < 3; i++) cell[i] = (rmax[i] - rmin[i]) / 255;
    } else if (sc->macro_nodes_root != RH_INVALID) {
        const rh_bvh_node &root = sc->nodes[sc->macro_nodes_root];
        for (int i = 0; i < 3; i++) {
            rmin[i] = root.bbox_min[i];
            cell[i] = (root.bbox_max[i] - root.bbox_min[i]) / 255;
        }
    }
    const size_t n = (size_t)rect[2] * rect[3];
    std::vector<rh_ray_packet> rays(n), sec(n);
    std::vector<rh_hit> hits(n);
    std::vector<float> temp((size_t)w * h * 4, 0.0f);
    for (int pass = 0; pass < passes; pass++) {
        const int it = ++(*iteration);
        if (have_halton < 0 || it % RH_HALTON_SEQ_LEN == 0) {
            halton_table_padded(pm, it, halton.data());
            have_halton = it;
        }
        rh_pass_info pi;
        memset(&pi, 0, sizeof(pi));
        pi.iteration = it;
        pi.bounce = 2;
        pi.settings = cam->pass_settings;
        pi.settings.max_total_depth = (uint8_t)std::min<int>(pi.settings.max_total_depth, g_max_bounces);
        Ctx c{*sc, halton.data()};

        if (cam->type == RH_CAM_GEO) {
            sample_mesh_in_texture_space(sc, cam, w, h, rect, it, halton.data(), rays.data(), hits.data());
        } else {
            orc_gen_primary(cam, rect, w, h, it, halton.data(), rays.data());
            orc_trace(sc, rays.data(), n, 0, nullptr, hits.data(), nullptr);
            if (ray_counts) ray_counts[0] += n;
        }
        int count = 0;
        for (size_t i = 0; i < n; i++) {
            const int x = (rays[i].xy >> 16) & 0xffff, y = rays[i].xy & 0xffff;
            pi.index = y * w + x;
            pi.rand_index = rand_index_of(pi, x, y);
            bool e;
            float col[4];
            shade(c, pi, hits[i], rays[i], col, &sec[count], e);
            if (e) count++;
            memcpy(&temp[(size_t)pi.index * 4], col, 16);
        }
        const bool sh = sh_temp && (cam->pass_settings.flags & RH_OUTPUT_SH);
        if (sh) {
            for (int y = rect[1]; y < rect[1] + rect[3]; y++)
                for (int x = rect[0]; x < rect[0] + rect[2]; x++) sh_temp[12 * ((size_t)y * w + x) + 4] = 0.0f;
            for (int i = 0; i < count; i++) {
                const int x = (sec[i].xy >> 16) & 0xffff, y = sec[i].xy & 0xffff;
                float *s = &sh_temp[12 * ((size_t)y * w + x)];
                s[0] = 0.282094806f;
                s[1] = 0.488602519f * sec[i].d[1];
                s[2] = 0.488602519f * sec[i].d[2];
                s[3] = 0.488602519f * sec[i].d[0];
                s[4] = sec[i].c[0];
            }
        }
        for (int b = 0; b < pi.settings.max_total_depth && count && !(pi.settings.flags & RH_SKIP_INDIRECT_LIGHT); b++) {
            orc_sort_rays(sec.data(), (size_t)count, rmin, cell);
            orc_trace(sc, sec.data(), (size_t)count, 0, nullptr, hits.data(), nullptr);
            if (ray_counts) ray_counts[1] += count;
            std::swap(rays, sec);
            const int rc = count;
            count = 0;
            pi.bounce = b + 3;
            for (int i = 0; i < rc; i++) {
                const int x = (rays[i].xy >> 16) & 0xffff, y = rays[i].xy & 0xffff;
                pi.index = y * w + x;
                pi.rand_index = rand_index_of(pi, x, y);
                bool e;
                float col[4];
                shade(c, pi, hits[i], rays[i], col, &sec[count], e);
                if (e) count++;
                float *t = &temp[(size_t)pi.index * 4];
                t[0] += col[0];
                t[1] += col[1];
                t[2] += col[2];
                t[3] += 0.0f;
            }
        }
        if (ray_counts) ray_counts[2] += c.shadow;
        // MixWith + tonemap (FramebufferRef.cpp:80-86, RendererRef.cpp:301-324)
        const float k = 1.0f / it;
        for (int y = rect[1]; y < rect[1] + rect[3]; y++) {
            for (int x = rect[0]; x < rect[0] + rect[2]; x++) {
                const size_t p = ((size_t)y * w + x) * 4;
                for (int j = 0; j < 4; j++) clean[p + j] += (temp[p + j] - clean[p + j]) * k;
                if (sh) {
                    // ComputeSHData then MixWith_SH
                    float *s = &sh_temp[3 * p], *o = &sh_clean[3 * p];
                    const float inv_weight = s[4] > RH_FLT_EPS ? (2.0f * RH_PI / s[4]) : 0.0f;
                    const float pr = temp[p] * inv_weight, pg = temp[p + 1] * inv_weight,
                                pb = temp[p + 2] * inv_weight;
                    const float coeff[4] = {s[0], s[1], s[2], s[3]};
                    for (int j = 0; j < 4; j++) {
                        s[4 + j] = coeff[j] * pg;
                        s[8 + j] = coeff[j] * pb;
                        s[j] = coeff[j] * pr;
                    }
                    for (int j = 0; j < 12; j++) o[j] += (s[j] - o[j]) * k;
                }
                float v[4] = {clean[p], clean[p + 1], clean[p + 2], clean[p + 3]};
                if (cam->dtype == RH_DEVICE_SRGB)
                    for (int j = 0; j < 3; j++) v[j] = v[j] > 0.0031308f ? std::pow(1.055f * v[j], (1.0f / 2.4f)) - 0.055f : 12.92f * v[j];
                if (cam->gamma != 1.0f)
                    for (float &x_ : v) x_ = std::pow(x_, 1.0f / cam->gamma);
                if (cam->pass_settings.flags & RH_CLAMP)
                    for (float &x_ : v) x_ = clampf(x_, 0.0f, 1.0f);
                memcpy(&final_rgba[p], v, 16);
            }
        }
    }
    return 0;
}

// FP64 reference gradient of a loss through the rendered mean image, by path
// replay: passes 1..passes are replayed as orc_render renders them; every
// shaded vertex j of a pixel's path records its colour C_j (the value added to
// the pass image) and the material whose main_color scales C_j and all later
// contributions once.  d(pass value)/d main_color[m] = S_m / main_color[m] with
// S_m the suffix sum of the path's colours from the vertex on, accumulated here
// in double: grad[3 m + c] += scale * adj[pixel][c] * S_m[c] / main_color[m][c].
int orc_adjoint(const rh_scene_view *sc, const rh_camera *cam, int w, int h, const int32_t rect[4], int passes,
                const float *adj_rgba, double scale, double *grad, double *texel_grad) {
    static const std::vector<uint16_t> pm = perms();
    std::vector<float> halton(RH_HALTON_TABLE_FLOATS);
    int have_halton = -1;
    const size_t n = (size_t)rect[2] * rect[3];
    std::vector<rh_ray_packet> rays(n), sec(n);
    std::vector<rh_hit> hits(n);
    struct Vtx {
        int mat;
        float c[3];
        AlbedoSample as;
    };
    std::vector<std::vector<Vtx>> path((size_t)w * h);
    for (int it = 1; it <= passes; it++) {
        if (have_halton < 0 || it % RH_HALTON_SEQ_LEN == 0) {
            halton_table_padded(pm, it, halton.data());
            have_halton = it;
        }
        rh_pass_info pi;
        memset(&pi, 0, sizeof(pi));
        pi.iteration = it;
        pi.bounce = 2;
        pi.settings = cam->pass_settings;
        pi.settings.max_total_depth = (uint8_t)std::min<int>(pi.settings.max_total_depth, g_max_bounces);
        Ctx c{*sc, halton.data()};
        for (auto &p : path) p.clear();

        orc_gen_primary(cam, rect, w, h, it, halton.data(), rays.data());
        orc_trace(sc, rays.data(), n, 0, nullptr, hits.data(), nullptr);
        int count = 0;
        auto shade_all = [&](int rc) {
            count = 0;
            for (int i = 0; i < rc; i++) {
                const int x = (rays[i].xy >> 16) & 0xffff, y = rays[i].xy & 0xffff;
                pi.index = y * w + x;
                pi.rand_index = rand_index_of(pi, x, y);
                bool e;
                float col[4];
                Vtx vx;
                shade(c, pi, hits[i], rays[i], col, &sec[count], e, &vx.mat, &vx.as);
                if (e) count++;
                for (int k = 0; k < 3; k++) vx.c[k] = col[k];
                path[(size_t)pi.index].push_back(vx);
            }
        };
        shade_all((int)n);
        for (int b = 0; b < pi.settings.max_total_depth && count && !(pi.settings.flags & RH_SKIP_INDIRECT_LIGHT); b++) {
            orc_trace(sc, sec.data(), (size_t)count, 0, nullptr, hits.data(), nullptr);
            std::swap(rays, sec);
            pi.bounce = b + 3;
            shade_all(count);
        }
        for (int y = rect[1]; y < rect[1] + rect[3]; y++) {
            for (int x = rect[0]; x < rect[0] + rect[2]; x++) {
                const size_t p = (size_t)y * w + x;
                const auto &v = path[p];
                double S[3] = {0, 0, 0};
                for (size_t k = v.size(); k-- > 0;) {
                    for (int ch = 0; ch < 3; ch++) S[ch] += (double)v[k].c[ch];
                    const int m = v[k].mat;
                    if (m < 0) continue;
                    for (int ch = 0; ch < 3; ch++) {
                        const double mc = (double)sc->materials[m].main_color[ch];
                        if (mc != 0.0) grad[3 * (size_t)m + ch] += scale * (double)adj_rgba[4 * p + ch] * S[ch] / mc;
                    }
                    if (texel_grad) {
                        // albedo = lin(raw) * main_color, raw = sum of texel * weight
                        const AlbedoSample &as = v[k].as;
                        const double kx = as.foot.kx, ky = as.foot.ky;
                        const double wq[4] = {(1 - kx) * (1 - ky), kx * (1 - ky), (1 - kx) * ky, kx * ky};
                        for (int ch = 0; ch < 3; ch++) {
                            const double r = as.raw[ch];
                            double lin = r, dlin = 1.0;
                            if (as.srgb) {
                                if (r > 0.04045) {
                                    lin = std::pow((r + 0.055) / 1.055, 2.4);
                                    dlin = 2.4 / 1.055 * std::pow((r + 0.055) / 1.055, 1.4);
                                } else {
                                    lin = r / 12.92;
                                    dlin = 1.0 / 12.92;
                                }
                            }
                            if (lin == 0.0) continue;
                            const double g = scale * (double)adj_rgba[4 * p + ch] * S[ch] / lin * dlin;
                            for (int q = 0; q < 4; q++) texel_grad[4 * (size_t)as.foot.idx[q] + ch] += g * wq[q];
                        }
                    }
                }
            }
        }
    }
    return 0;
}

} // extern "C"
