// Host-side scene preprocessing; see core.h.  Each routine follows the
// reference routine cited above it step for step, because the BVH node arrays,
// triangle records and index orders must come out identical to the
// reference's: the GPU walks them in the reference's order and must report the
// same hit primitive on identical rays.  Compiled with -ffp-contract=off.
#include "core.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <deque>
#include <functional>
#include <limits>
#include <random>

namespace RayHip {

namespace {
const float axis_aligned_normal_eps = 0.000001f; // Core.cpp:13

inline vec3 cross3(const vec3 &a, const vec3 &b) {
    return {{a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}};
}
// sequential sums, as the generic simd_vec<float, 3> (simd_vec.h:319-327, 559-563)
inline float dot3(const vec3 &a, const vec3 &b) {
    float r = 0;
    r += a[0] * b[0];
    r += a[1] * b[1];
    r += a[2] * b[2];
    return r;
}
inline float length3(const vec3 &a) { return std::sqrt(dot3(a, a)); }
inline vec3 div3(const vec3 &a, float s) { return {{a[0] / s, a[1] / s, a[2] / s}}; }
inline vec3 normalize3(const vec3 &a) { return div3(a, length3(a)); }
inline vec3 load3(const float *p) { return {{p[0], p[1], p[2]}}; }

// bbox_t (BVHSplit.cpp:8-23): "surface area" is 2 * (dx + dy + dz)
struct bbox_t {
    vec3 min = {{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(), std::numeric_limits<float>::max()}};
    vec3 max = {{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
                 std::numeric_limits<float>::lowest()}};
    bbox_t() = default;
    bbox_t(const vec3 &a, const vec3 &b) : min(a), max(b) {}
    float surface_area() const { return surface_area(min, max); }
    static float surface_area(const vec3 &mn, const vec3 &mx) {
        float d0 = mx[0] - mn[0], d1 = mx[1] - mn[1], d2 = mx[2] - mn[2];
        return 2 * (d0 + d1 + d2);
    }
};

struct split_data_t {
    std::vector<uint32_t> left_indices, right_indices;
    bbox_t left_bounds, right_bounds;
};

// ---- spatial splits (BVHSplit.cpp:6-124, 220-369) ----
const float kSpatialSplitAlpha = 0.00001f; // BVHSplit.cpp:6
const int kSpatialBins = 256;              // BVHSplit.cpp:7

struct dvec3 {
    double v[3];
};

// One Sutherland-Hodgman step (BVHSplit.cpp:26-70): the convex polygon in[0, n)
// clipped to the half space axis >= pos (keep_above) or axis <= pos; the new
// vertices on the plane get exactly pos on that axis.  Returns the new count.
int clip_polygon(const dvec3 *in, int n, dvec3 *out, int axis, double pos, bool keep_above) {
    if (n < 3) return 0;
    const double sign = keep_above ? 1.0 : -1.0;
    auto cut = [&](const dvec3 &a, const dvec3 &b) {
        const double t = (pos - a.v[axis]) / (b.v[axis] - a.v[axis]);
        dvec3 p;
        for (int k = 0; k < 3; k++) p.v[k] = a.v[k] + (b.v[k] - a.v[k]) * t;
        p.v[axis] = pos;
        return p;
    };
    int m = 0;
    dvec3 a = in[0];
    bool a_in = sign * (a.v[axis] - pos) >= 0;
    for (int i = 0; i < n; i++) {
        const dvec3 &b = in[i + 1 == n ? 0 : i + 1];
        const bool b_in = sign * (b.v[axis] - pos) >= 0;
        if (a_in && b_in) {
            out[m++] = b;
        } else if (a_in) {
            out[m++] = cut(a, b);
        } else if (b_in) {
            out[m++] = cut(a, b);
            out[m++] = b;
        }
        a = b;
        a_in = b_in;
    }
    return m;
}

// the float nearest below / above a double (castflt_down / castflt_up,
// BVHSplit.cpp:72-96): one ulp step on the bit pattern when rounding went the
// wrong way
float float_below(double x) {
    float f = (float)x;
    if ((double)f > x) {
        int32_t b;
        memcpy(&b, &f, 4);
        b += f > 0 ? -1 : 1;
        memcpy(&f, &b, 4);
    }
    return f;
}
float float_above(double x) {
    float f = (float)x;
    if ((double)f < x) {
        int32_t b;
        memcpy(&b, &f, 4);
        b += f < 0 ? -1 : 1;
        memcpy(&f, &b, 4);
    }
    return f;
}

// The box of triangle (a, b, c) clipped to `limits` (GetClippedAABB,
// BVHSplit.cpp:98-123): the polygon clipped by the six planes in double
// precision, its vertices rounded outwards.
bbox_t clipped_box(const vec3 &a, const vec3 &b, const vec3 &c, const bbox_t &limits) {
    dvec3 p[9], q[9];
    int n = 3;
    for (int k = 0; k < 3; k++) {
        p[0].v[k] = (double)a[k];
        p[1].v[k] = (double)b[k];
        p[2].v[k] = (double)c[k];
    }
    for (int axis = 0; axis < 3; axis++) {
        n = clip_polygon(p, n, q, axis, limits.min[axis], true);
        n = clip_polygon(q, n, p, axis, limits.max[axis], false);
    }
    bbox_t box;
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            box.min.v[k] = fmin_std(box.min[k], float_below(p[i].v[k]));
            box.max.v[k] = fmax_std(box.max[k], float_above(p[i].v[k]));
        }
    return box;
}

// the three corners of primitive p
struct TriCorners {
    vec3 a, b, c;
};
inline TriCorners corners(const prim_t &p, const float *positions, size_t stride) {
    return {load3(&positions[p.i0 * stride]), load3(&positions[p.i1 * stride]), load3(&positions[p.i2 * stride])};
}

// SplitPrimitives_SAH (BVHSplit.cpp:126-390): the full sorted sweep along each
// axis (primitives ordered by bbox max with std::sort) for the best object
// split; with s.allow_spatial_splits the primitives' boxes are first clipped to
// the node's box, and where the best object split leaves the children
// overlapping by more than a small fraction of the root's area, 256 split
// planes per axis are tried too, a triangle crossing the winning plane going to
// both children with its box clipped to each side.
split_data_t SplitPrimitives_SAH(const prim_t *primitives, const std::vector<uint32_t> &prim_indices,
                                 const float *positions, size_t stride, const vec3 &bbox_min, const vec3 &bbox_max,
                                 const vec3 &root_min, const vec3 &root_max, const bvh_settings_t &s) {
    const size_t num_tris = prim_indices.size();
    const bbox_t whole_box(bbox_min, bbox_max);

    std::vector<uint32_t> axis_lists[3];
    for (int axis = 0; axis < 3; axis++) {
        axis_lists[axis].resize(num_tris);
        for (uint32_t i = 0; i < (uint32_t)num_tris; i++) axis_lists[axis][i] = i;
    }

    // primitive boxes as the sweep sees them: clipped to the node with spatial splits
    const bool clip = s.allow_spatial_splits && positions;
    std::vector<bbox_t> clipped;
    if (clip) {
        clipped.resize(num_tris);
        for (size_t i = 0; i < num_tris; i++) {
            const TriCorners t = corners(primitives[prim_indices[i]], positions, stride);
            clipped[i] = clipped_box(t.a, t.b, t.c, whole_box);
        }
    }
    auto box_min = [&](uint32_t i) -> const vec3 & { return clip ? clipped[i].min : primitives[prim_indices[i]].bbox_min; };
    auto box_max = [&](uint32_t i) -> const vec3 & { return clip ? clipped[i].max : primitives[prim_indices[i]].bbox_max; };

    std::vector<bbox_t> right_bounds(num_tris);

    float res_sah = s.oversplit_threshold * whole_box.surface_area() * num_tris;
    int div_axis = -1;
    uint32_t div_index = 0;
    bbox_t res_left_bounds, res_right_bounds;

    for (int axis = 0; axis < 3; axis++) {
        std::vector<uint32_t> &list = axis_lists[axis];
        if (clip) {
            std::sort(list.begin(), list.end(), [axis, &clipped](uint32_t p1, uint32_t p2) -> bool {
                return clipped[p1].max[axis] < clipped[p2].max[axis];
            });
        } else {
            std::sort(list.begin(), list.end(), [axis, primitives, &prim_indices](uint32_t p1, uint32_t p2) -> bool {
                return primitives[prim_indices[p1]].bbox_max[axis] < primitives[prim_indices[p2]].bbox_max[axis];
            });
        }

        bbox_t cur_right_bounds;
        for (size_t i = list.size() - 1; i > 0; i--) {
            cur_right_bounds.min = vmin(cur_right_bounds.min, box_min(list[i]));
            cur_right_bounds.max = vmax(cur_right_bounds.max, box_max(list[i]));
            right_bounds[i - 1] = cur_right_bounds;
        }

        bbox_t left_bounds;
        for (size_t i = 1; i < list.size(); i++) {
            left_bounds.min = vmin(left_bounds.min, box_min(list[i - 1]));
            left_bounds.max = vmax(left_bounds.max, box_max(list[i - 1]));

            const float sah = s.node_traversal_cost * whole_box.surface_area() + left_bounds.surface_area() * i +
                              right_bounds[i - 1].surface_area() * (list.size() - i);
            if (sah < res_sah) {
                res_sah = sah;
                div_axis = axis;
                div_index = (uint32_t)i;
                res_left_bounds = left_bounds;
                res_right_bounds = right_bounds[i - 1];
            }
        }
    }

    // The spatial split.  (The reference also enters it with allow_spatial_splits
    // and no positions, where it would then read clipped boxes it never made;
    // mesh builds always pass positions and the other trees never ask for it.)
    const bbox_t overlap(vmax(res_left_bounds.min, res_right_bounds.min), vmin(res_left_bounds.max, res_right_bounds.max));
    const bool disjoint_free = !(overlap.max[0] <= overlap.min[0]) && !(overlap.max[1] <= overlap.min[1]) &&
                               !(overlap.max[2] <= overlap.min[2]);
    if (clip && disjoint_free &&
        overlap.surface_area() > kSpatialSplitAlpha * bbox_t::surface_area(root_min, root_max)) {
        struct Bin {
            bbox_t extends, limits;
            uint32_t enter = 0, exit = 0, prims = 0;
        };
        int spatial_split = -1;
        std::vector<Bin> bins(kSpatialBins);
        for (int axis = 0; axis < 3; axis++) {
            const double bin_size = double(whole_box.max[axis] - whole_box.min[axis]) / kSpatialBins;
            if (bin_size < RH_FLT_EPS) continue; // flat along this axis
            for (int i = 0; i < kSpatialBins; i++) {
                Bin &b = bins[i];
                b = Bin();
                b.limits = whole_box;
                b.limits.min.v[axis] = float_below(whole_box.min[axis] + i * bin_size);
                b.limits.max.v[axis] = float_above(whole_box.min[axis] + (i + 1) * bin_size);
            }
            bins[kSpatialBins - 1].limits.max.v[axis] = whole_box.max[axis];

            for (const uint32_t i : axis_lists[axis]) {
                const float lo = clipped[i].min[axis], hi = clipped[i].max[axis];
                int first = 0, last = 0;
                for (int j = 0; j < kSpatialBins && lo >= bins[j].limits.min[axis]; j++) first = j;
                for (int j = first; j < kSpatialBins && hi >= bins[j].limits.min[axis]; j++) last = j;
                bins[first].enter++;
                bins[last].exit++;
                const TriCorners t = corners(primitives[prim_indices[i]], positions, stride);
                for (int j = first; j <= last; j++) {
                    const bbox_t box = clipped_box(t.a, t.b, t.c, bins[j].limits);
                    bins[j].extends.min = vmin(bins[j].extends.min, box.min);
                    bins[j].extends.max = vmax(bins[j].extends.max, box.max);
                    bins[j].prims++;
                }
            }
            for (Bin &b : bins) {
                if (!b.prims) {
                    b.extends = b.limits;
                } else {
                    b.extends.min = vmax(b.extends.min, b.limits.min);
                    b.extends.max = vmin(b.extends.max, b.limits.max);
                }
            }
            for (int split = 1; split < kSpatialBins; split++) {
                bbox_t left, right;
                int n_left = 0, n_right = 0;
                for (int i = 0; i < split; i++) {
                    if (!bins[i].prims) continue;
                    left.min = vmin(left.min, bins[i].extends.min);
                    left.max = vmax(left.max, bins[i].extends.max);
                    n_left += (int)bins[i].enter;
                }
                for (int i = split; i < kSpatialBins; i++) {
                    if (!bins[i].prims) continue;
                    right.min = vmin(right.min, bins[i].extends.min);
                    right.max = vmax(right.max, bins[i].extends.max);
                    n_right += (int)bins[i].exit;
                }
                const float sah =
                    s.node_traversal_cost + left.surface_area() * n_left + right.surface_area() * n_right;
                if (sah < res_sah) {
                    res_sah = sah;
                    spatial_split = split;
                    div_axis = axis;
                    res_left_bounds = left;
                    res_right_bounds = right;
                }
            }
        }
        if (spatial_split != -1) {
            // a primitive reaching into both sides is referenced by both children
            split_data_t out;
            for (const uint32_t i : axis_lists[div_axis]) {
                if (clipped[i].min[div_axis] <= res_left_bounds.max[div_axis]) out.left_indices.push_back(prim_indices[i]);
                if (clipped[i].max[div_axis] >= res_right_bounds.min[div_axis])
                    out.right_indices.push_back(prim_indices[i]);
            }
            out.left_bounds = res_left_bounds;
            out.right_bounds = res_right_bounds;
            return out;
        }
    }

    split_data_t out;
    if (div_axis != -1) {
        out.left_indices.reserve(div_index);
        out.right_indices.reserve(prim_indices.size() - div_index);
        for (size_t i = 0; i < div_index; i++) out.left_indices.push_back(prim_indices[axis_lists[div_axis][i]]);
        for (size_t i = div_index; i < axis_lists[div_axis].size(); i++)
            out.right_indices.push_back(prim_indices[axis_lists[div_axis][i]]);
    } else {
        out.left_indices = prim_indices;
        res_left_bounds = whole_box;
    }
    out.left_bounds = res_left_bounds;
    out.right_bounds = res_right_bounds;
    return out;
}
} // namespace

bool PreprocessTri(const float *p, int stride, rh_tri_accel *acc) {
    if (!stride) stride = 3;

    float e0[3] = {p[stride] - p[0], p[stride + 1] - p[1], p[stride + 2] - p[2]},
          e1[3] = {p[stride * 2] - p[0], p[stride * 2 + 1] - p[1], p[stride * 2 + 2] - p[2]};

    float n[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2], e0[0] * e1[1] - e0[1] * e1[0]};

    int w = 2, u = 0, v = 1;
    if (std::abs(n[0]) >= std::abs(n[1]) && std::abs(n[0]) >= std::abs(n[2])) {
        w = 0;
        u = 1;
        v = 2;
    } else if (std::abs(n[1]) >= std::abs(n[0]) && std::abs(n[1]) >= std::abs(n[2])) {
        w = 1;
        u = 0;
        v = 2;
    }

    if (std::abs(n[w]) > RH_FLT_EPS) {
        acc->nu = n[u] / n[w];
        acc->nv = n[v] / n[w];
        acc->pu = p[u];
        acc->pv = p[v];
        acc->np = acc->nu * acc->pu + acc->nv * acc->pv + p[w];

        const float sign = (w == 1) ? -1.0f : 1.0f;
        acc->e0u = sign * e0[u] / n[w];
        acc->e0v = sign * e0[v] / n[w];
        acc->e1u = sign * e1[u] / n[w];
        acc->e1v = sign * e1[v] / n[w];

        acc->ci = w;
        if (std::abs(acc->nu) < axis_aligned_normal_eps && std::abs(acc->nv) < axis_aligned_normal_eps) {
            acc->ci |= RH_TRI_AXIS_ALIGNED_BIT;
        }
        if (n[w] < 0) acc->ci |= RH_TRI_INV_NORMAL_BIT;
        return true;
    }
    // InvalidTriangle (Core.cpp:95-103)
    const float qnan = std::numeric_limits<float>::quiet_NaN();
    acc->nu = acc->nv = acc->np = acc->pu = acc->pv = qnan;
    acc->ci = 0;
    acc->e0u = acc->e0v = acc->e1u = acc->e1v = qnan;
    acc->mi = acc->back_mi = 0xffffffffu;
    return false;
}

void ExtractPlaneNormal(const rh_tri_accel &tri, float *out_normal) {
    static const int next_u[] = {1, 0, 0}, next_v[] = {2, 2, 1};
    const int iw = tri.ci & RH_TRI_W_BITS;
    out_normal[iw] = 1.0f;
    out_normal[next_u[iw]] = tri.nu;
    out_normal[next_v[iw]] = tri.nv;
    float inv_l = std::sqrt(out_normal[0] * out_normal[0] + out_normal[1] * out_normal[1] + out_normal[2] * out_normal[2]);
    if (tri.ci & RH_TRI_INV_NORMAL_BIT) inv_l = -inv_l;
    out_normal[0] *= inv_l;
    out_normal[1] *= inv_l;
    out_normal[2] *= inv_l;
}

uint32_t PreprocessMesh(const float *attrs, const uint32_t *vtx_indices, size_t vtx_indices_count, int layout,
                        int base_vertex, const bvh_settings_t &s, std::vector<rh_bvh_node> &out_nodes,
                        std::vector<rh_tri_accel> &out_tris, std::vector<uint32_t> &out_tri_indices,
                        const LbvhBottom *device, const SahBoxes *sah) {
    static const size_t attr_strides[] = {8, 10, 11, 13};
    std::vector<prim_t> primitives;
    primitives.reserve(vtx_indices_count / 3);

    const size_t tris_start = out_tris.size();
    const size_t tris_count = vtx_indices_count / 3;
    out_tris.resize(tris_start + tris_count);

    const float *positions = attrs;
    const size_t attr_stride = attr_strides[layout];

    for (size_t j = 0; j < vtx_indices_count; j += 3) {
        float p[9];
        const uint32_t i0 = vtx_indices[j + 0] + base_vertex, i1 = vtx_indices[j + 1] + base_vertex,
                       i2 = vtx_indices[j + 2] + base_vertex;
        memcpy(&p[0], &positions[i0 * attr_stride], 3 * sizeof(float));
        memcpy(&p[3], &positions[i1 * attr_stride], 3 * sizeof(float));
        memcpy(&p[6], &positions[i2 * attr_stride], 3 * sizeof(float));

        PreprocessTri(&p[0], 0, &out_tris[tris_start + j / 3]);

        const vec3 p0 = load3(&p[0]), p1 = load3(&p[3]), p2 = load3(&p[6]);
        primitives.push_back({i0, i1, i2, vmin(p0, vmin(p1, p2)), vmax(p0, vmax(p1, p2))});
    }

    const size_t indices_start = out_tri_indices.size();
    const uint32_t num_out_nodes =
        s.use_fast_bvh_build
            ? PreprocessPrims_HLBVH(primitives.data(), primitives.size(), out_nodes, out_tri_indices, device, sah)
            : PreprocessPrims_SAH(primitives.data(), primitives.size(), positions, attr_stride, s, out_nodes,
                                  out_tri_indices);
    for (size_t i = indices_start; i < out_tri_indices.size(); i++) out_tri_indices[i] += (uint32_t)tris_start;
    return num_out_nodes;
}

uint32_t PreprocessPrims_SAH(const prim_t *prims, size_t prims_count, const float *positions, size_t stride,
                             const bvh_settings_t &s, std::vector<rh_bvh_node> &out_nodes,
                             std::vector<uint32_t> &out_indices) {
    struct prims_coll_t {
        std::vector<uint32_t> indices;
        vec3 min = {{std::numeric_limits<float>::max(), std::numeric_limits<float>::max(),
                     std::numeric_limits<float>::max()}};
        vec3 max = {{std::numeric_limits<float>::lowest(), std::numeric_limits<float>::lowest(),
                     std::numeric_limits<float>::lowest()}};
        prims_coll_t() = default;
        prims_coll_t(std::vector<uint32_t> &&ind, const vec3 &mn, const vec3 &mx)
            : indices(std::move(ind)), min(mn), max(mx) {}
    };

    // breadth-first: new lists go to the front, the oldest is taken from the back
    std::deque<prims_coll_t> prim_lists;
    prim_lists.emplace_back();

    size_t num_nodes = out_nodes.size();
    const uint32_t root_node_index = (uint32_t)num_nodes;

    for (uint32_t j = 0; j < (uint32_t)prims_count; j++) {
        prim_lists.back().indices.push_back(j);
        prim_lists.back().min = vmin(prim_lists.back().min, prims[j].bbox_min);
        prim_lists.back().max = vmax(prim_lists.back().max, prims[j].bbox_max);
    }

    const vec3 root_min = prim_lists.back().min, root_max = prim_lists.back().max;
    while (!prim_lists.empty()) {
        split_data_t split = SplitPrimitives_SAH(prims, prim_lists.back().indices, positions, stride,
                                                 prim_lists.back().min, prim_lists.back().max, root_min, root_max, s);
        prim_lists.pop_back();

        if (split.right_indices.empty()) {
            rh_bvh_node n;
            n.prim_index = RH_LEAF_NODE_BIT + (uint32_t)out_indices.size();
            n.prim_count = (uint32_t)split.left_indices.size();
            memcpy(n.bbox_min, split.left_bounds.min.v, 3 * sizeof(float));
            memcpy(n.bbox_max, split.left_bounds.max.v, 3 * sizeof(float));
            out_nodes.push_back(n);
            out_indices.insert(out_indices.end(), split.left_indices.begin(), split.left_indices.end());
        } else {
            const uint32_t index = (uint32_t)num_nodes;

            uint32_t space_axis = 0;
            const vec3 &lb0 = split.left_bounds.min, &lb1 = split.left_bounds.max;
            const vec3 &rb0 = split.right_bounds.min, &rb1 = split.right_bounds.max;
            const vec3 c_left = {{(lb0[0] + lb1[0]) / 2, (lb0[1] + lb1[1]) / 2, (lb0[2] + lb1[2]) / 2}};
            const vec3 c_right = {{(rb0[0] + rb1[0]) / 2, (rb0[1] + rb1[1]) / 2, (rb0[2] + rb1[2]) / 2}};
            const vec3 dist = {{std::abs(c_left[0] - c_right[0]), std::abs(c_left[1] - c_right[1]),
                                std::abs(c_left[2] - c_right[2])}};
            if (dist[0] > dist[1] && dist[0] > dist[2]) {
                space_axis = 0;
            } else if (dist[1] > dist[0] && dist[1] > dist[2]) {
                space_axis = 1;
            } else {
                space_axis = 2;
            }

            const vec3 bmin = vmin(lb0, rb0), bmax = vmax(lb1, rb1);
            rh_bvh_node n;
            n.prim_index = index + 1;
            n.prim_count = (space_axis << 30) + index + 2;
            memcpy(n.bbox_min, bmin.v, 3 * sizeof(float));
            memcpy(n.bbox_max, bmax.v, 3 * sizeof(float));
            out_nodes.push_back(n);

            prim_lists.emplace_front(std::move(split.left_indices), lb0, lb1);
            prim_lists.emplace_front(std::move(split.right_indices), rb0, rb1);
            num_nodes += 2;
        }
    }
    return (uint32_t)(out_nodes.size() - root_node_index);
}

uint32_t PreprocessPrims_SAH_Boxes(const prim_t *prims, size_t prims_count, const bvh_settings_t &s,
                                   std::vector<rh_bvh_node> &out_nodes, std::vector<uint32_t> &out_indices,
                                   const SahBoxes *sah) {
    if (sah && prims_count) {
        std::vector<float> boxes(6 * prims_count);
        for (size_t i = 0; i < prims_count; i++) {
            memcpy(&boxes[6 * i], prims[i].bbox_min.v, 3 * sizeof(float));
            memcpy(&boxes[6 * i + 3], prims[i].bbox_max.v, 3 * sizeof(float));
        }
        std::vector<rh_bvh_node> nodes;
        std::vector<uint32_t> indices;
        bvh_settings_t plain = s;
        plain.allow_spatial_splits = false;
        if ((*sah)(boxes.data(), prims_count, plain, nodes, indices)) {
            const uint32_t node_base = (uint32_t)out_nodes.size(), index_base = (uint32_t)out_indices.size();
            for (rh_bvh_node n : nodes) {
                if (n.prim_index & RH_LEAF_NODE_BIT) {
                    n.prim_index = RH_LEAF_NODE_BIT + (n.prim_index & RH_PRIM_INDEX_BITS) + index_base;
                } else {
                    n.prim_index += node_base;
                    n.prim_count = (n.prim_count & RH_SEP_AXIS_BITS) + (n.prim_count & RH_RIGHT_CHILD_BITS) + node_base;
                }
                out_nodes.push_back(n);
            }
            out_indices.insert(out_indices.end(), indices.begin(), indices.end());
            return (uint32_t)nodes.size();
        }
    }
    bvh_settings_t plain = s;
    plain.allow_spatial_splits = false;
    return PreprocessPrims_SAH(prims, prims_count, nullptr, 0, plain, out_nodes, out_indices);
}

// ---- the fast build: HLBVH (mesh_desc_t::use_fast_bvh_build) ----

uint32_t F2U32_x86(float f) {
    // x86-64 converts float to uint32_t through a 64-bit truncation and keeps
    // the low half; NaN and values outside the int64 range give 0x8000...0
    if (!(f > -9.2233715e18f && f < 9.2233715e18f)) return 0u;
    return (uint32_t)(int64_t)f;
}

static uint32_t spread_bits3(uint32_t x) {
    // the 10 low bits of x moved to bits 0, 3, 6, ..., 27
    x &= 0x3ffu;
    x = (x | (x << 16)) & 0x030000ffu;
    x = (x | (x << 8)) & 0x0300f00fu;
    x = (x | (x << 4)) & 0x030c30c3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

uint32_t MortonCode3(uint32_t x, uint32_t y, uint32_t z) {
    return (spread_bits3(z) << 2) + (spread_bits3(y) << 1) + spread_bits3(x);
}

void MortonCodes(const prim_t *prims, size_t n, uint32_t *codes) {
    const float big = std::numeric_limits<float>::max();
    vec3 lo = {{big, big, big}}, hi = {{-big, -big, -big}};
    for (size_t j = 0; j < n; j++) {
        lo = vmin(lo, prims[j].bbox_min);
        hi = vmax(hi, prims[j].bbox_max);
    }
    float scale[3];
    for (int a = 0; a < 3; a++) scale[a] = (float)(1 << kMortonBitsPerAxis) / (hi[a] - lo[a]);
    for (size_t j = 0; j < n; j++) {
        uint32_t c[3];
        for (int a = 0; a < 3; a++) {
            const float centre = 0.5f * (prims[j].bbox_min[a] + prims[j].bbox_max[a]);
            c[a] = F2U32_x86((centre - lo[a]) * scale[a]);
        }
        codes[j] = MortonCode3(c[0], c[1], c[2]);
    }
}

static void emit_lbvh_node(const prim_t *prims, const uint32_t *indices, const uint32_t *codes, uint32_t node,
                           uint32_t first, uint32_t count, uint32_t index_offset, int bit,
                           std::vector<rh_bvh_node> &out) {
    const uint32_t last = first + count - 1;
    while (bit >= 0 && count >= 8 && ((codes[first] >> bit) & 1u) == ((codes[last] >> bit) & 1u)) bit--;
    if (bit < 0 || count < 8) {
        const float big = std::numeric_limits<float>::max();
        vec3 lo = {{big, big, big}}, hi = {{-big, -big, -big}};
        for (uint32_t i = first; i <= last; i++) {
            lo = vmin(lo, prims[indices[i]].bbox_min);
            hi = vmax(hi, prims[indices[i]].bbox_max);
        }
        rh_bvh_node &n = out[node];
        n.prim_index = RH_LEAF_NODE_BIT + first + index_offset;
        n.prim_count = count;
        memcpy(n.bbox_min, lo.v, 3 * sizeof(float));
        memcpy(n.bbox_max, hi.v, 3 * sizeof(float));
        return;
    }
    // codes are sorted and agree above `bit`: the range is 0...0 1...1 in it
    const uint32_t split =
        (uint32_t)(std::partition_point(codes + first, codes + last + 1,
                                        [bit](uint32_t c) { return ((c >> bit) & 1u) == 0u; }) - codes);
    uint32_t left = (uint32_t)out.size(), right = left + 1;
    out.resize(out.size() + 2);
    memset(&out[left], 0, 2 * sizeof(rh_bvh_node));
    emit_lbvh_node(prims, indices, codes, left, first, split - first, index_offset, bit - 1, out);
    emit_lbvh_node(prims, indices, codes, right, split, last + 1 - split, index_offset, bit - 1, out);
    rh_bvh_node &n = out[node];
    for (int a = 0; a < 3; a++) {
        n.bbox_min[a] = fmin_std(out[left].bbox_min[a], out[right].bbox_min[a]);
        n.bbox_max[a] = fmax_std(out[left].bbox_max[a], out[right].bbox_max[a]);
    }
    const uint32_t axis = (uint32_t)bit % 3;
    if (out[left].bbox_min[axis] > out[right].bbox_min[axis]) std::swap(left, right);
    n.prim_index = left;
    n.prim_count = right + (axis << 30);
}

uint32_t EmitLBVH(const prim_t *prims, const uint32_t *indices, const uint32_t *codes, uint32_t first, uint32_t count,
                  uint32_t index_offset, int bit, std::vector<rh_bvh_node> &out) {
    // Depth first: a node splits its (code-sorted) range where the highest
    // bit that differs inside it turns from 0 to 1; the two children are
    // appended together, the lower half is built first.  Ranges of fewer
    // than 8 primitives, or with identical codes, become leaves.
    const uint32_t node = (uint32_t)out.size();
    out.emplace_back();
    memset(&out[node], 0, sizeof(rh_bvh_node));
    emit_lbvh_node(prims, indices, codes, node, first, count, index_offset, bit, out);
    return node;
}

uint32_t AssembleHLBVH(const std::vector<HlbvhTreelet> &treelets, const std::vector<rh_bvh_node> &bottom,
                       const std::vector<rh_bvh_node> &top, const std::vector<uint32_t> &top_indices,
                       std::vector<rh_bvh_node> &out_nodes) {
    // The top tree's leaves (one treelet each) are dropped and its interior
    // nodes point at the treelets' roots instead; the treelets' nodes follow
    // the top tree's interior nodes.
    const uint32_t base = (uint32_t)out_nodes.size();
    std::vector<uint32_t> dropped_before(top.size() + 1, 0);
    for (size_t i = 0; i < top.size(); i++)
        dropped_before[i + 1] = dropped_before[i] + ((top[i].prim_index & RH_LEAF_NODE_BIT) ? 1u : 0u);
    const uint32_t interior = (uint32_t)top.size() - dropped_before[top.size()];
    const uint32_t bottom_base = base + interior;
    auto child_ref = [&](uint32_t c) -> uint32_t {
        const rh_bvh_node &ch = top[c];
        if (ch.prim_index & RH_LEAF_NODE_BIT)
            return bottom_base + treelets[top_indices[ch.prim_index & RH_PRIM_INDEX_BITS]].node_index;
        return base + c - dropped_before[c];
    };
    for (size_t i = 0; i < top.size(); i++) {
        rh_bvh_node n = top[i];
        if (n.prim_index & RH_LEAF_NODE_BIT) continue;
        n.prim_index = child_ref(n.prim_index);
        n.prim_count = (n.prim_count & ~RH_RIGHT_CHILD_BITS) + child_ref(n.prim_count & RH_RIGHT_CHILD_BITS);
        out_nodes.push_back(n);
    }
    for (rh_bvh_node n : bottom) {
        if (!(n.prim_index & RH_LEAF_NODE_BIT)) {
            n.prim_index += bottom_base;
            n.prim_count += bottom_base;
        }
        out_nodes.push_back(n);
    }
    return (uint32_t)(out_nodes.size() - base);
}

void TopLevelSAH(const std::vector<HlbvhTreelet> &treelets, const std::vector<rh_bvh_node> &bottom,
                 std::vector<rh_bvh_node> &top, std::vector<uint32_t> &top_indices, const SahBoxes *sah) {
    std::vector<prim_t> top_prims(treelets.size());
    for (size_t i = 0; i < treelets.size(); i++) {
        const rh_bvh_node &r = bottom[treelets[i].node_index];
        top_prims[i] = prim_t{0, 0, 0, load3(r.bbox_min), load3(r.bbox_max)};
    }
    // every treelet its own leaf
    bvh_settings_t s;
    s.oversplit_threshold = std::numeric_limits<float>::max();
    s.node_traversal_cost = 0.0f;
    s.allow_spatial_splits = false;
    PreprocessPrims_SAH_Boxes(top_prims.data(), top_prims.size(), s, top, top_indices, sah);
}

uint32_t PreprocessPrims_HLBVH(const prim_t *prims, size_t prims_count, std::vector<rh_bvh_node> &out_nodes,
                               std::vector<uint32_t> &out_indices, const LbvhBottom *device, const SahBoxes *sah) {
    if (prims_count == 0) return 0;
    const uint32_t indices_start = (uint32_t)out_indices.size();
    if (device) {
        std::vector<uint32_t> order;
        std::vector<HlbvhTreelet> treelets;
        std::vector<rh_bvh_node> bottom, top;
        std::vector<uint32_t> top_indices;
        (*device)(prims, prims_count, indices_start, order, bottom, treelets);
        out_indices.insert(out_indices.end(), order.begin(), order.end());
        TopLevelSAH(treelets, bottom, top, top_indices, sah);
        return AssembleHLBVH(treelets, bottom, top, top_indices, out_nodes);
    }
    std::vector<uint32_t> codes(prims_count), order(prims_count);
    MortonCodes(prims, prims_count, codes.data());
    for (uint32_t j = 0; j < (uint32_t)prims_count; j++) order[j] = j;
    // stable: equal codes keep their primitive order
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
    std::vector<uint32_t> sorted(prims_count);
    for (size_t j = 0; j < prims_count; j++) sorted[j] = codes[order[j]];
    out_indices.insert(out_indices.end(), order.begin(), order.end());

    std::vector<HlbvhTreelet> treelets;
    std::vector<rh_bvh_node> bottom;
    for (uint32_t start = 0; start < (uint32_t)prims_count;) {
        uint32_t end = start + 1;
        while (end < (uint32_t)prims_count && (sorted[end] & kTreeletMask) == (sorted[start] & kTreeletMask)) end++;
        treelets.push_back({start, end - start, 0});
        start = end;
    }
    for (HlbvhTreelet &t : treelets)
        t.node_index = EmitLBVH(prims, &out_indices[indices_start], sorted.data(), t.first, t.count, indices_start,
                                kTreeletBit, bottom);
    std::vector<rh_bvh_node> top;
    std::vector<uint32_t> top_indices;
    TopLevelSAH(treelets, bottom, top, top_indices);
    return AssembleHLBVH(treelets, bottom, top, top_indices, out_nodes);
}

void ConstructCamera(int type, int filter, int dtype, const float origin[3], const float fwd[3], const float up[3],
                     float fov, float gamma, float focus_distance, float focus_factor, rh_camera *cam) {
    if (type == RH_CAM_PERSP) {
        vec3 o = load3(origin), f = load3(fwd), u = load3(up);
        if (dot3(u, u) < RH_FLT_EPS) {
            if (std::abs(f[1]) >= 0.999f) {
                u = {{1.0f, 0.0f, 0.0f}};
            } else {
                u = {{0.0f, 1.0f, 0.0f}};
            }
        }
        const vec3 s = normalize3(cross3(f, u));
        u = cross3(s, f);

        cam->type = type;
        cam->filter = filter;
        cam->dtype = dtype;
        cam->fov = fov;
        cam->gamma = gamma;
        cam->focus_distance = focus_distance;
        cam->focus_factor = focus_factor;
        memcpy(cam->origin, o.v, 3 * sizeof(float));
        memcpy(cam->fwd, f.v, 3 * sizeof(float));
        memcpy(cam->side, s.v, 3 * sizeof(float));
        memcpy(cam->up, u.v, 3 * sizeof(float));
    }
}

void TransformBoundingBox(const float bbox_min[3], const float bbox_max[3], const float *xform, float out_min[3],
                          float out_max[3]) {
    out_min[0] = out_max[0] = xform[12];
    out_min[1] = out_max[1] = xform[13];
    out_min[2] = out_max[2] = xform[14];
    for (int j = 0; j < 3; j++) {
        for (int i = 0; i < 3; i++) {
            const float a = xform[i * 4 + j] * bbox_min[i];
            const float b = xform[i * 4 + j] * bbox_max[i];
            if (a < b) {
                out_min[j] += a;
                out_max[j] += b;
            } else {
                out_min[j] += b;
                out_max[j] += a;
            }
        }
    }
}

// Cofactor expansion with the reference's term order (Core.cpp:909-950): the
// inverse feeds the object-space rays, so it must round identically.
void InverseMatrix(const float m[16], float o[16]) {
    const float A2323 = m[10] * m[15] - m[11] * m[14];
    const float A1323 = m[9] * m[15] - m[11] * m[13];
    const float A1223 = m[9] * m[14] - m[10] * m[13];
    const float A0323 = m[8] * m[15] - m[11] * m[12];
    const float A0223 = m[8] * m[14] - m[10] * m[12];
    const float A0123 = m[8] * m[13] - m[9] * m[12];
    const float A2313 = m[6] * m[15] - m[7] * m[14];
    const float A1313 = m[5] * m[15] - m[7] * m[13];
    const float A1213 = m[5] * m[14] - m[6] * m[13];
    const float A2312 = m[6] * m[11] - m[7] * m[10];
    const float A1312 = m[5] * m[11] - m[7] * m[9];
    const float A1212 = m[5] * m[10] - m[6] * m[9];
    const float A0313 = m[4] * m[15] - m[7] * m[12];
    const float A0213 = m[4] * m[14] - m[6] * m[12];
    const float A0312 = m[4] * m[11] - m[7] * m[8];
    const float A0212 = m[4] * m[10] - m[6] * m[8];
    const float A0113 = m[4] * m[13] - m[5] * m[12];
    const float A0112 = m[4] * m[9] - m[5] * m[8];

    const float inv_det = 1.0f / (m[0] * (m[5] * A2323 - m[6] * A1323 + m[7] * A1223) -
                                  m[1] * (m[4] * A2323 - m[6] * A0323 + m[7] * A0223) +
                                  m[2] * (m[4] * A1323 - m[5] * A0323 + m[7] * A0123) -
                                  m[3] * (m[4] * A1223 - m[5] * A0223 + m[6] * A0123));

    o[0] = inv_det * (m[5] * A2323 - m[6] * A1323 + m[7] * A1223);
    o[1] = inv_det * -(m[1] * A2323 - m[2] * A1323 + m[3] * A1223);
    o[2] = inv_det * (m[1] * A2313 - m[2] * A1313 + m[3] * A1213);
    o[3] = inv_det * -(m[1] * A2312 - m[2] * A1312 + m[3] * A1212);
    o[4] = inv_det * -(m[4] * A2323 - m[6] * A0323 + m[7] * A0223);
    o[5] = inv_det * (m[0] * A2323 - m[2] * A0323 + m[3] * A0223);
    o[6] = inv_det * -(m[0] * A2313 - m[2] * A0313 + m[3] * A0213);
    o[7] = inv_det * (m[0] * A2312 - m[2] * A0312 + m[3] * A0212);
    o[8] = inv_det * (m[4] * A1323 - m[5] * A0323 + m[7] * A0123);
    o[9] = inv_det * -(m[0] * A1323 - m[1] * A0323 + m[3] * A0123);
    o[10] = inv_det * (m[0] * A1313 - m[1] * A0313 + m[3] * A0113);
    o[11] = inv_det * -(m[0] * A1312 - m[1] * A0312 + m[3] * A0112);
    o[12] = inv_det * -(m[4] * A1223 - m[5] * A0223 + m[6] * A0123);
    o[13] = inv_det * (m[0] * A1223 - m[1] * A0223 + m[2] * A0123);
    o[14] = inv_det * -(m[0] * A1213 - m[1] * A0213 + m[2] * A0113);
    o[15] = inv_det * (m[0] * A1212 - m[1] * A0212 + m[2] * A0112);
}

void ComputeTangentBasis(size_t vtx_offset, size_t vtx_start, std::vector<rh_vertex> &vertices,
                         std::vector<uint32_t> &new_vtx_indices, const uint32_t *indices, size_t indices_count) {
    // per-vertex "twin" copies by binormal-sign class, and accumulated binormals
    std::vector<std::array<uint32_t, 3>> twin_verts(vertices.size(), {{0, 0, 0}});
    std::vector<vec3> binormals(vertices.size(), vec3{{0.0f, 0.0f, 0.0f}});

    for (size_t i = 0; i < indices_count; i += 3) {
        rh_vertex *v0 = &vertices[indices[i + 0]];
        rh_vertex *v1 = &vertices[indices[i + 1]];
        rh_vertex *v2 = &vertices[indices[i + 2]];

        vec3 &b0 = binormals[indices[i + 0]];
        vec3 &b1 = binormals[indices[i + 1]];
        vec3 &b2 = binormals[indices[i + 2]];

        const vec3 dp1 = {{v1->p[0] - v0->p[0], v1->p[1] - v0->p[1], v1->p[2] - v0->p[2]}};
        const vec3 dp2 = {{v2->p[0] - v0->p[0], v2->p[1] - v0->p[1], v2->p[2] - v0->p[2]}};
        const float dt1[2] = {v1->t[0][0] - v0->t[0][0], v1->t[0][1] - v0->t[0][1]};
        const float dt2[2] = {v2->t[0][0] - v0->t[0][0], v2->t[0][1] - v0->t[0][1]};

        vec3 tangent, binormal;
        const float det = dt1[0] * dt2[1] - dt1[1] * dt2[0];
        if (std::abs(det) > RH_FLT_EPS) {
            const float inv_det = 1.0f / det;
            for (int k = 0; k < 3; k++) {
                tangent[k] = (dp1[k] * dt2[1] - dp2[k] * dt1[1]) * inv_det;
                binormal[k] = (dp2[k] * dt1[0] - dp1[k] * dt2[0]) * inv_det;
            }
        } else {
            const vec3 plane_N = cross3(dp1, dp2);
            int w = 2;
            tangent = {{0.0f, 1.0f, 0.0f}};
            if (std::abs(plane_N[0]) <= std::abs(plane_N[1]) && std::abs(plane_N[0]) <= std::abs(plane_N[2])) {
                tangent = {{1.0f, 0.0f, 0.0f}};
                w = 1;
            } else if (std::abs(plane_N[2]) <= std::abs(plane_N[0]) && std::abs(plane_N[2]) <= std::abs(plane_N[1])) {
                tangent = {{0.0f, 0.0f, 1.0f}};
                w = 0;
            }
            if (std::abs(plane_N[w]) > RH_FLT_EPS) {
                binormal = normalize3(cross3(plane_N, tangent));
                tangent = normalize3(cross3(plane_N, binormal));
            } else {
                binormal = {{0.0f, 0.0f, 0.0f}};
                tangent = {{0.0f, 0.0f, 0.0f}};
            }
        }

        // corner 0
        {
            const int i1 = (v0->b[0] * tangent[0] + v0->b[1] * tangent[1] + v0->b[2] * tangent[2]) < 0;
            const int i2 = 2 * (b0[0] * binormal[0] + b0[1] * binormal[1] + b0[2] * binormal[2] < 0);
            if (i1 || i2) {
                uint32_t index = twin_verts[indices[i + 0]][i1 + i2 - 1];
                if (index == 0) {
                    index = (uint32_t)(vtx_offset + vertices.size());
                    vertices.push_back(*v0);
                    memset(&vertices.back().b[0], 0, 3 * sizeof(float));
                    twin_verts[indices[i + 0]][i1 + i2 - 1] = index;
                    v1 = &vertices[indices[i + 1]];
                    v2 = &vertices[indices[i + 2]];
                }
                new_vtx_indices[i] = index;
                v0 = &vertices[index - vtx_offset];
            } else {
                b0 = binormal;
            }
            v0->b[0] += tangent[0];
            v0->b[1] += tangent[1];
            v0->b[2] += tangent[2];
        }
        // corner 1
        {
            const int i1 = v1->b[0] * tangent[0] + v1->b[1] * tangent[1] + v1->b[2] * tangent[2] < 0;
            const int i2 = 2 * (b1[0] * binormal[0] + b1[1] * binormal[1] + b1[2] * binormal[2] < 0);
            if (i1 || i2) {
                uint32_t index = twin_verts[indices[i + 1]][i1 + i2 - 1];
                if (index == 0) {
                    index = (uint32_t)(vtx_offset + vertices.size());
                    vertices.push_back(*v1);
                    memset(&vertices.back().b[0], 0, 3 * sizeof(float));
                    twin_verts[indices[i + 1]][i1 + i2 - 1] = index;
                    v0 = &vertices[indices[i + 0]];
                    v2 = &vertices[indices[i + 2]];
                }
                new_vtx_indices[i + 1] = index;
                v1 = &vertices[index - vtx_offset];
            } else {
                b1 = binormal;
            }
            v1->b[0] += tangent[0];
            v1->b[1] += tangent[1];
            v1->b[2] += tangent[2];
        }
        // corner 2
        {
            const int i1 = v2->b[0] * tangent[0] + v2->b[1] * tangent[1] + v2->b[2] * tangent[2] < 0;
            const int i2 = 2 * (b2[0] * binormal[0] + b2[1] * binormal[1] + b2[2] * binormal[2] < 0);
            if (i1 || i2) {
                uint32_t index = twin_verts[indices[i + 2]][i1 + i2 - 1];
                if (index == 0) {
                    index = (uint32_t)(vtx_offset + vertices.size());
                    vertices.push_back(*v2);
                    memset(&vertices.back().b[0], 0, 3 * sizeof(float));
                    twin_verts[indices[i + 2]][i1 + i2 - 1] = index;
                    v0 = &vertices[indices[i + 0]];
                    v1 = &vertices[indices[i + 1]];
                }
                new_vtx_indices[i + 2] = index;
                v2 = &vertices[index - vtx_offset];
            } else {
                b2 = binormal;
            }
            v2->b[0] += tangent[0];
            v2->b[1] += tangent[1];
            v2->b[2] += tangent[2];
        }
    }

    for (size_t i = vtx_start; i < vertices.size(); i++) {
        rh_vertex &v = vertices[i];
        if (std::abs(v.b[0]) > RH_FLT_EPS || std::abs(v.b[1]) > RH_FLT_EPS || std::abs(v.b[2]) > RH_FLT_EPS) {
            const vec3 tangent = load3(v.b);
            vec3 binormal = cross3(load3(v.n), tangent);
            const float l = length3(binormal);
            if (l > RH_FLT_EPS) {
                binormal = div3(binormal, l);
                memcpy(&v.b[0], binormal.v, 3 * sizeof(float));
            }
        }
    }
}

// 2x2 box filter of TextureUtilsRef.cpp:9-37, over the row-major texels with
// the reference's indexing: for an odd width the last pair of a row takes the
// next row's first texel, and for an odd size the output holds ceil(w/2) x
// ceil(h/2) texels while the next level is read as floor(w/2) x floor(h/2),
// as in the reference.  Where those reads run past the end of the input (the
// last row pair of an odd height) the reference reads outside its vector; here
// such a read takes the texel one row up instead (the last row repeated).
std::vector<uint32_t> DownsampleTexture(const std::vector<uint32_t> &tex, const int res[2]) {
    if (res[0] == 1 || res[1] == 1) return tex;
    std::vector<uint32_t> ret;
    ret.reserve((size_t)res[0] * res[1] / 4);
    auto ch = [](uint32_t p, int c) { return (int)((p >> (8 * c)) & 0xff); };
    const size_t n = tex.size(), row = (size_t)res[0];
    auto at = [&](size_t k) {
        while (k >= n && k >= row) k -= row;
        return k < n ? tex[k] : 0u;
    };
    for (int j = 0; j < res[1]; j += 2) {
        for (int i = 0; i < res[0]; i += 2) {
            const uint32_t a = at((size_t)(j + 0) * res[0] + i), b = at((size_t)(j + 0) * res[0] + i + 1),
                           c = at((size_t)(j + 1) * res[0] + i), d = at((size_t)(j + 1) * res[0] + i + 1);
            uint32_t out = 0;
            for (int k = 0; k < 4; k++) {
                const int s = ch(a, k) + ch(b, k) + ch(c, k) + ch(d, k);
                out |= (uint32_t)(uint8_t)(s / 4) << (8 * k);
            }
            ret.push_back(out);
        }
    }
    return ret;
}

std::vector<uint16_t> ComputeRadicalInversePermutations() {
    static const int primes[RH_HALTON_COUNT] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29, 31, 37,
                                                41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89};
    auto rand_func = std::bind(std::uniform_int_distribution<int>(), std::mt19937(0));
    int total = 0;
    for (int p : primes) total += p;
    std::vector<uint16_t> perms((size_t)total);
    uint16_t *p = perms.data();
    for (int k = 0; k < RH_HALTON_COUNT; k++) {
        const int count = primes[k];
        for (int j = 0; j < count; j++) p[j] = (uint16_t)j;
        // rejection-sampled shuffle (Halton.h:6-20)
        for (int i = 0; i < count; i++) {
            int other;
            const uint32_t thres = (~((uint32_t)count) + 1u) % count;
            while (true) {
                const uint32_t r = rand_func();
                if (r >= thres) {
                    other = r % count;
                    break;
                }
            }
            std::swap(p[i], p[other]);
        }
        p += count;
    }
    return perms;
}

namespace {
template <int base> float ScrambledRadicalInverse(const uint16_t *perm, uint64_t a) {
    const float inv_base = 1.0f / base;
    uint64_t reversed_digits = 0;
    float inv_base_n = 1;
    while (a) {
        const uint64_t next = a / base;
        const uint64_t digit = a - next * base;
        reversed_digits = reversed_digits * base + perm[digit];
        inv_base_n *= inv_base;
        a = next;
    }
    return std::min(inv_base_n * (reversed_digits + inv_base * perm[0] / (1 - inv_base)),
                    1.0f - std::numeric_limits<float>::epsilon());
}
} // namespace

void UpdateHaltonSequence(const std::vector<uint16_t> &perms, int iteration, float *seq) {
    const uint16_t *p = perms.data();
    for (int i = 0; i < RH_HALTON_SEQ_LEN; i++) {
        const uint64_t a = (uint64_t)(iteration + i);
        float *row = seq + i * RH_HALTON_COUNT;
        row[0] = ScrambledRadicalInverse<2>(p + 0, a);
        row[1] = ScrambledRadicalInverse<3>(p + 2, a);
        row[2] = ScrambledRadicalInverse<5>(p + 5, a);
        row[3] = ScrambledRadicalInverse<7>(p + 10, a);
        row[4] = ScrambledRadicalInverse<11>(p + 17, a);
        row[5] = ScrambledRadicalInverse<13>(p + 28, a);
        row[6] = ScrambledRadicalInverse<17>(p + 41, a);
        row[7] = ScrambledRadicalInverse<19>(p + 58, a);
        row[8] = ScrambledRadicalInverse<23>(p + 77, a);
        row[9] = ScrambledRadicalInverse<29>(p + 100, a);
        row[10] = ScrambledRadicalInverse<31>(p + 129, a);
        row[11] = ScrambledRadicalInverse<37>(p + 160, a);
        row[12] = ScrambledRadicalInverse<41>(p + 197, a);
        row[13] = ScrambledRadicalInverse<43>(p + 238, a);
        row[14] = ScrambledRadicalInverse<47>(p + 281, a);
        row[15] = ScrambledRadicalInverse<53>(p + 328, a);
        row[16] = ScrambledRadicalInverse<59>(p + 381, a);
        row[17] = ScrambledRadicalInverse<61>(p + 440, a);
        row[18] = ScrambledRadicalInverse<67>(p + 501, a);
        row[19] = ScrambledRadicalInverse<71>(p + 568, a);
        row[20] = ScrambledRadicalInverse<73>(p + 639, a);
        row[21] = ScrambledRadicalInverse<79>(p + 712, a);
        row[22] = ScrambledRadicalInverse<83>(p + 791, a);
        row[23] = ScrambledRadicalInverse<89>(p + 874, a);
    }
}

TextureSplitter::TextureSplitter(const int res[2]) {
    nodes_.emplace_back();
    nodes_.back().size[0] = res[0];
    nodes_.back().size[1] = res[1];
}

int TextureSplitter::Allocate(const int res[2], int pos[2]) {
    const int i = Insert_Recursive(0, res);
    if (i != -1) {
        pos[0] = nodes_[i].pos[0];
        pos[1] = nodes_[i].pos[1];
    }
    return i;
}

int TextureSplitter::Insert_Recursive(int i, const int res[2]) {
    if (!nodes_[i].is_free || res[0] > nodes_[i].size[0] || res[1] > nodes_[i].size[1]) return -1;

    int ch0 = nodes_[i].child[0], ch1 = nodes_[i].child[1];
    if (ch0 != -1) {
        const int new_node = Insert_Recursive(ch0, res);
        if (new_node != -1) return new_node;
        return Insert_Recursive(ch1, res);
    }
    if (res[0] == nodes_[i].size[0] && res[1] == nodes_[i].size[1]) {
        nodes_[i].is_free = false;
        return i;
    }
    nodes_[i].child[0] = ch0 = (int)nodes_.size();
    nodes_.emplace_back();
    nodes_[i].child[1] = ch1 = (int)nodes_.size();
    nodes_.emplace_back();

    const node_t &n = nodes_[i];
    const int dw = n.size[0] - res[0], dh = n.size[1] - res[1];
    if (dw > dh) {
        nodes_[ch0].pos[0] = n.pos[0];
        nodes_[ch0].pos[1] = n.pos[1];
        nodes_[ch0].size[0] = res[0];
        nodes_[ch0].size[1] = n.size[1];
        nodes_[ch1].pos[0] = n.pos[0] + res[0];
        nodes_[ch1].pos[1] = n.pos[1];
        nodes_[ch1].size[0] = n.size[0] - res[0];
        nodes_[ch1].size[1] = n.size[1];
    } else {
        nodes_[ch0].pos[0] = n.pos[0];
        nodes_[ch0].pos[1] = n.pos[1];
        nodes_[ch0].size[0] = n.size[0];
        nodes_[ch0].size[1] = res[1];
        nodes_[ch1].pos[0] = n.pos[0];
        nodes_[ch1].pos[1] = n.pos[1] + res[1];
        nodes_[ch1].size[0] = n.size[0];
        nodes_[ch1].size[1] = n.size[1] - res[1];
    }
    nodes_[ch0].parent = nodes_[ch1].parent = i;
    return Insert_Recursive(ch0, res);
}

// ---- 8-wide trees ----
namespace {
bool is_leaf(const rh_bvh_node &n) { return (n.prim_index & RH_LEAF_NODE_BIT) != 0; }
uint32_t right_of(const rh_bvh_node &n) { return n.prim_count & RH_RIGHT_CHILD_BITS; }

// the (up to 8) binary nodes that become the children of the wide node made
// for interior node n, in left-to-right order
int wide_children(const rh_bvh_node *nodes, const rh_bvh_node &n, uint32_t out[8]) {
    int count = 0;
    const uint32_t halves[2] = {n.prim_index, right_of(n)};
    for (uint32_t h : halves) {
        const rh_bvh_node &c = nodes[h];
        if (is_leaf(c)) {
            out[count++] = h;
            continue;
        }
        const uint32_t quarters[2] = {c.prim_index, right_of(c)};
        for (uint32_t q : quarters) {
            const rh_bvh_node &g = nodes[q];
            if (is_leaf(g)) {
                out[count++] = q;
            } else {
                out[count++] = g.prim_index;
                out[count++] = right_of(g);
            }
        }
    }
    return count;
}
} // namespace

uint32_t FlattenBVH(const rh_bvh_node *nodes, uint32_t root, std::vector<rh_mbvh_node> &out) {
    const rh_bvh_node &n = nodes[root];
    const uint32_t self = (uint32_t)out.size();
    out.emplace_back();
    if (is_leaf(n)) {
        rh_mbvh_node &w = out[self];
        memset(&w, 0, sizeof(w));
        for (int a = 0; a < 3; a++) {
            w.bbox_min[a][0] = n.bbox_min[a];
            w.bbox_max[a][0] = n.bbox_max[a];
        }
        w.child[0] = n.prim_index;
        w.child[1] = n.prim_count;
        return self;
    }
    uint32_t kids[8];
    const int count = wide_children(nodes, n, kids);

    // slot = Morton code of the child's box centre on a 2x2x2 grid over the
    // centres' bounding box (top widened by 0.001)
    vec3 centre[8];
    vec3 lo = {{FLT_MAX, FLT_MAX, FLT_MAX}}, hi = {{-FLT_MAX, -FLT_MAX, -FLT_MAX}};
    for (int i = 0; i < count; i++) {
        const rh_bvh_node &k = nodes[kids[i]];
        for (int a = 0; a < 3; a++) centre[i][a] = 0.5f * (k.bbox_min[a] + k.bbox_max[a]);
        lo = vmin(lo, centre[i]);
        hi = vmax(hi, centre[i]);
    }
    float scale[3];
    for (int a = 0; a < 3; a++) scale[a] = 2.0f / ((hi[a] + 0.001f) - lo[a]);
    uint32_t slot_of[8];
    for (int s = 0; s < 8; s++) slot_of[s] = 0xffffffffu;
    for (int i = 0; i < count; i++) {
        uint32_t code = 0;
        for (int a = 0; a < 3; a++) code |= (uint32_t)((centre[i][a] - lo[a]) * scale[a]) << a;
        code %= 8; // a coordinate of 2 (rounding at huge extents) stays in the table
        while (slot_of[code] != 0xffffffffu) code = (code + 1) % 8;
        slot_of[code] = kids[i];
    }
    uint32_t child[8];
    for (int s = 0; s < 8; s++) {
        child[s] = slot_of[s] != 0xffffffffu ? FlattenBVH(nodes, slot_of[s], out) : RH_MBVH_EMPTY;
    }
    rh_mbvh_node &w = out[self]; // out may have grown: take the reference now
    for (int s = 0; s < 8; s++) {
        w.child[s] = child[s];
        for (int a = 0; a < 3; a++) {
            if (child[s] != RH_MBVH_EMPTY) {
                w.bbox_min[a][s] = nodes[slot_of[s]].bbox_min[a];
                w.bbox_max[a][s] = nodes[slot_of[s]].bbox_max[a];
            } else {
                w.bbox_min[a][s] = w.bbox_max[a][s] = NAN;
            }
        }
    }
    return self;
}

void WideRootBounds(const rh_bvh_node *nodes, uint32_t root, float bmin[3], float bmax[3]) {
    const rh_bvh_node &n = nodes[root];
    if (is_leaf(n)) {
        for (int a = 0; a < 3; a++) {
            bmin[a] = n.bbox_min[a];
            bmax[a] = n.bbox_max[a];
        }
        return;
    }
    uint32_t kids[8];
    const int count = wide_children(nodes, n, kids);
    for (int a = 0; a < 3; a++) {
        bmin[a] = RH_MAX_DIST;
        bmax[a] = -RH_MAX_DIST;
    }
    for (int i = 0; i < count; i++) {
        for (int a = 0; a < 3; a++) {
            bmin[a] = std::min(bmin[a], nodes[kids[i]].bbox_min[a]);
            bmax[a] = std::max(bmax[a], nodes[kids[i]].bbox_max[a]);
        }
    }
}

} // namespace RayHip
