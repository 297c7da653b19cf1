// Host-side pieces shared by the single-scene context (rh_kernels.hip) and the
// batch context (rh_batch.hip): error reporting behind rh_last_error, device
// buffers, scene upload from an rh_scene_view, and per-pass constants.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "rh_wavefront.h"

namespace rh_rt {

// the calling thread's last error message (one instance for the library)
inline std::string &err() {
    static thread_local std::string s;
    return s;
}

inline int fail(hipError_t e, const char *what) {
    err() = std::string(what) + ": " + hipGetErrorString(e);
    return -(int)e;
}

inline int arg_error(const char *msg) {
    err() = msg;
    return -1001;
}

inline void free_ptr(void *p) {
    if (p) (void)hipFree(p);
}
// frees a device buffer and forgets it: a later failed allocation then leaves
// a null pointer behind, never a freed one
template <class T> inline void release(T *&p) {
    free_ptr((void *)p);
    p = nullptr;
}

// Uploads a 256-row Halton table into a device table of RH_HALTON_TABLE_FLOATS
// floats: the rows, then row 0 again as the extra row.
inline hipError_t upload_halton(float *dev, const float *seq, hipStream_t stream) {
    hipError_t e = hipMemcpyAsync(dev, seq, RH_HALTON_COUNT * RH_HALTON_SEQ_LEN * sizeof(float),
                                  hipMemcpyHostToDevice, stream);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(dev + RH_HALTON_COUNT * RH_HALTON_SEQ_LEN, seq, RH_HALTON_COUNT * sizeof(float),
                          hipMemcpyHostToDevice, stream);
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// The streams of the pass slots: kPassStreams per device (the device current
// when a device's first context asks), created once and shared by every
// context and batch of the process until it ends (all at the default
// priority).  A process gets few hardware queues per device; streams created
// and destroyed by context after context can come to share one queue, which
// serialises passes meant to overlap.  Contexts sharing a stream only add
// order between their work: each orders its own with events it recorded
// before waiting on them.
constexpr int kPassStreams = 4;
inline int pass_streams(int device, hipStream_t out[kPassStreams]) {
    static std::mutex mu;
    static std::vector<std::array<hipStream_t, kPassStreams>> pool;
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0) return arg_error("pass_streams: no device");
    if ((int)pool.size() <= device) pool.resize((size_t)device + 1, std::array<hipStream_t, kPassStreams>{});
    auto &p = pool[(size_t)device];
    for (int j = 0; j < kPassStreams; j++) {
        if (!p[j]) {
            const hipError_t e = hipStreamCreateWithFlags(&p[j], hipStreamNonBlocking);
            if (e != hipSuccess) {
                p[j] = nullptr;
                return fail(e, "hipStreamCreateWithFlags");
            }
        }
        out[j] = p[j];
    }
    return 0;
}

} // namespace rh_rt

#define RH_CHECK(call)                                                                                                 \
    do {                                                                                                               \
        hipError_t _e = (call);                                                                                        \
        if (_e != hipSuccess) return rh_rt::fail(_e, #call);                                                           \
    } while (0)

namespace {

template <typename T>
int upload_array(std::vector<rh_rt::DevBuf> &bufs, hipStream_t stream, const T *src, size_t count, const T **dst) {
    *dst = nullptr;
    if (!src || !count) return 0;
    rh_rt::DevBuf b;
    b.bytes = count * sizeof(T);
    RH_CHECK(hipMalloc(&b.p, b.bytes));
    RH_CHECK(hipMemcpyAsync(b.p, src, b.bytes, hipMemcpyHostToDevice, stream));
    bufs.push_back(b);
    *dst = static_cast<const T *>(b.p);
    return 0;
}

// The device form of the 8-wide trees (rh_device.h, "8-wide trees"): the
// interior nodes (and leaf roots) only, each as 8 slots of two float4 with the
// used slots first, a leaf child's prim range folded into its parent's slot,
// the meshes' and macro tree's roots renumbered; and the wide light tree's
// leaves in the order
// ComputeDirectLighting walks them (CoreRef.cpp:1949-1970: slots pushed in
// slot order, the last popped first).  A light leaf is reached only if the
// point is strictly inside its slot box and its ancestors' (which contain
// it), so testing the leaf boxes alone, in this order, takes the same lights
// in the same order; a leaf root has no box test at all.
inline int build_wide(const rh_scene_view *v, hipStream_t stream, std::vector<rh_rt::DevBuf> &bufs, Scene &s,
                      std::vector<rh_bvh_node> &leaves, std::vector<rh_mesh> &meshes, std::vector<float4> &nodes_out) {
    const uint32_t n = v->mnodes_count;
    auto is_leaf = [&](uint32_t i) { return (v->mnodes[i].child[0] & RH_LEAF_NODE_BIT) != 0; };
    // device nodes: the interior nodes and the tree roots (a leaf child lives in its parent's slot)
    std::vector<uint8_t> root(n, 0);
    for (uint32_t k = 0; k < v->meshes_count; k++) {
        if (v->meshes[k].node_index >= n) return rh_rt::arg_error("mesh tree root outside the wide node array");
        root[v->meshes[k].node_index] = 1;
    }
    if (s.macro_root != RH_INVALID) {
        if (s.macro_root >= n) return rh_rt::arg_error("macro tree root outside the wide node array");
        root[s.macro_root] = 1;
    }
    std::vector<uint32_t> dev(n, RH_INVALID);
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; i++)
        if (!is_leaf(i) || root[i]) dev[i] = m++;
    if (m >= (1u << 28)) return rh_rt::arg_error("too many wide nodes for the device form (2^28)");

    // kWideHalf on the words naming an interior node with at most 4 used slots
    auto half_flag = [&](uint32_t i) {
        int used = 0;
        for (int k = 0; k < 8; k++) used += v->mnodes[i].child[k] != RH_MBVH_EMPTY;
        return used <= 4 ? kWideHalf : 0u;
    };
    // leaf counts in the words (Scene::wide_leaf_counts) while every stacked
    // slot reference (node x 8 + slot) stays below 2^27
    const bool fold_counts = m < (1u << 24);
    s.wide_leaf_counts = fold_counts ? 1u : 0u;
    std::vector<float4> w((size_t)m * 16);
    const float nan = std::numeric_limits<float>::quiet_NaN();
    auto as_float = [](uint32_t u) {
        float f;
        memcpy(&f, &u, 4);
        return f;
    };
    for (uint32_t i = 0; i < n; i++) {
        if (dev[i] == RH_INVALID) continue;
        const rh_mbvh_node &nd = v->mnodes[i];
        float4 *out = &w[(size_t)dev[i] * 16];
        int j = 0; // next device slot: used slots first, in slot order
        for (int k = 0; k < 8; k++) {
            uint32_t word, count = 0;
            if (is_leaf(i)) {
                if (k > 0) break;
                word = nd.child[0];
                count = nd.child[1];
            } else {
                if (nd.child[k] == RH_MBVH_EMPTY) continue;
                const uint32_t c = nd.child[k];
                if (c >= n) return rh_rt::arg_error("wide node child outside the node array");
                if (is_leaf(c)) {
                    word = v->mnodes[c].child[0];
                    count = v->mnodes[c].child[1];
                    // small counts travel in the word too (kLeafCountBits)
                    if (fold_counts && count >= 1 && count <= 15 &&
                        (word & RH_PRIM_INDEX_BITS) < (1u << kLeafCountShift))
                        word |= count << kLeafCountShift;
                } else {
                    word = dev[c] | half_flag(c);
                }
            }
            out[2 * j] = make_float4(nd.bbox_min[0][k], nd.bbox_min[1][k], nd.bbox_min[2][k], as_float(word));
            out[2 * j + 1] = make_float4(nd.bbox_max[0][k], nd.bbox_max[1][k], nd.bbox_max[2][k], as_float(count));
            j++;
        }
        for (; j < 8; j++) {
            out[2 * j] = make_float4(nan, nan, nan, as_float(RH_MBVH_EMPTY));
            out[2 * j + 1] = make_float4(nan, nan, nan, as_float(0u));
        }
    }
    int rc = upload_array(bufs, stream, w.data(), w.size(), &s.wnodes);
    if (rc) return rc;
    nodes_out.swap(w); // the same buffer: it outlives the upload (the caller synchronises)
    // a tree root is named by its node's index, flagged when it is a leaf
    auto root_word = [&](uint32_t i) { return dev[i] | (is_leaf(i) ? kWideLeafRoot : half_flag(i)); };
    meshes.assign(v->meshes, v->meshes + v->meshes_count);
    for (rh_mesh &mesh : meshes) mesh.node_index = root_word(mesh.node_index);
    rc = upload_array(bufs, stream, meshes.data(), meshes.size(), &s.meshes);
    if (rc) return rc;
    if (s.macro_root != RH_INVALID) {
        const rh_mbvh_node &mr = v->mnodes[s.macro_root];
        s.macro_word = is_leaf(s.macro_root) ? mr.child[0] : root_word(s.macro_root);
        s.macro_count = is_leaf(s.macro_root) ? mr.child[1] : 0u;
        s.macro_root = root_word(s.macro_root);
    }

    if (s.light_root == RH_INVALID) return 0;
    if (s.light_root >= n) return rh_rt::arg_error("light tree root outside the node array");
    const rh_mbvh_node &lroot = v->mnodes[s.light_root];
    if (lroot.child[0] & RH_LEAF_NODE_BIT) {
        s.light_untested = 1;
        leaves.push_back(rh_bvh_node{{-RH_MAX_DIST, -RH_MAX_DIST, -RH_MAX_DIST}, lroot.child[0],
                                     {RH_MAX_DIST, RH_MAX_DIST, RH_MAX_DIST}, lroot.child[1]});
        return 0;
    }
    std::vector<std::pair<uint32_t, int>> st; // (node, slot) still to visit
    for (int k = 0; k < 8; k++)
        if (lroot.child[k] != RH_MBVH_EMPTY) st.push_back({s.light_root, k});
    while (!st.empty()) {
        const auto e = st.back();
        st.pop_back();
        const rh_mbvh_node &p = v->mnodes[e.first];
        const uint32_t ci = p.child[e.second];
        if (ci >= n) return rh_rt::arg_error("light tree node outside the node array");
        const rh_mbvh_node &c = v->mnodes[ci];
        if (c.child[0] & RH_LEAF_NODE_BIT) {
            const int k = e.second;
            leaves.push_back(rh_bvh_node{{p.bbox_min[0][k], p.bbox_min[1][k], p.bbox_min[2][k]}, c.child[0],
                                         {p.bbox_max[0][k], p.bbox_max[1][k], p.bbox_max[2][k]}, c.child[1]});
        } else {
            for (int k = 0; k < 8; k++)
                if (c.child[k] != RH_MBVH_EMPTY) st.push_back({ci, k});
        }
    }
    return 0;
}

// Device copies of the scene arrays an instance transform leaves as they are:
// triangles, vertices, materials, textures and lights (appended to bufs,
// described in s).
inline int build_device_static(const rh_scene_view *v, hipStream_t stream, std::vector<rh_rt::DevBuf> &bufs,
                               Scene &s) {
    int rc = 0;
    rc |= upload_array(bufs, stream, v->tris, v->tris_count, &s.tris);
    rc |= upload_array(bufs, stream, v->tri_indices, v->tri_indices_count, &s.tri_indices);
    if (v->tri_indices_count) {
        std::vector<rh_tri_accel> leaf_tris(v->tri_indices_count);
        for (size_t k = 0; k < v->tri_indices_count; k++) {
            if (v->tri_indices[k] >= v->tris_count) return rh_rt::arg_error("triangle index outside the triangle array");
            leaf_tris[k] = v->tris[v->tri_indices[k]];
            leaf_tris[k].mi = v->tri_indices[k]; // the leaf test's hit index in place of the material
        }
        rc |= upload_array(bufs, stream, leaf_tris.data(), leaf_tris.size(), &s.leaf_tris);
        // the upload is asynchronous: the staging vector must outlive it
        if (!rc) RH_CHECK(hipStreamSynchronize(stream));
    }
    rc |= upload_array(bufs, stream, v->vertices, v->vertices_count, &s.vertices);
    rc |= upload_array(bufs, stream, v->vtx_indices, v->vtx_indices_count, &s.vtx_indices);
    rc |= upload_array(bufs, stream, v->materials, v->materials_count, &s.materials);
    rc |= upload_array(bufs, stream, v->textures, v->textures_count, &s.textures);
    rc |= upload_array(bufs, stream, v->tex_mips, (size_t)v->textures_count * RH_NUM_MIP_LEVELS, &s.tex_mips);
    rc |= upload_array(bufs, stream, v->texels, v->texels_count, &s.texels);
    rc |= upload_array(bufs, stream, v->lights, v->lights_count, &s.lights);
    rc |= upload_array(bufs, stream, v->li_indices, v->li_indices_count, &s.li_indices);
    if (rc) return rc;
    if (v->env) s.env = *v->env;
    s.all_solid = 1;
    for (uint32_t i = 0; i < v->tris_count; i++)
        if (!((uint32_t)v->tris[i].ci & RH_TRI_SOLID_BIT)) s.all_solid = 0;
    return 0;
}

// Device copies of the arrays that follow the instances: the node arrays (a
// rebuilt instance tree renumbers them), meshes, instances, transforms, the
// tree roots and the light tree's leaves, listed in the reference's
// depth-first order (right child first) for the device's stackless light walk.
inline int build_device_instances(const rh_scene_view *v, hipStream_t stream, std::vector<rh_rt::DevBuf> &bufs,
                                  Scene &s) {
    s.nodes = nullptr;
    s.meshes = nullptr;
    s.wnodes = nullptr;
    s.light_untested = 0;
    int rc = 0;
    rc |= upload_array(bufs, stream, v->nodes, v->nodes_count, &s.nodes);
    if (!v->mnodes) rc |= upload_array(bufs, stream, v->meshes, v->meshes_count, &s.meshes); // wide: renumbered
    rc |= upload_array(bufs, stream, v->mesh_instances, v->mesh_instances_count, &s.mesh_instances);
    rc |= upload_array(bufs, stream, v->mi_indices, v->mi_indices_count, &s.mi_indices);
    rc |= upload_array(bufs, stream, v->transforms, v->transforms_count, &s.transforms);
    if (rc) return rc;
    s.macro_root = v->macro_nodes_root;
    s.light_root = v->light_nodes_root;
    std::vector<rh_bvh_node> leaves;
    std::vector<rh_mesh> meshes; // as the device walks name their roots
    std::vector<float4> wide_nodes; // host copy of the device wide nodes
    if (v->mnodes) {
        rc = build_wide(v, stream, bufs, s, leaves, meshes, wide_nodes);
        if (rc) return rc;
    } else {
        meshes.assign(v->meshes, v->meshes + v->meshes_count);
    }
    // the instance records (load_instance), in mi_indices order
    std::vector<float4> recs((size_t)v->mi_indices_count * 5);
    for (size_t k = 0; k < v->mi_indices_count; k++) {
        const uint32_t mi = v->mi_indices[k];
        if (mi >= v->mesh_instances_count) return rh_rt::arg_error("instance index outside the instance array");
        const rh_mesh_instance &in = v->mesh_instances[mi];
        if (in.tr_index >= v->transforms_count || in.mesh_index >= meshes.size())
            return rh_rt::arg_error("instance names a transform or mesh outside its array");
        const float *x = v->transforms[in.tr_index].inv_xform;
        uint32_t root = meshes[in.mesh_index].node_index;
        float4 *r = &recs[k * 5];
        r[0] = make_float4(in.bbox_min[0], in.bbox_min[1], in.bbox_min[2], 0.0f);
        r[1] = make_float4(in.bbox_max[0], in.bbox_max[1], in.bbox_max[2], 0.0f);
        memcpy(&r[0].w, &root, 4);
        memcpy(&r[1].w, &mi, 4);
        r[2] = make_float4(x[0], x[4], x[8], x[12]);
        r[3] = make_float4(x[1], x[5], x[9], x[13]);
        r[4] = make_float4(x[2], x[6], x[10], x[14]);
    }
    rc = upload_array(bufs, stream, recs.data(), recs.size(), &s.inst_recs);
    if (rc) return rc;
    // a leaf macro root of one instance: its record in the scene record
    s.macro_single = 0;
    if (v->mnodes && s.macro_root != RH_INVALID && (s.macro_word & RH_LEAF_NODE_BIT) && s.macro_count == 1) {
        const size_t first = s.macro_word & RH_PRIM_INDEX_BITS;
        if (first >= v->mi_indices_count) return rh_rt::arg_error("macro leaf outside the instance list");
        for (int k = 0; k < 5; k++) s.macro_inst[k] = recs[first * 5 + k];
        s.macro_single = 1;
    }
    // and its tree root's device node, when the root is interior
    auto root_of = [](float f) {
        uint32_t u;
        memcpy(&u, &f, 4);
        return u;
    };
    s.root_cached = 0;
    if (s.macro_single && !(root_of(s.macro_inst[0].w) & kWideLeafRoot)) {
        const uint32_t idx = root_of(s.macro_inst[0].w) & kWideIndex;
        if ((size_t)idx >= wide_nodes.size() / 16) return rh_rt::arg_error("instance tree root outside the wide nodes");
        for (int k = 0; k < 16; k++) s.root_node[k] = wide_nodes[(size_t)idx * 16 + k];
        s.root_cached = 1;
    }
    if (!v->mnodes && s.light_root != RH_INVALID && v->nodes && s.light_root < v->nodes_count) {
        std::vector<uint32_t> st{s.light_root};
        while (!st.empty()) {
            const uint32_t cur = st.back();
            st.pop_back();
            if (cur >= v->nodes_count) return rh_rt::arg_error("light tree node outside the node array");
            const rh_bvh_node &n = v->nodes[cur];
            if (n.prim_index & RH_LEAF_NODE_BIT) {
                leaves.push_back(n);
            } else {
                st.push_back(n.prim_index);
                st.push_back(n.prim_count & RH_RIGHT_CHILD_BITS);
            }
        }
    }
    rc = upload_array(bufs, stream, leaves.data(), leaves.size(), &s.light_leaves);
    if (rc) return rc;
    s.light_leaf_count = (uint32_t)leaves.size();
    if (!v->mnodes && s.macro_root != RH_INVALID && (!v->nodes || s.macro_root >= v->nodes_count))
        return rh_rt::arg_error("macro tree root outside the node array");
    // the staging vectors above go out of scope when this returns
    if (hipStreamSynchronize(stream) != hipSuccess) return rh_rt::arg_error("scene upload failed");
    return 0;
}

// Device copies of all the scene arrays of v (appended to bufs) described by out.
inline int build_device_scene(const rh_scene_view *v, hipStream_t stream, std::vector<rh_rt::DevBuf> &bufs,
                              Scene &out) {
    Scene s = {};
    int rc = build_device_static(v, stream, bufs, s);
    if (rc) return rc;
    rc = build_device_instances(v, stream, bufs, s);
    if (rc) return rc;
    out = s;
    return 0;
}

// per-pass constants of a w x h frame; host-side float math as
// GeneratePrimaryRays (CoreRef.cpp:430-431)
inline PassConst make_pass_const(int w, int h, const rh_pass_desc *p) {
    PassConst pc;
    memset(&pc, 0, sizeof(pc));
    pc.cam = p->cam;
    memcpy(pc.rect, p->rect, sizeof(pc.rect));
    pc.w = w;
    pc.h = h;
    pc.iteration = p->iteration;
    pc.aspect = float(w) / h;
    pc.fov_k = std::tan(0.5f * p->cam.fov * RH_PI / 180.0f) * p->cam.focus_distance;
    memcpy(pc.root_min, p->root_min, sizeof(pc.root_min));
    memcpy(pc.cell_size, p->cell_size, sizeof(pc.cell_size));
    pc.mix_k = 1.0f / p->iteration;
    return pc;
}

} // namespace
