// Host-side pieces shared by the single-scene context (rh_kernels.hip) and the
// batch context (rh_batch.hip): error reporting behind rh_last_error, device
// buffers, scene upload from an rh_scene_view, and per-pass constants.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
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

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

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

// Device copies of the scene arrays of v (appended to bufs) described by out.
// The light tree's leaves are listed in the reference's depth-first order
// (right child first) for the device's stackless light walk.
inline int build_device_scene(const rh_scene_view *v, hipStream_t stream, std::vector<rh_rt::DevBuf> &bufs,
                              Scene &out) {
    Scene s = {};
    int rc = 0;
    rc |= upload_array(bufs, stream, v->nodes, v->nodes_count, &s.nodes);
    rc |= upload_array(bufs, stream, v->tris, v->tris_count, &s.tris);
    rc |= upload_array(bufs, stream, v->tri_indices, v->tri_indices_count, &s.tri_indices);
    rc |= upload_array(bufs, stream, v->meshes, v->meshes_count, &s.meshes);
    rc |= upload_array(bufs, stream, v->mesh_instances, v->mesh_instances_count, &s.mesh_instances);
    rc |= upload_array(bufs, stream, v->mi_indices, v->mi_indices_count, &s.mi_indices);
    rc |= upload_array(bufs, stream, v->transforms, v->transforms_count, &s.transforms);
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
    s.macro_root = v->macro_nodes_root;
    s.light_root = v->light_nodes_root;
    std::vector<rh_bvh_node> leaves;
    if (s.light_root != RH_INVALID && v->nodes && s.light_root < v->nodes_count) {
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
    if (s.macro_root != RH_INVALID && (!v->nodes || s.macro_root >= v->nodes_count))
        return rh_rt::arg_error("macro tree root outside the node array");
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
