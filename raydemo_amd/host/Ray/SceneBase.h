// Scene interface of the Ray renderer API, declared for the HIP backend.
// Mirrors the reference's src/Ray/SceneBase.h:15-285 (descriptor structs and
// the SceneBase virtual interface, with the non-virtual camera store).
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include "Types.h"

namespace Ray {
enum ePrimType { TriangleList };

enum eVertexLayout { PxyzNxyzTuv = 0, PxyzNxyzTuvTuv, PxyzNxyzBxyzTuv, PxyzNxyzBxyzTuvTuv };
const size_t AttrStrides[] = {8, 10, 11, 13};

enum eMaterialType { DiffuseMaterial, GlossyMaterial, RefractiveMaterial, EmissiveMaterial, MixMaterial, TransparentMaterial };

struct mat_desc_t {
    eMaterialType type;
    float main_color[3] = {1, 1, 1};
    uint32_t main_texture;
    uint32_t normal_map = 0xffffffff;
    uint32_t mix_materials[2] = {0xffffffff};
    float roughness = 0;
    float strength = 1;
    float fresnel = 1;
    float int_ior = 1;
    float ext_ior = 1;
};

struct shape_desc_t {
    uint32_t mat_index;
    uint32_t back_mat_index;
    size_t vtx_start;
    size_t vtx_count;
    shape_desc_t(uint32_t m, uint32_t bm, size_t s, size_t c) : mat_index(m), back_mat_index(bm), vtx_start(s), vtx_count(c) {}
    shape_desc_t(uint32_t m, size_t s, size_t c) : mat_index(m), back_mat_index(0xffffffff), vtx_start(s), vtx_count(c) {}
};

struct mesh_desc_t {
    ePrimType prim_type;
    eVertexLayout layout;
    const float *vtx_attrs;
    size_t vtx_attrs_count;
    const uint32_t *vtx_indices;
    size_t vtx_indices_count;
    int base_vertex = 0;
    std::vector<shape_desc_t> shapes;
    bool allow_spatial_splits = false;
    bool use_fast_bvh_build = false;
};

struct tex_desc_t {
    const pixel_color8_t *data;
    int w, h;
    bool is_srgb = true;
    bool generate_mipmaps;
};

enum eLightType { PointLight, SpotLight, DirectionalLight };

struct light_desc_t {
    eLightType type;
    float position[3], radius;
    float color[3];
    float direction[3], angle;
};

struct camera_desc_t {
    eCamType type = Persp;
    eFilterType filter = Tent;
    eDeviceType dtype = SRGB;
    float origin[3];
    float fwd[3] = {};
    float up[3];
    float fov, gamma = 1.0f;
    float focus_distance = 1.0f;
    float focus_factor = 0.0f;
    uint32_t mi_index, uv_index = 0;
    bool lighting_only = false;
    bool skip_direct_lighting = false;
    bool skip_indirect_lighting = false;
    bool no_background = false;
    bool clamp = false;
    bool output_sh = false;
    bool use_coherent_sampling = false;
    uint8_t max_diff_depth = 4;
    uint8_t max_glossy_depth = 4;
    uint8_t max_refr_depth = 4;
    uint8_t max_transp_depth = 4;
    uint8_t max_total_depth = 4;
    uint8_t termination_start_depth = 3;
};

struct environment_desc_t {
    float env_col[3] = {0.0f};
    float env_clamp = 0.0f;
    uint32_t env_map = 0xffffffff;
};

class SceneBase {
protected:
    union cam_storage_t {
        camera_t cam;
        uint32_t next_free;
    };
    std::vector<cam_storage_t> cams_;
    uint32_t cam_first_free_ = 0xffffffff;
    uint32_t current_cam_ = 0xffffffff;

public:
    virtual ~SceneBase() = default;

    virtual void GetEnvironment(environment_desc_t &env) = 0;
    virtual void SetEnvironment(const environment_desc_t &env) = 0;

    virtual uint32_t AddTexture(const tex_desc_t &t) = 0;
    virtual void RemoveTexture(uint32_t i) = 0;

    virtual uint32_t AddMaterial(const mat_desc_t &m) = 0;
    virtual void RemoveMaterial(uint32_t i) = 0;

    virtual uint32_t AddMesh(const mesh_desc_t &m) = 0;
    virtual void RemoveMesh(uint32_t i) = 0;

    virtual uint32_t AddLight(const light_desc_t &l) = 0;
    virtual void RemoveLight(uint32_t i) = 0;

    virtual uint32_t AddMeshInstance(uint32_t m_index, const float *xform) = 0;
    virtual void SetMeshInstanceTransform(uint32_t mi_index, const float *xform) = 0;
    virtual void RemoveMeshInstance(uint32_t mi_index) = 0;

    uint32_t AddCamera(const camera_desc_t &c);
    void GetCamera(uint32_t i, camera_desc_t &c) const;
    void SetCamera(uint32_t i, const camera_desc_t &c);
    void RemoveCamera(uint32_t i);

    uint32_t current_cam() { return current_cam_; }
    void set_current_cam(uint32_t i) { current_cam_ = i; }

    virtual uint32_t triangle_count() = 0;
    virtual uint32_t node_count() = 0;
};
} // namespace Ray
