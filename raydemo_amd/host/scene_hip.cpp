// Ray::Hip::Scene: scene storage and host-side preprocessing for the HIP
// backend.  Record layouts and build order follow the reference's Ref scene
// (src/Ray/internal/SceneRef.cpp) so that triangles, BVH trees and index
// arrays match it; textures are placed in the same virtual atlas positions but
// stored compactly per mip level (see rh_tex_mip).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hip_backend.h"

namespace Ray {

// ---- SceneBase camera store (SceneBase.cpp:8-89) ----
uint32_t SceneBase::AddCamera(const camera_desc_t &c) {
    uint32_t i;
    if (cam_first_free_ == 0xffffffff) {
        i = (uint32_t)cams_.size();
        cams_.emplace_back();
        memset(&cams_.back(), 0, sizeof(cam_storage_t));
    } else {
        i = cam_first_free_;
        cam_first_free_ = cams_[i].next_free;
    }
    SetCamera(i, c);
    if (current_cam_ == 0xffffffff) current_cam_ = i;
    return i;
}

void SceneBase::GetCamera(uint32_t i, camera_desc_t &c) const {
    const camera_t &cam = cams_[i].cam;
    c.type = cam.type;
    c.dtype = cam.dtype;
    c.gamma = cam.gamma;
    if (c.type != Geo) {
        c.filter = cam.filter;
        memcpy(c.origin, cam.origin, 3 * sizeof(float));
        memcpy(c.fwd, cam.fwd, 3 * sizeof(float));
        memcpy(c.up, cam.up, 3 * sizeof(float));
        c.fov = cam.fov;
        c.focus_distance = cam.focus_distance;
        c.focus_factor = cam.focus_factor;
    } else {
        c.mi_index = cam.mi_index;
        c.uv_index = cam.uv_index;
    }
    const uint32_t f = cam.pass_settings.flags;
    c.lighting_only = (f & LightingOnly) != 0;
    c.skip_direct_lighting = (f & SkipDirectLight) != 0;
    c.skip_indirect_lighting = (f & SkipIndirectLight) != 0;
    c.no_background = (f & NoBackground) != 0;
    c.clamp = (f & Clamp) != 0;
    c.output_sh = (f & OutputSH) != 0;
    c.use_coherent_sampling = (f & UseCoherentSampling) != 0;
    c.max_diff_depth = cam.pass_settings.max_diff_depth;
    c.max_glossy_depth = cam.pass_settings.max_glossy_depth;
    c.max_refr_depth = cam.pass_settings.max_refr_depth;
    c.max_transp_depth = cam.pass_settings.max_transp_depth;
    c.max_total_depth = cam.pass_settings.max_total_depth;
    c.termination_start_depth = cam.pass_settings.termination_start_depth;
}

void SceneBase::SetCamera(uint32_t i, const camera_desc_t &c) {
    camera_t &cam = cams_[i].cam;
    if (c.type != Geo) {
        RayHip::ConstructCamera(c.type, c.filter, c.dtype, c.origin, c.fwd, c.up, c.fov, c.gamma, c.focus_distance,
                                c.focus_factor, reinterpret_cast<rh_camera *>(&cam));
    } else {
        cam.type = Geo;
        cam.gamma = c.gamma;
        cam.mi_index = c.mi_index;
        cam.uv_index = c.uv_index;
    }
    uint32_t f = 0;
    if (c.lighting_only) f |= LightingOnly;
    if (c.skip_direct_lighting) f |= SkipDirectLight;
    if (c.skip_indirect_lighting) f |= SkipIndirectLight;
    if (c.no_background) f |= NoBackground;
    if (c.clamp) f |= Clamp;
    if (c.output_sh) f |= OutputSH;
    if (c.use_coherent_sampling) f |= UseCoherentSampling;
    cam.pass_settings.flags = f;
    cam.pass_settings.max_diff_depth = c.max_diff_depth;
    cam.pass_settings.max_glossy_depth = c.max_glossy_depth;
    cam.pass_settings.max_refr_depth = c.max_refr_depth;
    cam.pass_settings.max_transp_depth = c.max_transp_depth;
    cam.pass_settings.max_total_depth = c.max_total_depth;
    cam.pass_settings.termination_start_depth = c.termination_start_depth;
}

void SceneBase::RemoveCamera(uint32_t i) {
    cams_[i].next_free = cam_first_free_;
    cam_first_free_ = i;
}

namespace Hip {

static_assert(sizeof(camera_t) == sizeof(rh_camera), "camera_t layout");

Scene::Scene(bool use_wide_bvh, bool device_build, int device)
    : wide_(use_wide_bvh), device_build_(device_build), build_device_(device) {
    device_bottom_ = DeviceLbvhBottom(&builder_, build_device_);
    device_sah_ = DeviceSahBoxes(&builder_, build_device_);
    const int atlas_res[2] = {RH_TEXTURE_ATLAS_SIZE, RH_TEXTURE_ATLAS_SIZE};
    splitters_.emplace_back(atlas_res);

    // default environment map, white, RGBE exponent 128 (SceneRef.cpp:9-28)
    static const pixel_color8_t default_env_map = {255, 255, 255, 128};
    tex_desc_t t;
    t.data = &default_env_map;
    t.w = t.h = 1;
    t.generate_mipmaps = false;
    default_env_texture_ = AddTexture(t);
    environment_desc_t desc;
    desc.env_col[0] = desc.env_col[1] = desc.env_col[2] = 0.0f;
    desc.env_map = default_env_texture_;
    SetEnvironment(desc);

    // default flat normal map (SceneRef.cpp:30-44)
    static const pixel_color8_t default_normalmap = {127, 127, 255, 0};
    t.data = &default_normalmap;
    default_normals_texture_ = AddTexture(t);
}

void Scene::GetEnvironment(environment_desc_t &env) {
    memcpy(env.env_col, env_.env_col, 3 * sizeof(float));
    env.env_clamp = env_.env_clamp;
    env.env_map = env_.env_map;
}

// The reference stores the indices a caller passes as they are and the
// renderers read through them; an index past the scene's textures or
// materials reads outside the device arrays.  They are checked here instead
// (std::invalid_argument), with the sentinel 0xffffffff allowed where the
// reference gives it a meaning.
void Scene::CheckTexture(uint32_t tex, bool optional, const char *what) const {
    if (optional && tex == 0xffffffff) return;
    if (tex >= textures_.size()) {
        throw std::invalid_argument(std::string(what) + ": texture " + std::to_string(tex) + " does not exist (" +
                                    std::to_string(textures_.size()) + " textures)");
    }
}

void Scene::CheckMaterial(uint32_t mat, bool optional, const char *what) const {
    if (optional && mat == 0xffffffff) return;
    if (mat >= materials_.size()) {
        throw std::invalid_argument(std::string(what) + ": material " + std::to_string(mat) + " does not exist (" +
                                    std::to_string(materials_.size()) + " materials)");
    }
}

void Scene::SetEnvironment(const environment_desc_t &env) {
    CheckTexture(env.env_map, true, "SetEnvironment env_map");
    memcpy(env_.env_col, env.env_col, 3 * sizeof(float));
    env_.env_clamp = env.env_clamp;
    env_.env_map = env.env_map == 0xffffffff ? default_env_texture_ : env.env_map;
    Touch();
}

// TextureAtlasTiled::Allocate placement (TextureAtlasRef.cpp:192-231): first
// page whose splitter fits the bordered block; the page count doubles when
// none does.
int Scene::AllocateInAtlas(const int res[2], int pos[2]) {
    const int bres[2] = {res[0] + 2, res[1] + 2};
    if (bres[0] > RH_TEXTURE_ATLAS_SIZE || bres[1] > RH_TEXTURE_ATLAS_SIZE) return -1;
    while (true) {
        for (int page = 0; page < (int)splitters_.size(); page++) {
            if (splitters_[page].Allocate(bres, pos) != -1) return page;
        }
        const int atlas_res[2] = {RH_TEXTURE_ATLAS_SIZE, RH_TEXTURE_ATLAS_SIZE};
        const size_t n = splitters_.size();
        for (size_t i = 0; i < n; i++) splitters_.emplace_back(atlas_res);
    }
}

// bordered block: texel (x, y) of the block is data[(y-1) mod h][(x-1) mod w],
// the one-texel wrap border the reference writes around every block
void Scene::WriteBlock(const rh_tex_mip &m, const std::vector<uint32_t> &data, const int res[2]) {
    uint32_t *dst = &texels_[m.offset];
    for (int y = 0; y < m.bh; y++) {
        const int sy = (y - 1 + res[1]) % res[1];
        for (int x = 0; x < m.bw; x++) {
            const int sx = (x - 1 + res[0]) % res[0];
            dst[(size_t)y * m.bw + x] = data[(size_t)sy * res[0] + sx];
        }
    }
}

bool Scene::SetTextureData(uint32_t tex, const uint32_t *rgba, uint32_t &first, uint32_t &count) {
    if (tex >= textures_.size()) return false;
    const rh_texture &t = textures_[tex];
    int res[2] = {(int)(t.width & RH_TEXTURE_WIDTH_BITS), (int)t.height};
    std::vector<uint32_t> data(rgba, rgba + (size_t)res[0] * res[1]);
    const rh_tex_mip *mips = &tex_mips_[(size_t)tex * RH_NUM_MIP_LEVELS];
    first = mips[0].offset;
    uint32_t end = first;
    for (int lev = 0; lev < RH_NUM_MIP_LEVELS; lev++) {
        if (lev > 0 && mips[lev].offset == mips[lev - 1].offset) break; // levels past the last repeat it
        if (lev > 0) {
            data = RayHip::DownsampleTexture(data, res);
            res[0] /= 2;
            res[1] /= 2;
        }
        WriteBlock(mips[lev], data, res);
        end = mips[lev].offset + (uint32_t)mips[lev].bw * mips[lev].bh;
    }
    count = end - first;
    return true;
}

uint32_t Scene::AddTexture(const tex_desc_t &_t) {
    const uint32_t tex_index = (uint32_t)textures_.size();
    rh_texture t;
    memset(&t, 0, sizeof(t));
    t.width = (uint16_t)_t.w;
    t.height = (uint16_t)_t.h;
    if (_t.is_srgb) t.width |= RH_TEXTURE_SRGB_BIT;

    std::vector<rh_tex_mip> mips(RH_NUM_MIP_LEVELS);
    int mip = 0;
    int res[2] = {_t.w, _t.h};
    std::vector<uint32_t> data((size_t)_t.w * _t.h);
    memcpy(data.data(), _t.data, data.size() * sizeof(uint32_t));

    while (res[0] >= 1 && res[1] >= 1) {
        int pos[2];
        const int page = AllocateInAtlas(res, pos);
        if (page == -1) return 0xffffffff;
        t.page[mip] = (uint8_t)page;
        t.pos[mip][0] = (uint16_t)pos[0];
        t.pos[mip][1] = (uint16_t)pos[1];

        const int bw = res[0] + 2, bh = res[1] + 2;
        mips[mip].offset = (uint32_t)texels_.size();
        mips[mip].bw = (uint16_t)bw;
        mips[mip].bh = (uint16_t)bh;
        texels_.resize(texels_.size() + (size_t)bw * bh);
        WriteBlock(mips[mip], data, res);

        mip++;
        if (_t.generate_mipmaps) {
            data = RayHip::DownsampleTexture(data, res);
            res[0] /= 2;
            res[1] /= 2;
        } else {
            break;
        }
    }
    for (int i = mip; i < RH_NUM_MIP_LEVELS; i++) {
        t.page[i] = t.page[mip - 1];
        t.pos[i][0] = t.pos[mip - 1][0];
        t.pos[i][1] = t.pos[mip - 1][1];
        mips[i] = mips[mip - 1];
    }
    textures_.push_back(t);
    tex_mips_.insert(tex_mips_.end(), mips.begin(), mips.end());
    Touch();
    return tex_index;
}

uint32_t Scene::AddMaterial(const mat_desc_t &m) {
    if ((int)m.type < (int)DiffuseMaterial || (int)m.type > (int)TransparentMaterial) {
        throw std::invalid_argument("AddMaterial: unknown material type " + std::to_string((int)m.type));
    }
    CheckTexture(m.main_texture, false, "AddMaterial main_texture");
    CheckTexture(m.normal_map, true, "AddMaterial normal_map");
    if (m.type == MixMaterial) {
        CheckMaterial(m.mix_materials[0], false, "AddMaterial mix_materials[0]");
        CheckMaterial(m.mix_materials[1], false, "AddMaterial mix_materials[1]");
    }
    rh_material mat;
    memset(&mat, 0, sizeof(mat));
    mat.type = m.type;
    mat.textures[RH_MAIN_TEXTURE] = m.main_texture;
    memcpy(mat.main_color, m.main_color, 3 * sizeof(float));
    mat.int_ior = m.int_ior;
    mat.ext_ior = m.ext_ior;
    if (m.type == DiffuseMaterial || m.type == GlossyMaterial || m.type == RefractiveMaterial) {
        mat.roughness = m.roughness;
    } else if (m.type == EmissiveMaterial) {
        mat.roughness = m.strength;
    } else if (m.type == MixMaterial) {
        mat.roughness = m.strength;
        mat.textures[RH_MIX_MAT1] = m.mix_materials[0];
        mat.textures[RH_MIX_MAT2] = m.mix_materials[1];
    }
    mat.textures[RH_NORMALS_TEXTURE] = m.normal_map != 0xffffffff ? m.normal_map : default_normals_texture_;
    materials_.push_back(mat);
    Touch();
    return (uint32_t)(materials_.size() - 1);
}

// SceneRef::AddMesh (SceneRef.cpp:156-263), binary BVH
void DeviceLbvhBoxes(rh_builder *builder, const float *boxes, size_t n, uint32_t index_offset,
                     std::vector<uint32_t> &order, std::vector<rh_bvh_node> &bottom,
                     std::vector<RayHip::HlbvhTreelet> &treelets) {
    order.resize(n);
    bottom.resize(2 * n);
    std::vector<uint32_t> tl(3 * 4096);
    uint32_t nodes = 0, count = 0;
    if (rh_build_lbvh(builder, boxes, n, index_offset, order.data(), bottom.data(), bottom.size(), &nodes, tl.data(),
                      &count) != 0)
        throw std::runtime_error(std::string("fast BVH build: ") + rh_last_error());
    bottom.resize(nodes);
    treelets.resize(count);
    for (uint32_t t = 0; t < count; t++) treelets[t] = {tl[3 * t], tl[3 * t + 1], tl[3 * t + 2]};
}

RayHip::LbvhBottom DeviceLbvhBottom(rh_builder **builder, int device) {
    return [builder, device](const RayHip::prim_t *prims, size_t n, uint32_t index_offset,
                             std::vector<uint32_t> &order, std::vector<rh_bvh_node> &bottom,
                             std::vector<RayHip::HlbvhTreelet> &treelets) {
        if (!*builder && rh_builder_create(device, builder) != 0) {
            *builder = nullptr;
            throw std::runtime_error(std::string("fast BVH build: ") + rh_last_error());
        }
        std::vector<float> boxes(6 * n);
        for (size_t i = 0; i < n; i++) {
            memcpy(&boxes[6 * i], prims[i].bbox_min.v, 3 * sizeof(float));
            memcpy(&boxes[6 * i + 3], prims[i].bbox_max.v, 3 * sizeof(float));
        }
        DeviceLbvhBoxes(*builder, boxes.data(), n, index_offset, order, bottom, treelets);
    };
}

RayHip::SahBoxes DeviceSahBoxes(rh_builder **builder, int device) {
    return [builder, device](const float *boxes, size_t n, const RayHip::bvh_settings_t &s,
                             std::vector<rh_bvh_node> &nodes, std::vector<uint32_t> &indices) {
        if (!*builder && rh_builder_create(device, builder) != 0) {
            *builder = nullptr;
            throw std::runtime_error(std::string("SAH build: ") + rh_last_error());
        }
        nodes.resize(2 * n);
        indices.resize(n);
        uint32_t count = 0;
        int host_needed = 0;
        if (rh_build_sah(*builder, boxes, n, s.oversplit_threshold, s.node_traversal_cost, nodes.data(), nodes.size(),
                         &count, indices.data(), &host_needed) != 0)
            throw std::runtime_error(std::string("SAH build: ") + rh_last_error());
        if (host_needed) return false;
        nodes.resize(count);
        return true;
    };
}

Scene::~Scene() {
    if (builder_) rh_builder_destroy(builder_);
}

uint32_t Scene::AddMesh(const mesh_desc_t &_m) {
    if (!_m.vtx_indices_count || _m.vtx_indices_count % 3) return 0xffffffff;
    if ((int)_m.layout < (int)PxyzNxyzTuv || (int)_m.layout > (int)PxyzNxyzBxyzTuvTuv) {
        throw std::invalid_argument("AddMesh: unknown vertex layout " + std::to_string((int)_m.layout));
    }
    for (size_t i = 0; i < _m.vtx_indices_count; i++) {
        const int64_t v = (int64_t)_m.vtx_indices[i] + _m.base_vertex;
        if (v < 0 || v >= (int64_t)_m.vtx_attrs_count) {
            throw std::invalid_argument("AddMesh: index " + std::to_string(i) + " names vertex " + std::to_string(v) +
                                        " of " + std::to_string(_m.vtx_attrs_count));
        }
    }
    for (const shape_desc_t &sh : _m.shapes) {
        CheckMaterial(sh.mat_index, false, "AddMesh shape mat_index");
        CheckMaterial(sh.back_mat_index, true, "AddMesh shape back_mat_index");
        if (sh.vtx_start > _m.vtx_indices_count || sh.vtx_count > _m.vtx_indices_count - sh.vtx_start) {
            throw std::invalid_argument("AddMesh: shape range [" + std::to_string(sh.vtx_start) + ", +" +
                                        std::to_string(sh.vtx_count) + ") past the " +
                                        std::to_string(_m.vtx_indices_count) + " indices");
        }
    }
    meshes_.emplace_back();
    rh_mesh &m = meshes_.back();
    const uint32_t tris_start = (uint32_t)tris_.size();

    RayHip::bvh_settings_t s;
    s.node_traversal_cost = 0.025f;
    s.oversplit_threshold = 0.95f;
    s.allow_spatial_splits = _m.allow_spatial_splits;
    s.use_fast_bvh_build = _m.use_fast_bvh_build;

    m.node_index = (uint32_t)nodes_.size();
    const size_t nodes_before = nodes_.size(), tri_indices_before = tri_indices_.size();
    try {
        m.node_count = RayHip::PreprocessMesh(_m.vtx_attrs, _m.vtx_indices, _m.vtx_indices_count, _m.layout,
                                              _m.base_vertex, s, nodes_, tris_, tri_indices_,
                                              device_build_ ? &device_bottom_ : nullptr,
                                              device_build_ ? &device_sah_ : nullptr);
    } catch (...) {
        nodes_.resize(nodes_before);
        tris_.resize(tris_start);
        tri_indices_.resize(tri_indices_before);
        meshes_.pop_back();
        throw;
    }

    // triangle materials and the solid bit (no transparent material reachable)
    // (SceneRef.cpp:185-203 walks the mix tree with a 32-entry stack; here the
    // walk has no depth limit and visits each material once)
    std::vector<uint32_t> material_stack;
    std::vector<uint8_t> seen;
    for (const shape_desc_t &sh : _m.shapes) {
        bool is_solid = true;
        material_stack.assign(1, sh.mat_index);
        seen.assign(materials_.size(), 0);
        while (!material_stack.empty()) {
            const uint32_t mi = material_stack.back();
            material_stack.pop_back();
            if (seen[mi]) continue;
            seen[mi] = 1;
            const rh_material &mat = materials_[mi];
            if (mat.type == MixMaterial) {
                material_stack.push_back(mat.textures[RH_MIX_MAT1]);
                material_stack.push_back(mat.textures[RH_MIX_MAT2]);
            } else if (mat.type == TransparentMaterial) {
                is_solid = false;
                break;
            }
        }
        for (size_t i = sh.vtx_start; i < sh.vtx_start + sh.vtx_count; i += 3) {
            rh_tri_accel &tri = tris_[tris_start + i / 3];
            if (is_solid) {
                tri.ci = (int32_t)((uint32_t)tri.ci | RH_TRI_SOLID_BIT);
            } else {
                tri.ci = (int32_t)((uint32_t)tri.ci & ~RH_TRI_SOLID_BIT);
            }
            tri.mi = sh.mat_index;
            tri.back_mi = sh.back_mat_index;
        }
    }

    m.tris_index = tris_start;
    m.tris_count = (uint32_t)tris_.size() - tris_start;

    std::vector<uint32_t> new_vtx_indices;
    new_vtx_indices.reserve(_m.vtx_indices_count);
    for (size_t i = 0; i < _m.vtx_indices_count; i++) {
        new_vtx_indices.push_back(_m.vtx_indices[i] + _m.base_vertex + (uint32_t)vertices_.size());
    }

    const size_t stride = AttrStrides[_m.layout];
    const size_t new_vertices_start = vertices_.size();
    vertices_.resize(new_vertices_start + _m.vtx_attrs_count);
    for (size_t i = 0; i < _m.vtx_attrs_count; i++) {
        rh_vertex &v = vertices_[new_vertices_start + i];
        const float *a = _m.vtx_attrs + i * stride;
        memcpy(v.p, a, 3 * sizeof(float));
        memcpy(v.n, a + 3, 3 * sizeof(float));
        if (_m.layout == PxyzNxyzTuv) {
            memcpy(v.t[0], a + 6, 2 * sizeof(float));
            v.t[1][0] = v.t[1][1] = 0.0f;
            v.b[0] = v.b[1] = v.b[2] = 0.0f;
        } else if (_m.layout == PxyzNxyzTuvTuv) {
            memcpy(v.t[0], a + 6, 2 * sizeof(float));
            memcpy(v.t[1], a + 8, 2 * sizeof(float));
            v.b[0] = v.b[1] = v.b[2] = 0.0f;
        } else if (_m.layout == PxyzNxyzBxyzTuv) {
            memcpy(v.b, a + 6, 3 * sizeof(float));
            memcpy(v.t[0], a + 9, 2 * sizeof(float));
            v.t[1][0] = v.t[1][1] = 0.0f;
        } else {
            memcpy(v.b, a + 6, 3 * sizeof(float));
            memcpy(v.t[0], a + 9, 2 * sizeof(float));
            memcpy(v.t[1], a + 11, 2 * sizeof(float));
        }
    }
    if (_m.layout == PxyzNxyzTuv || _m.layout == PxyzNxyzTuvTuv) {
        RayHip::ComputeTangentBasis(0, new_vertices_start, vertices_, new_vtx_indices, new_vtx_indices.data(),
                                    new_vtx_indices.size());
    }
    vtx_indices_.insert(vtx_indices_.end(), new_vtx_indices.begin(), new_vtx_indices.end());
    wide_dirty_ = true;
    wide_meshes_dirty_ = true;
    Touch();
    return (uint32_t)(meshes_.size() - 1);
}

// Removes a mesh, its instances, triangles, tree nodes and index ranges, and
// re-bases every reference past the removed ranges.  (The reference's
// SceneRef::RemoveMesh, SceneRef.cpp:265-303, leaves tri_indices pointing at
// shifted triangles; this removal keeps the scene consistent.)
void Scene::RemoveMesh(uint32_t i) {
    if (i >= meshes_.size()) return;
    Finalize();
    wide_dirty_ = true;
    wide_meshes_dirty_ = true;
    const rh_mesh m = meshes_[i];

    // the mesh's triangle index range: leaves of its tree cover one contiguous range
    uint32_t ti_begin = 0xffffffff, ti_end = 0;
    for (uint32_t n = m.node_index; n < m.node_index + m.node_count; n++) {
        if (nodes_[n].prim_index & RH_LEAF_NODE_BIT) {
            const uint32_t b = nodes_[n].prim_index & RH_PRIM_INDEX_BITS;
            ti_begin = std::min(ti_begin, b);
            ti_end = std::max(ti_end, b + nodes_[n].prim_count);
        }
    }
    if (ti_begin == 0xffffffff) ti_begin = ti_end = 0;
    const uint32_t ti_count = ti_end - ti_begin;

    // instances of this mesh go; later meshes shift down by one
    std::vector<rh_mesh_instance> kept;
    for (const rh_mesh_instance &mi : mesh_instances_) {
        if (mi.mesh_index == i) continue;
        rh_mesh_instance c = mi;
        if (c.mesh_index > i) c.mesh_index--;
        kept.push_back(c);
    }
    mesh_instances_.swap(kept);

    tris_.erase(tris_.begin() + m.tris_index, tris_.begin() + m.tris_index + m.tris_count);
    tri_indices_.erase(tri_indices_.begin() + ti_begin, tri_indices_.begin() + ti_end);
    for (uint32_t &ti : tri_indices_) {
        if (ti >= m.tris_index + m.tris_count) ti -= m.tris_count;
    }
    // re-base leaves of the other mesh trees that point past the removed index range
    for (size_t k = 0; k < meshes_.size(); k++) {
        if (k == i) continue;
        for (uint32_t n = meshes_[k].node_index; n < meshes_[k].node_index + meshes_[k].node_count; n++) {
            rh_bvh_node &nd = nodes_[n];
            if ((nd.prim_index & RH_LEAF_NODE_BIT) && (nd.prim_index & RH_PRIM_INDEX_BITS) >= ti_end) nd.prim_index -= ti_count;
        }
        if (meshes_[k].tris_index > m.tris_index) meshes_[k].tris_index -= m.tris_count;
    }
    meshes_.erase(meshes_.begin() + i);
    RemoveNodes(m.node_index, m.node_count);
    macro_dirty_ = true;
    Touch();
}

uint32_t Scene::AddLight(const light_desc_t &_l) {
    rh_light l;
    memcpy(l.pos, _l.position, 3 * sizeof(float));
    l.radius = _l.radius;
    memcpy(l.col, _l.color, 3 * sizeof(float));
    if (_l.type == SpotLight) {
        l.dir[0] = -_l.direction[0];
        l.dir[1] = -_l.direction[1];
        l.dir[2] = -_l.direction[2];
        l.spot = std::cos(_l.angle * RH_PI / 180.0f);
    } else if (_l.type == PointLight) {
        l.dir[0] = l.dir[2] = 0.0f;
        l.dir[1] = 1.0f;
        l.spot = -1.0f;
    } else {
        const float dist = 9999999.0f;
        l.dir[0] = -_l.direction[0];
        l.dir[1] = -_l.direction[1];
        l.dir[2] = -_l.direction[2];
        l.pos[0] = l.dir[0] * dist;
        l.pos[1] = l.dir[1] * dist;
        l.pos[2] = l.dir[2] * dist;
        l.radius = dist * std::tan(_l.angle * RH_PI / 180.0f) + 1.0f;
        float k = 1.0f + dist / l.radius;
        k *= k;
        l.col[0] = _l.color[0] * k;
        l.col[1] = _l.color[1] * k;
        l.col[2] = _l.color[2] * k;
        l.spot = 0.0f;
    }
    l.brightness = std::max(l.col[0], std::max(l.col[1], l.col[2]));
    lights_.push_back(l);
    light_dirty_ = true;
    Touch();
    return (uint32_t)(lights_.size() - 1);
}

void Scene::RemoveLight(uint32_t i) {
    // the reference leaves this unimplemented (SceneRef.cpp:351-354)
    (void)i;
}

uint32_t Scene::AddMeshInstance(uint32_t mesh_index, const float *xform) {
    if (mesh_index >= meshes_.size()) return 0xffffffff;
    const uint32_t mi_index = (uint32_t)mesh_instances_.size();
    mesh_instances_.emplace_back();
    rh_mesh_instance &mi = mesh_instances_.back();
    memset(&mi, 0, sizeof(mi));
    mi.mesh_index = mesh_index;
    mi.tr_index = (uint32_t)transforms_.size();
    transforms_.emplace_back();
    SetMeshInstanceTransform(mi_index, xform);
    Touch();
    return mi_index;
}

// SceneRef::SetMeshInstanceTransform (SceneRef.cpp:370-414)
void Scene::SetMeshInstanceTransform(uint32_t mi_index, const float *xform) {
    rh_mesh_instance &mi = mesh_instances_[mi_index];
    rh_transform &tr = transforms_[mi.tr_index];
    memcpy(tr.xform, xform, 16 * sizeof(float));
    RayHip::InverseMatrix(tr.xform, tr.inv_xform);
    const rh_mesh &m = meshes_[mi.mesh_index];
    if (wide_) {
        // the bounds of the mesh's wide root (SceneRef.cpp:383-406)
        float bmin[3], bmax[3];
        RayHip::WideRootBounds(nodes_.data(), m.node_index, bmin, bmax);
        RayHip::TransformBoundingBox(bmin, bmax, xform, mi.bbox_min, mi.bbox_max);
    } else {
        const rh_bvh_node &n = nodes_[m.node_index];
        RayHip::TransformBoundingBox(n.bbox_min, n.bbox_max, xform, mi.bbox_min, mi.bbox_max);
    }
    macro_dirty_ = true;
    version_++;
}

void Scene::RemoveMeshInstance(uint32_t i) {
    if (i >= mesh_instances_.size()) return;
    mesh_instances_.erase(mesh_instances_.begin() + i);
    macro_dirty_ = true;
    Touch();
}

// SceneRef::RemoveNodes (SceneRef.cpp:437-490), binary BVH
void Scene::RemoveNodes(uint32_t node_index, uint32_t node_count) {
    if (!node_count) return;
    nodes_.erase(nodes_.begin() + node_index, nodes_.begin() + node_index + node_count);
    if (node_index != nodes_.size()) {
        for (rh_mesh &m : meshes_) {
            if (m.node_index > node_index) m.node_index -= node_count;
        }
        for (uint32_t i = node_index; i < nodes_.size(); i++) {
            rh_bvh_node &n = nodes_[i];
            if ((n.prim_index & RH_LEAF_NODE_BIT) == 0) {
                if (n.prim_index > node_index) n.prim_index -= node_count;
                if ((n.prim_count & RH_RIGHT_CHILD_BITS) > node_index) n.prim_count -= node_count;
            }
        }
        if (macro_nodes_root_ != 0xffffffff && macro_nodes_root_ > node_index) macro_nodes_root_ -= node_count;
        if (light_nodes_root_ != 0xffffffff && light_nodes_root_ > node_index) light_nodes_root_ -= node_count;
    }
}

// SceneRef::RebuildMacroBVH (SceneRef.cpp:492-516)
void Scene::RebuildMacroBVH() {
    if (macro_nodes_root_ != 0xffffffff) RemoveNodes(macro_nodes_root_, macro_nodes_count_);
    mi_indices_.clear();
    macro_nodes_root_ = 0xffffffff;
    macro_nodes_count_ = 0;
    if (mesh_instances_.empty()) return;

    std::vector<RayHip::prim_t> primitives;
    primitives.reserve(mesh_instances_.size());
    for (const rh_mesh_instance &mi : mesh_instances_) {
        primitives.push_back({0, 0, 0, {{mi.bbox_min[0], mi.bbox_min[1], mi.bbox_min[2]}},
                              {{mi.bbox_max[0], mi.bbox_max[1], mi.bbox_max[2]}}});
    }
    macro_nodes_root_ = (uint32_t)nodes_.size();
    // a renderer's scene builds the instance tree on the device (the same tree)
    macro_nodes_count_ = RayHip::PreprocessPrims_SAH_Boxes(primitives.data(), primitives.size(), {}, nodes_,
                                                           mi_indices_, device_build_ ? &device_sah_ : nullptr);
}

// SceneRef::RebuildLightBVH (SceneRef.cpp:518-587): each light's influence
// region, transformed into world space, is one primitive of the light tree
void Scene::RebuildLightBVH() {
    if (light_nodes_root_ != 0xffffffff) RemoveNodes(light_nodes_root_, light_nodes_count_);
    li_indices_.clear();
    light_nodes_root_ = 0xffffffff;
    light_nodes_count_ = 0;
    if (lights_.empty()) return;

    using RayHip::vec3;
    std::vector<RayHip::prim_t> primitives;
    primitives.reserve(lights_.size());
    for (const rh_light &l : lights_) {
        const float influence = l.radius * (std::sqrt(l.brightness / RH_LIGHT_ATTEN_CUTOFF) - 1.0f);
        vec3 bbox_min = {{0.0f, 0.0f, 0.0f}}, bbox_max = {{0.0f, 0.0f, 0.0f}};
        const vec3 p1 = {{-l.dir[0] * influence, -l.dir[1] * influence, -l.dir[2] * influence}};
        bbox_min = RayHip::vmin(bbox_min, p1);
        bbox_max = RayHip::vmax(bbox_max, p1);
        const vec3 p2 = {{-l.dir[0] * l.spot * influence, -l.dir[1] * l.spot * influence, -l.dir[2] * l.spot * influence}};
        const float d = std::sqrt(1.0f - l.spot * l.spot) * influence;
        bbox_min = RayHip::vmin(bbox_min, vec3{{p2[0] - d, p2[1] - 0.0f, p2[2] - d}});
        bbox_max = RayHip::vmax(bbox_max, vec3{{p2[0] + d, p2[1] + 0.0f, p2[2] + d}});
        if (l.spot < 0.0f) {
            bbox_min = RayHip::vmin(bbox_min, vec3{{p1[0] - influence, p1[1] - 0.0f, p1[2] - influence}});
            bbox_max = RayHip::vmax(bbox_max, vec3{{p1[0] + influence, p1[1] + 0.0f, p1[2] + influence}});
        }
        vec3 up = {{1.0f, 0.0f, 0.0f}};
        if (std::abs(l.dir[1]) < std::abs(l.dir[2]) && std::abs(l.dir[1]) < std::abs(l.dir[0])) {
            up = {{0.0f, 1.0f, 0.0f}};
        } else if (std::abs(l.dir[2]) < std::abs(l.dir[0]) && std::abs(l.dir[2]) < std::abs(l.dir[1])) {
            up = {{0.0f, 0.0f, 1.0f}};
        }
        const vec3 side = {{-l.dir[1] * up[2] + l.dir[2] * up[1], -l.dir[2] * up[0] + l.dir[0] * up[2],
                            -l.dir[0] * up[1] + l.dir[1] * up[0]}};
        const float xform[16] = {side[0], l.dir[0], up[0], 0.0f, side[1], l.dir[1], up[1], 0.0f,
                                 side[2], l.dir[2], up[2], 0.0f, l.pos[0], l.pos[1], l.pos[2], 1.0f};
        RayHip::prim_t prim;
        prim.i0 = prim.i1 = prim.i2 = 0;
        RayHip::TransformBoundingBox(bbox_min.v, bbox_max.v, xform, prim.bbox_min.v, prim.bbox_max.v);
        primitives.push_back(prim);
    }
    light_nodes_root_ = (uint32_t)nodes_.size();
    light_nodes_count_ =
        RayHip::PreprocessPrims_SAH(primitives.data(), primitives.size(), nullptr, 0, {}, nodes_, li_indices_);
}

void Scene::Finalize() {
    if (light_dirty_) {
        RebuildLightBVH();
        light_dirty_ = false;
        wide_dirty_ = true;
    }
    if (macro_dirty_) {
        RebuildMacroBVH();
        macro_dirty_ = false;
        wide_dirty_ = true;
    }
    if (wide_ && wide_dirty_) RebuildWide();
}

// FlattenBVH of every tree: the mesh trees in mesh order, then the macro and
// light trees.  Node indices differ from the reference's array (which the
// reference edits in place as meshes, instances and lights come and go), the
// trees themselves do not.
void Scene::RebuildWide() {
    if (wide_meshes_dirty_ || wide_meshes_.size() != meshes_.size()) {
        mnodes_.clear();
        wide_meshes_ = meshes_;
        for (rh_mesh &m : wide_meshes_) {
            const uint32_t before = (uint32_t)mnodes_.size();
            m.node_index = RayHip::FlattenBVH(nodes_.data(), m.node_index, mnodes_);
            m.node_count = (uint32_t)mnodes_.size() - before;
        }
        wide_mesh_end_ = (uint32_t)mnodes_.size();
        wide_meshes_dirty_ = false;
    } else {
        // only the instance or light tree changed: the meshes' wide nodes,
        // which do not depend on where their binary nodes sit, stay
        mnodes_.resize(wide_mesh_end_);
    }
    wide_macro_root_ = macro_nodes_root_ != 0xffffffff ? RayHip::FlattenBVH(nodes_.data(), macro_nodes_root_, mnodes_)
                                                       : 0xffffffff;
    wide_light_root_ = light_nodes_root_ != 0xffffffff ? RayHip::FlattenBVH(nodes_.data(), light_nodes_root_, mnodes_)
                                                       : 0xffffffff;
    wide_dirty_ = false;
}

uint32_t Scene::node_count() {
    Finalize();
    return (uint32_t)(wide_ ? mnodes_.size() : nodes_.size());
}

// camera record of the current camera; the ray-sort hash grid spans the
// macro-tree root box in 255 cells per axis (RendererRef.cpp:66-77)
void Scene::PassCamera(rh_pass_desc &pass) {
    memcpy(&pass.cam, &camera(current_cam()), sizeof(rh_camera));
    if (wide_ && wide_macro_root_ != 0xffffffff) {
        // the wide root's box as the reference's wide path reads it: slot 0 of
        // a leaf, else the box of the last used slot (RendererRef.cpp:52-66)
        const rh_mbvh_node &root = mnodes_[wide_macro_root_];
        float lo[3] = {RH_MAX_DIST, RH_MAX_DIST, RH_MAX_DIST}, hi[3] = {-RH_MAX_DIST, -RH_MAX_DIST, -RH_MAX_DIST};
        for (int j = 0; j < 8; j++) {
            if (!(root.child[0] & RH_LEAF_NODE_BIT) && root.child[j] == RH_MBVH_EMPTY) continue;
            for (int i = 0; i < 3; i++) {
                lo[i] = root.bbox_min[i][j];
                hi[i] = root.bbox_max[i][j];
            }
            if (root.child[0] & RH_LEAF_NODE_BIT) break;
        }
        for (int i = 0; i < 3; i++) {
            pass.root_min[i] = lo[i];
            pass.cell_size[i] = (hi[i] - lo[i]) / 255;
        }
    } else if (macro_nodes_root_ != 0xffffffff) {
        const rh_bvh_node &root = nodes_[macro_nodes_root_];
        for (int i = 0; i < 3; i++) {
            pass.root_min[i] = root.bbox_min[i];
            pass.cell_size[i] = (root.bbox_max[i] - root.bbox_min[i]) / 255;
        }
    }
}

rh_scene_view Scene::View() {
    Finalize();
    rh_scene_view v;
    memset(&v, 0, sizeof(v));
#define RH_VIEW(field, vec)                                                                                            \
    v.field = vec.empty() ? nullptr : vec.data();                                                                     \
    v.field##_count = (uint32_t)vec.size();
    v.env = &env_;
    RH_VIEW(mesh_instances, mesh_instances_)
    RH_VIEW(mi_indices, mi_indices_)
    RH_VIEW(meshes, meshes_)
    RH_VIEW(transforms, transforms_)
    RH_VIEW(vtx_indices, vtx_indices_)
    RH_VIEW(vertices, vertices_)
    if (wide_) {
        RH_VIEW(meshes, wide_meshes_)
        RH_VIEW(mnodes, mnodes_)
    } else {
        RH_VIEW(nodes, nodes_)
    }
    RH_VIEW(tris, tris_)
    RH_VIEW(tri_indices, tri_indices_)
    RH_VIEW(materials, materials_)
    RH_VIEW(textures, textures_)
    RH_VIEW(texels, texels_)
    RH_VIEW(lights, lights_)
    RH_VIEW(li_indices, li_indices_)
#undef RH_VIEW
    v.tex_mips = tex_mips_.empty() ? nullptr : tex_mips_.data();
    v.macro_nodes_root = wide_ ? wide_macro_root_ : macro_nodes_root_;
    v.light_nodes_root = wide_ ? wide_light_root_ : light_nodes_root_;
    return v;
}

} // namespace Hip
} // namespace Ray
