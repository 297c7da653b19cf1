// Ray::Hip::Renderer and Ray::Hip::Scene — the HIP backend behind the Ray
// renderer/scene API.  They take the place of Ray::Ocl::Renderer/Scene
// (src/Ray/internal/RendererOCL.h:20-190, SceneOCL.h) and follow the CPU Ref
// backend's semantics (src/Ray/internal/RendererRef.cpp, SceneRef.cpp).
// Host-side scene preprocessing happens here in C++; the device is reached
// only through the C ABI of include/rh_hip.h.
#pragma once

#include <cstdint>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "Ray/RendererBase.h"
#include "core.h"
#include "rh_hip.h"

namespace Ray {
// settings_t::device_index (the HIP device ordinal) where the build declares
// it: the reference declares platform_index/device_index only when its OpenCL
// backend is compiled in (RendererBase.h:24-29); otherwise device 0.
template <class S> inline auto settings_device(const S &s, int) -> decltype(s.device_index, int()) {
    return s.device_index;
}
template <class S> inline int settings_device(const S &, long) { return 0; }
template <class S> inline auto set_settings_device(S &s, int d, int) -> decltype(s.device_index = d, void()) {
    s.device_index = d;
}
template <class S> inline void set_settings_device(S &, int, long) {}

namespace Hip {

class Renderer;

class Scene : public SceneBase {
    friend class Renderer;

    std::vector<rh_bvh_node> nodes_;
    std::vector<rh_tri_accel> tris_;
    std::vector<uint32_t> tri_indices_;
    std::vector<rh_transform> transforms_;
    std::vector<rh_mesh> meshes_;
    std::vector<rh_mesh_instance> mesh_instances_;
    std::vector<uint32_t> mi_indices_;
    std::vector<rh_vertex> vertices_;
    std::vector<uint32_t> vtx_indices_;
    std::vector<rh_material> materials_;
    std::vector<rh_texture> textures_;
    std::vector<rh_tex_mip> tex_mips_;
    std::vector<uint32_t> texels_;
    std::vector<rh_light> lights_;
    std::vector<uint32_t> li_indices_;
    rh_environment env_ = {};

    uint32_t macro_nodes_root_ = 0xffffffff, macro_nodes_count_ = 0;
    uint32_t light_nodes_root_ = 0xffffffff, light_nodes_count_ = 0;
    uint32_t default_normals_texture_ = 0xffffffff, default_env_texture_ = 0xffffffff;

    // texture atlas placement (TextureAtlasTiled, TextureAtlasRef.cpp:178-266)
    std::vector<RayHip::TextureSplitter> splitters_;

    // the macro tree is rebuilt lazily (once before use) instead of after every
    // instance change; the light tree likewise
    bool macro_dirty_ = false, light_dirty_ = false;
    // version_ changes with every edit; static_version_ with every edit but
    // an instance transform (then the device keeps triangles, textures, ...)
    uint64_t version_ = 1, static_version_ = 1;
    void Touch() {
        version_++;
        static_version_++;
    }

    // 8-wide trees (settings use_wide_bvh, the reference's default): the
    // binary trees above stay the build's working form, and the wide arrays
    // are flattened from them before use, mesh trees first, then the macro
    // and light trees (SceneRef.cpp:172-181, 506-515, 577-586)
    bool wide_ = false;
    bool wide_dirty_ = true;
    std::vector<rh_mbvh_node> mnodes_;
    std::vector<rh_mesh> wide_meshes_;
    uint32_t wide_macro_root_ = 0xffffffff, wide_light_root_ = 0xffffffff;
    bool wide_meshes_dirty_ = true; // a mesh came or went: every tree is flattened again
    uint32_t wide_mesh_end_ = 0;    // the mesh trees' part of mnodes_
    void RebuildWide();

    // fast (HLBVH) builds run on the device when the scene comes from a
    // renderer; a standalone scene builds on the host
    bool device_build_ = false;
    int build_device_ = -1;
    rh_builder *builder_ = nullptr;
    RayHip::LbvhBottom device_bottom_;
    RayHip::SahBoxes device_sah_; // the instance tree and the fast build's top tree

    void RemoveNodes(uint32_t node_index, uint32_t node_count);
    void WriteBlock(const rh_tex_mip &m, const std::vector<uint32_t> &data, const int res[2]);
    void RebuildMacroBVH();
    void RebuildLightBVH();
    int AllocateInAtlas(const int res[2], int pos[2]);
    // std::invalid_argument unless tex (mat) names an existing texture
    // (material), or is 0xffffffff where optional
    void CheckTexture(uint32_t tex, bool optional, const char *what) const;
    void CheckMaterial(uint32_t mat, bool optional, const char *what) const;

public:
    explicit Scene(bool use_wide_bvh = false, bool device_build = false, int device = -1);
    ~Scene() override;
    Scene(const Scene &) = delete;
    Scene &operator=(const Scene &) = delete;
    bool wide() const { return wide_; }
    bool device_build() const { return device_build_; }

    void GetEnvironment(environment_desc_t &env) override;
    void SetEnvironment(const environment_desc_t &env) override;

    uint32_t AddTexture(const tex_desc_t &t) override;
    void RemoveTexture(uint32_t) override {}

    uint32_t AddMaterial(const mat_desc_t &m) override;
    void RemoveMaterial(uint32_t) override {}

    uint32_t AddMesh(const mesh_desc_t &m) override;
    void RemoveMesh(uint32_t i) override;

    uint32_t AddLight(const light_desc_t &l) override;
    void RemoveLight(uint32_t i) override;

    uint32_t AddMeshInstance(uint32_t m_index, const float *xform) override;
    void SetMeshInstanceTransform(uint32_t mi_index, const float *xform) override;
    void RemoveMeshInstance(uint32_t mi_index) override;

    uint32_t triangle_count() override { return (uint32_t)tris_.size(); }
    uint32_t node_count() override;

    // Brings the lazily rebuilt trees up to date and describes the arrays.
    void Finalize();
    rh_scene_view View();
    // The current camera and the sort hash grid (macro-tree root box) of a pass.
    void PassCamera(rh_pass_desc &pass);
    // New level-0 texels (RGBA8, the texture's size) of texture tex: its
    // stored blocks (borders, and mips when it has them) are rebuilt in place;
    // [first, first + count) is the range of the texel array they occupy.
    bool SetTextureData(uint32_t tex, const uint32_t *rgba, uint32_t &first, uint32_t &count);
    uint64_t version() const { return version_; }
    uint64_t static_version() const { return static_version_; }
    const camera_t &camera(uint32_t i) const { return cams_[i].cam; }
};

// The fast build's bottom half (Morton sort, treelets) on GPU `device`, with
// the builder created at first use in *builder (kept by the caller).
RayHip::LbvhBottom DeviceLbvhBottom(rh_builder **builder, int device);
// The SAH build of boxes on GPU `device` (rh_build_sah), declining where the
// host's order of equal keys decides.
RayHip::SahBoxes DeviceSahBoxes(rh_builder **builder, int device);
// the same for boxes already packed as 6 floats each
void DeviceLbvhBoxes(rh_builder *builder, const float *boxes, size_t n, uint32_t index_offset,
                     std::vector<uint32_t> &order, std::vector<rh_bvh_node> &bottom,
                     std::vector<RayHip::HlbvhTreelet> &treelets);

class Renderer : public RendererBase {
    rh_context *ctx_ = nullptr;
    int device_ = -1;
    int w_ = 0, h_ = 0;
    bool sort_rays_ = true;
    bool use_wide_bvh_ = false;

    std::vector<uint16_t> permutations_;
    mutable std::vector<pixel_color_t> host_pixels_;
    mutable std::vector<shl1_data_t> host_sh_;
    mutable bool host_dirty_ = true;
    const Scene *uploaded_scene_ = nullptr;
    uint64_t uploaded_version_ = 0, uploaded_static_version_ = 0;
    const float *uploaded_halton_ = nullptr;
    int uploaded_halton_epoch_ = -1;
    stats_t stats_ = {};

    bool PreparePass(Scene &s, RegionContext &region, rh_pass_desc &pass);

public:
    Renderer(const settings_t &s, int device_index);
    ~Renderer() override;

    eRendererType type() const override { return RendererHIP; }
    std::pair<int, int> size() const override { return std::make_pair(w_, h_); }
    const pixel_color_t *get_pixels_ref() const override;
    // null until a pass with OutputSH has run (since the last Resize)
    const shl1_data_t *get_sh_data_ref() const override;
    void Resize(int w, int h) override;
    void Clear(const pixel_color_t &c) override;
    std::shared_ptr<SceneBase> CreateScene() override;
    void RenderScene(const std::shared_ptr<SceneBase> &s, RegionContext &region) override;
    void GetStats(stats_t &st) override;
    void ResetStats() override;

    // HIP-specific
    rh_context *context() { return ctx_; }
    // Differentiation.  AdjointScene replays the pass a RenderScene call with
    // the same region state would render and accumulates d loss / d
    // main_color (see rh_adjoint_pass); SetMaterialColors patches main_color
    // (host rgb triples) in the scene and in its device copy;
    // ReadAccumulated copies the linear accumulated image.
    bool AdjointScene(const std::shared_ptr<SceneBase> &s, RegionContext &region, const float *d_adj_rgba,
                      float *d_grad, float *d_texel_grad, float scale);
    bool SetMaterialColors(const std::shared_ptr<SceneBase> &s, uint32_t first, uint32_t n, const float *rgb);
    bool SetTextureData(const std::shared_ptr<SceneBase> &s, uint32_t tex, const uint8_t *rgba);
    bool ReadAccumulated(float *dst, size_t count);
    void set_sort_rays(bool on) { sort_rays_ = on; }
    std::string last_error;
};

// Many scenes rendered together: each pass renders one progressive pass of
// every scene that still has samples to take, in one set of persistent
// kernels (include/rh_hip.h, rh_batch_*).  Each scene's image equals what
// Renderer::RenderScene renders for it with a full-frame region, pass for
// pass.  Scenes are copied when added; later changes to them are not seen.
class BatchRenderer {
    rh_batch *b_ = nullptr;
    std::vector<uint16_t> permutations_;
    std::vector<float> halton_;
    std::vector<std::pair<int, int>> sizes_;
    std::vector<std::vector<pixel_color_t>> pixels_;

public:
    explicit BatchRenderer(int device_index);
    ~BatchRenderer();
    BatchRenderer(const BatchRenderer &) = delete;
    BatchRenderer &operator=(const BatchRenderer &) = delete;

    // Adds a scene of this backend (its current camera) at w x h for spp
    // passes; returns its index, or -1 (see last_error).
    int AddScene(const std::shared_ptr<SceneBase> &s, int w, int h, int spp);
    int passes() const { return rh_batch_passes(b_); }
    int passes_done() const { return rh_batch_passes_done(b_); }
    int scenes() const { return (int)sizes_.size(); }
    // One pass over the batch; false on error or when every scene is done.
    bool RenderPass();
    // Back to pass 0 with cleared images; the scenes stay on the device.
    bool Restart();
    const pixel_color_t *get_pixels_ref(int scene);
    std::pair<int, int> size(int scene) const { return sizes_[scene]; }
    rh_batch *handle() { return b_; }
    std::string last_error;
};

} // namespace Hip
} // namespace Ray
