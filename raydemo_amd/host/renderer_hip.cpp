// Ray::Hip::Renderer: drives one RenderScene pass on the device through the
// C ABI of include/rh_hip.h.  Pass set-up follows the Ref renderer
// (src/Ray/internal/RendererRef.cpp:19-325): region iteration and Halton
// refresh, depth clamp, hash-grid box from the macro-tree root; the stages
// themselves (generate, trace, sort, shade, accumulate) run as HIP kernels.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "Ray/RendererFactory.h"
#include "hip_backend.h"

namespace Ray {
namespace Hip {

Renderer::Renderer(const settings_t &s, int device_index) {
    const int rc = rh_create(device_index, &ctx_);
    if (rc != 0) {
        throw std::runtime_error(std::string("HIP backend: ") + rh_last_error());
    }
    device_ = device_index;
    permutations_ = RayHip::ComputeRadicalInversePermutations();
    use_wide_bvh_ = s.use_wide_bvh;
    Resize(s.w, s.h);
}

Renderer::~Renderer() {
    if (ctx_) rh_destroy(ctx_);
}

void Renderer::Resize(int w, int h) {
    if (w == w_ && h == h_) return;
    if (rh_resize(ctx_, w, h) != 0) throw std::runtime_error(std::string("HIP backend: ") + rh_last_error());
    w_ = w;
    h_ = h;
    host_pixels_.assign((size_t)w * h, pixel_color_t{0, 0, 0, 0});
    host_dirty_ = true;
}

void Renderer::Clear(const pixel_color_t &c) {
    const float rgba[4] = {c.r, c.g, c.b, c.a};
    if (rh_clear(ctx_, rgba) != 0) last_error = rh_last_error();
    host_dirty_ = true;
}

const pixel_color_t *Renderer::get_pixels_ref() const {
    if (host_dirty_) {
        if (rh_read_pixels(ctx_, reinterpret_cast<float *>(host_pixels_.data()), host_pixels_.size() * 4) == 0) {
            host_dirty_ = false;
        }
    }
    return host_pixels_.data();
}

const shl1_data_t *Renderer::get_sh_data_ref() const {
    host_sh_.resize((size_t)w_ * h_);
    const int rc = rh_read_sh(ctx_, reinterpret_cast<float *>(host_sh_.data()), host_sh_.size() * 12);
    return rc == 0 ? host_sh_.data() : nullptr;
}

std::shared_ptr<SceneBase> Renderer::CreateScene() { return std::make_shared<Scene>(use_wide_bvh_, true, device_); }

// Uploads the scene if it changed, advances the region's iteration and
// refreshes/uploads its Halton table exactly as the Ref renderer does
// (RendererRef.cpp:19-96), and describes the pass.  false: nothing to render.
bool Renderer::PreparePass(Scene &s, RegionContext &region, rh_pass_desc &pass) {
    if (s.current_cam() == 0xffffffff) return false;

    s.Finalize();
    if (uploaded_scene_ != &s || uploaded_version_ != s.version()) {
        const rh_scene_view v = s.View();
        // only instance transforms moved: the device keeps the rest
        const bool moved_only = uploaded_scene_ == &s && uploaded_static_version_ == s.static_version();
        if (!(moved_only && rh_update_instances(ctx_, &v) == 0) && rh_upload_scene(ctx_, &v) != 0) {
            last_error = rh_last_error();
            uploaded_scene_ = nullptr;
            return false;
        }
        uploaded_scene_ = &s;
        uploaded_version_ = s.version();
        uploaded_static_version_ = s.static_version();
    }

    memset(&pass, 0, sizeof(pass));
    s.PassCamera(pass);

    rect_t rect = region.rect();
    if (rect.w == 0 || rect.h == 0) rect = {0, 0, w_, h_};
    if (rect.x < 0 || rect.y < 0 || rect.x + rect.w > w_ || rect.y + rect.h > h_) {
        last_error = "region outside the framebuffer";
        return false;
    }

    region.iteration++;
    if (!region.halton_seq || region.iteration % RH_HALTON_SEQ_LEN == 0) {
        if (!region.halton_seq) region.halton_seq.reset(new float[RH_HALTON_COUNT * RH_HALTON_SEQ_LEN]);
        RayHip::UpdateHaltonSequence(permutations_, region.iteration, region.halton_seq.get());
        uploaded_halton_ = nullptr;
    }
    const int epoch = region.iteration / RH_HALTON_SEQ_LEN;
    if (uploaded_halton_ != region.halton_seq.get() || uploaded_halton_epoch_ != epoch) {
        if (rh_set_halton(ctx_, region.halton_seq.get(), RH_HALTON_COUNT * RH_HALTON_SEQ_LEN) != 0) {
            last_error = rh_last_error();
            return false;
        }
        uploaded_halton_ = region.halton_seq.get();
        uploaded_halton_epoch_ = epoch;
    }

    pass.rect[0] = rect.x;
    pass.rect[1] = rect.y;
    pass.rect[2] = rect.w;
    pass.rect[3] = rect.h;
    pass.iteration = region.iteration;
    pass.sort_rays = sort_rays_ ? 1 : 0;
    // the device clamps to its "max_bounces" option (the reference's 10 by default)
    pass.cam.pass_settings.max_total_depth =
        std::min<uint8_t>(pass.cam.pass_settings.max_total_depth, RH_MAX_BOUNCES_EXT);
    return true;
}

void Renderer::RenderScene(const std::shared_ptr<SceneBase> &_s, RegionContext &region) {
    const auto s = std::dynamic_pointer_cast<Scene>(_s);
    if (!s) return; // scene of another backend: no-op, as RendererRef.cpp:20-21
    rh_pass_desc pass;
    if (!PreparePass(*s, region, pass)) return;
    if (rh_render_pass(ctx_, &pass) != 0) {
        last_error = rh_last_error();
        return;
    }
    host_dirty_ = true;
}

bool Renderer::AdjointScene(const std::shared_ptr<SceneBase> &_s, RegionContext &region, const float *d_adj_rgba,
                            float *d_grad, float *d_texel_grad, float scale) {
    const auto s = std::dynamic_pointer_cast<Scene>(_s);
    if (!s) return false;
    rh_pass_desc pass;
    if (!PreparePass(*s, region, pass)) return false;
    if (rh_adjoint_pass(ctx_, &pass, d_adj_rgba, d_grad, d_texel_grad, scale) != 0) {
        last_error = rh_last_error();
        return false;
    }
    return true;
}

bool Renderer::SetMaterialColors(const std::shared_ptr<SceneBase> &_s, uint32_t first, uint32_t n, const float *rgb) {
    const auto s = std::dynamic_pointer_cast<Scene>(_s);
    if (!s) return false;
    if ((size_t)first + n > s->materials_.size()) {
        last_error = "material range outside the scene";
        return false;
    }
    for (uint32_t k = 0; k < n; k++)
        for (int ch = 0; ch < 3; ch++) s->materials_[first + k].main_color[ch] = rgb[3 * k + ch];
    // the device copy is patched in place when this scene's materials are the
    // uploaded ones (moved instances aside); otherwise the next pass uploads
    // the updated host arrays
    if (uploaded_scene_ == s.get() && uploaded_static_version_ == s->static_version()) {
        if (rh_set_material_colors(ctx_, first, n, rgb) != 0) {
            last_error = rh_last_error();
            return false;
        }
    }
    return true;
}

bool Renderer::SetTextureData(const std::shared_ptr<SceneBase> &_s, uint32_t tex, const uint8_t *rgba) {
    const auto s = std::dynamic_pointer_cast<Scene>(_s);
    if (!s) return false;
    uint32_t first = 0, count = 0;
    if (!s->SetTextureData(tex, reinterpret_cast<const uint32_t *>(rgba), first, count)) {
        last_error = "texture index out of range";
        return false;
    }
    if (uploaded_scene_ == s.get() && uploaded_static_version_ == s->static_version()) {
        if (rh_set_texels(ctx_, first, count, &s->texels_[first]) != 0) {
            last_error = rh_last_error();
            return false;
        }
    }
    return true;
}

bool Renderer::ReadAccumulated(float *dst, size_t count) {
    if (rh_read_accum(ctx_, dst, count) != 0) {
        last_error = rh_last_error();
        return false;
    }
    return true;
}

void Renderer::GetStats(stats_t &st) {
    rh_stats rs;
    if (rh_get_stats(ctx_, &rs) == 0) {
        static_assert(sizeof(rs) == sizeof(stats_t), "stats layout");
        memcpy(&st, &rs, sizeof(st));
    } else {
        st = stats_;
    }
}

void Renderer::ResetStats() { rh_reset_stats(ctx_); }

} // namespace Hip

std::shared_ptr<RendererBase> CreateRenderer(const settings_t &s, uint32_t flags, std::ostream &log_stream) {
    if (flags & RendererHIP) {
        log_stream << "Ray: Creating HIP renderer " << s.w << "x" << s.h << std::endl;
        return std::make_shared<Hip::Renderer>(s, settings_device(s, 0));
    }
    throw std::runtime_error("this library provides only the HIP backend (RendererHIP)");
}

} // namespace Ray
