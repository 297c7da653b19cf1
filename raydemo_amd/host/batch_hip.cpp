// Ray::Hip::BatchRenderer: many scenes per pass through the batch ABI of
// include/rh_hip.h.  Pass numbering and the Halton refresh follow a fresh
// RegionContext per scene (RendererRef.cpp:84-96): pass k is iteration k, and
// the table is rebuilt at k == 1 and whenever k % 256 == 0.
#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "hip_backend.h"

namespace Ray {
namespace Hip {

BatchRenderer::BatchRenderer(int device_index) {
    if (rh_batch_create(device_index, &b_) != 0) throw std::runtime_error(std::string("HIP batch: ") + rh_last_error());
    permutations_ = RayHip::ComputeRadicalInversePermutations();
    halton_.assign((size_t)RH_HALTON_COUNT * RH_HALTON_SEQ_LEN, 0.0f);
}

BatchRenderer::~BatchRenderer() {
    if (b_) rh_batch_destroy(b_);
}

int BatchRenderer::AddScene(const std::shared_ptr<SceneBase> &_s, int w, int h, int spp) {
    const auto s = std::dynamic_pointer_cast<Scene>(_s);
    if (!s) {
        last_error = "scene of another backend";
        return -1;
    }
    if (s->current_cam() == 0xffffffff) {
        last_error = "scene has no camera";
        return -1;
    }
    if (passes_done() > 0) {
        last_error = "scenes are added before the first pass";
        return -1;
    }
    const rh_scene_view v = s->View();
    rh_pass_desc pass;
    memset(&pass, 0, sizeof(pass));
    s->PassCamera(pass);
    if (pass.cam.type == Geo || (pass.cam.pass_settings.flags & OutputSH)) {
        last_error = "the batch renders perspective cameras without SH output";
        return -1;
    }
    pass.rect[2] = w;
    pass.rect[3] = h;
    pass.iteration = 1;
    const int idx = rh_batch_add_scene(b_, &v, &pass, w, h, spp);
    if (idx < 0) {
        last_error = rh_last_error();
        return -1;
    }
    sizes_.emplace_back(w, h);
    pixels_.emplace_back();
    return idx;
}

bool BatchRenderer::RenderPass() {
    // the device's count: a knob that lays the batch out again restarts it
    const int done = passes_done();
    if (done >= passes()) {
        last_error = "every scene has rendered its passes";
        return false;
    }
    const int it = done + 1;
    if (it == 1 || it % RH_HALTON_SEQ_LEN == 0) {
        RayHip::UpdateHaltonSequence(permutations_, it, halton_.data());
        if (rh_batch_set_halton(b_, halton_.data(), halton_.size()) != 0) {
            last_error = rh_last_error();
            return false;
        }
    }
    if (rh_batch_render_pass(b_) != 0) {
        last_error = rh_last_error();
        return false;
    }
    return true;
}

bool BatchRenderer::Restart() {
    if (rh_batch_restart(b_) != 0) {
        last_error = rh_last_error();
        return false;
    }
    return true;
}

const pixel_color_t *BatchRenderer::get_pixels_ref(int scene) {
    if (scene < 0 || scene >= (int)sizes_.size()) return nullptr;
    auto &px = pixels_[scene];
    px.resize((size_t)sizes_[scene].first * sizes_[scene].second);
    if (rh_batch_read_pixels(b_, scene, reinterpret_cast<float *>(px.data()), px.size() * 4) != 0) {
        last_error = rh_last_error();
        return nullptr;
    }
    return px.data();
}

} // namespace Hip
} // namespace Ray
