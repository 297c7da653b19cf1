// Renderer interface of the Ray renderer API, declared for the HIP backend.
// Mirrors the reference's src/Ray/RendererBase.h:14-103; the HIP backend adds
// the RendererHIP type bit next to RendererOCL.
#pragma once

#include <memory>
#include <utility>

#include "SceneBase.h"
#include "Types.h"

namespace Ray {
enum eRendererType {
    RendererRef = (1 << 0),
    RendererSSE2 = (1 << 1),
    RendererAVX = (1 << 2),
    RendererAVX2 = (1 << 3),
    RendererNEON = (1 << 4),
    RendererOCL = (1 << 5),
    RendererHIP = (1 << 6),
};

struct settings_t {
    int w, h;
#if !defined(DISABLE_OCL)
    int platform_index = -1, device_index = -1;
#endif
    bool use_wide_bvh = true;
};

class RegionContext {
    rect_t rect_;

public:
    int iteration = 0;
    std::unique_ptr<float[]> halton_seq;

    explicit RegionContext(const rect_t &rect) : rect_(rect) {}
    rect_t rect() const { return rect_; }
    void Clear() {
        iteration = 0;
        halton_seq = nullptr;
    }
};

class RendererBase {
public:
    virtual ~RendererBase() = default;
    virtual eRendererType type() const = 0;
    virtual std::pair<int, int> size() const = 0;
    virtual const pixel_color_t *get_pixels_ref() const = 0;
    virtual const shl1_data_t *get_sh_data_ref() const = 0;
    virtual void Resize(int w, int h) = 0;
    virtual void Clear(const pixel_color_t &c = {0, 0, 0, 0}) = 0;
    virtual std::shared_ptr<SceneBase> CreateScene() = 0;
    virtual void RenderScene(const std::shared_ptr<SceneBase> &s, RegionContext &region) = 0;

    struct stats_t {
        unsigned long long time_primary_ray_gen_us;
        unsigned long long time_primary_trace_us;
        unsigned long long time_primary_shade_us;
        unsigned long long time_secondary_sort_us;
        unsigned long long time_secondary_trace_us;
        unsigned long long time_secondary_shade_us;
    };
    virtual void GetStats(stats_t &st) = 0;
    virtual void ResetStats() = 0;
};
} // namespace Ray
