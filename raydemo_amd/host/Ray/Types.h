// Public value types of the Ray renderer API, declared for the HIP backend.
// Same names, members and meaning as the reference's src/Ray/Types.h:15-71,
// so code written against the reference compiles against this header.
#pragma once

#include <cstdint>

namespace Ray {
struct pixel_color_t {
    float r, g, b, a;
};
static_assert(sizeof(pixel_color_t) == 16, "RGBA32F pixel");

struct pixel_color8_t {
    uint8_t r, g, b, a;
};
static_assert(sizeof(pixel_color8_t) == 4, "RGBA8 pixel");

struct shl1_data_t {
    float coeff_r[4], coeff_g[4], coeff_b[4];
};

struct rect_t {
    int x, y, w, h;
};

enum eCamType { Persp, Ortho, Geo };
enum eFilterType { Box, Tent };
enum eDeviceType { None, SRGB };

enum ePassFlags {
    SkipDirectLight = (1 << 0),
    SkipIndirectLight = (1 << 1),
    LightingOnly = (1 << 2),
    NoBackground = (1 << 3),
    Clamp = (1 << 4),
    OutputSH = (1 << 5),
    UseCoherentSampling = (1 << 6)
};

struct pass_settings_t {
    uint8_t max_diff_depth, max_glossy_depth, max_refr_depth, max_transp_depth, max_total_depth;
    uint8_t termination_start_depth;
    uint8_t pad[2];
    uint32_t flags;
};

struct camera_t {
    eCamType type;
    eFilterType filter;
    eDeviceType dtype;
    float fov, gamma;
    float focus_distance, focus_factor;
    float origin[3], fwd[3], side[3], up[3];
    uint32_t mi_index, uv_index;
    pass_settings_t pass_settings;
};
} // namespace Ray
