// Renderer factory of the Ray renderer API for the HIP backend library.
// Mirrors src/Ray/RendererFactory.h:14-24: CreateRenderer tries the HIP
// backend when its bit is set and throws (so a caller's factory can fall back)
// when no usable device exists.
#pragma once

#include <iostream>
#include <memory>

#include "RendererBase.h"

namespace Ray {
const uint32_t default_renderer_flags = RendererHIP;

std::shared_ptr<RendererBase> CreateRenderer(const settings_t &s, uint32_t flags = default_renderer_flags,
                                             std::ostream &log_stream = std::cout);
} // namespace Ray
