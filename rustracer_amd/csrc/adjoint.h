// adjoint.h — launch interface of the path-replay adjoint (adjoint.hip).
#pragma once

#include "render.h"

namespace rt {

// Bounces one adjoint lane records per sample (paths are at most this deep).
constexpr uint32_t kMaxAdjointBounces = 16;

// Accumulates into d_grads[i] (device buffers in the layout of RGBA32F texture i, or
// NULL) the gradient of sum(grad_accum * accum) over the frame `fa` describes.
// d_grads is a device array of num_textures pointers.  `rec` (or NULL) holds the
// frame's path records from the forward trace; the adjoint then replays them
// instead of tracing the paths again.
int launch_adjoint(const FrameArgs& fa, const float* grad_accum, float* const* d_grads, uint32_t num_textures,
                   const PathRecord* rec, hipStream_t st, bool any_hit);

}  // namespace rt
