// adjoint.h — launch interface of the path-replay adjoint (adjoint.hip).
#pragma once

#include "render.h"

namespace rt {

// Deepest path the adjoint replays: one bit per bounce records the forward's shadow
// results in a 64-bit mask.
constexpr uint32_t kMaxAdjointBounces = 64;

// Where the adjoint writes (all device memory; each entry may be NULL).
struct AdjointOut {
    float* const* grads;      // device array of num_textures pointers: gradient buffer of
                              // RGBA32F texture i (its texel layout), or NULL
    uint32_t num_textures;
    float* grad_plight;       // d/d intensity of point light i (num_plights floats)
    float* grad_dlight;       // d/d intensity of directional light i (num_dlights floats)
};

// Accumulates into `out` the gradient of sum(grad_accum * accum) over the frame `fa`
// describes.  `table_bounces`: contributions of vertices at bounces below this are
// summed per wave in LDS before their atomics (see adjoint.hip).  `next_pixel`: one
// device counter the persistent kernel takes pixels from (zeroed on `st` here).
int launch_adjoint(const FrameArgs& fa, const float* grad_accum, const AdjointOut& out, uint32_t table_bounces,
                   hipStream_t st, bool any_hit, bool stats, uint32_t* next_pixel);

}  // namespace rt
