// skin.h — vertex skinning (the reference's AnimationCompute.comp) for animated scenes.
#pragma once

#include "common.h"

namespace rt {

constexpr uint32_t kBonesPerSkin = 256;  // MAX_BONES_PER_SKIN, AnimationCompute.comp:4

// out[i] = in[i], skinned by the weighted joint matrices of skin in[i].skin_index when
// it is >= 0.  `skins`: num_skins x 256 column-major mat4.  A vertex whose skin index
// or joint is out of range is copied unchanged and raises *bad.
int launch_skin(hipStream_t st, const rt_vertex* in, rt_vertex* out, uint32_t n, const float* skins,
                uint32_t num_skins, uint32_t* bad);

}  // namespace rt
