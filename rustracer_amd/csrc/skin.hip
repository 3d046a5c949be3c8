// skin.hip — linear-blend skinning of the vertex buffer, one thread per vertex
// (crates/examples/gltf_viewer/shaders/AnimationCompute.comp:14-38).  Streams 128 B in
// and 128 B out per vertex; the joint matrices (16 KB per skin) stay in cache.
#include "skin.h"

namespace rt {

namespace {

constexpr int kSkinBlock = 256;

// column-major mat4 times (x, y, z, w)
__device__ inline float4 mat_vec(const float m[16], float x, float y, float z, float w) {
    return make_float4(m[0] * x + m[4] * y + m[8] * z + m[12] * w, m[1] * x + m[5] * y + m[9] * z + m[13] * w,
                       m[2] * x + m[6] * y + m[10] * z + m[14] * w, m[3] * x + m[7] * y + m[11] * z + m[15] * w);
}

__global__ __launch_bounds__(kSkinBlock) void k_skin(const rt_vertex* in, rt_vertex* out, uint32_t n,
                                                      const float* skins, uint32_t num_skins, uint32_t* bad) {
    const uint32_t i = blockIdx.x * kSkinBlock + threadIdx.x;
    if (i >= n) return;
    const float4* src = reinterpret_cast<const float4*>(in + i);
    float4 w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = src[k];
    // w[0] pos, w[1] normal, w[2] tangent, w[4] weights, w[5] joints, w[7].x skin_index
    const int skin = int(__float_as_uint(w[7].x));
    if (skin >= 0) {
        const uint32_t j[4] = {__float_as_uint(w[5].x), __float_as_uint(w[5].y), __float_as_uint(w[5].z),
                               __float_as_uint(w[5].w)};
        if (uint32_t(skin) >= num_skins || j[0] >= kBonesPerSkin || j[1] >= kBonesPerSkin || j[2] >= kBonesPerSkin ||
            j[3] >= kBonesPerSkin) {
            atomicOr(bad, kErrBadSkin);
        } else {
            // transform = weights.x * bones[joints.x] + ... + weights.w * bones[joints.w]
            const float* b = skins + size_t(skin) * kBonesPerSkin * 16;
            const float wt[4] = {w[4].x, w[4].y, w[4].z, w[4].w};
            float m[16];
#pragma unroll
            for (int e = 0; e < 16; e++) m[e] = wt[0] * b[16 * j[0] + e];
#pragma unroll
            for (int k = 1; k < 4; k++)
#pragma unroll
                for (int e = 0; e < 16; e++) m[e] = m[e] + wt[k] * b[16 * j[k] + e];
            const float4 p = mat_vec(m, w[0].x, w[0].y, w[0].z, 1.f);
            w[0].x = p.x;
            w[0].y = p.y;
            w[0].z = p.z;
            const float4 nn = mat_vec(m, w[1].x, w[1].y, w[1].z, 0.f);
            const float nl = 1.0f / sqrtf(nn.x * nn.x + nn.y * nn.y + nn.z * nn.z);
            w[1].x = nn.x * nl;
            w[1].y = nn.y * nl;
            w[1].z = nn.z * nl;
            const float4 t = mat_vec(m, w[2].x, w[2].y, w[2].z, 0.f);
            const float tl = 1.0f / sqrtf(t.x * t.x + t.y * t.y + t.z * t.z + t.w * t.w);
            w[2] = make_float4(t.x * tl, t.y * tl, t.z * tl, w[2].w);  // tangent.w kept
        }
    }
    float4* dst = reinterpret_cast<float4*>(out + i);
#pragma unroll
    for (int k = 0; k < 8; k++) dst[k] = w[k];
}

}  // namespace

int launch_skin(hipStream_t st, const rt_vertex* in, rt_vertex* out, uint32_t n, const float* skins,
                uint32_t num_skins, uint32_t* bad) {
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_skin, dim3((n + kSkinBlock - 1) / kSkinBlock), dim3(kSkinBlock), 0, st, in, out, n, skins,
                       num_skins, bad);
    return int(hipGetLastError());
}

}  // namespace rt
