// output.h — the display image RayTracing.rgen writes after accumulation
// (crates/examples/gltf_viewer/shaders/RayTracing.rgen:132-166): the mean radiance
// acc / TotalNumberOfSamples, tone mapping (shaders/lib/Tonemapping.glsl:1-53), the
// HEAT and DISTANCE mappings (shaders/lib/Heatmap.glsl:4-33, rgen:151-163), and the
// store into an 8-bit UNORM image with alpha 1.
//
// Evaluated in fp32 without contraction, in the shaders' operation order, by the
// kernel that retires the pixel (fused with the accumulation write).
#pragma once

#include "common.h"

namespace rt {

struct OutputArgs {
    uint32_t* out;       // width * height packed pixels (NULL: no display image)
    float heat_scale;    // ubo.HeatmapScale
    uint32_t tone_mode;  // ubo.tone_mapping_mode
    uint32_t mapping;    // ubo.mapping
    uint32_t total;      // ubo.TotalNumberOfSamples
    uint32_t bgra;       // byte order B, G, R, A (the swapchain's B8G8R8A8_UNORM)
};

#pragma clang fp contract(off)

// Tonemapping.glsl:9-11 (GAMMA 2.2)
__device__ inline float gamma_encode(float c) { return powf(c, 1.0f / 2.2f); }

// Tonemapping.glsl:15-23
__device__ inline float uncharted2(float c) {
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    return ((c * (A * c + C * B) + D * E) / (c * (A * c + B) + D * F)) - E / F;
}

__device__ inline float tone_map(float c, uint32_t mode) {
    switch (mode) {
        case 0:  // defaultToneMap, :49-52
            return gamma_encode(c / (c + 1.0f));
        case 1: {  // toneMapUncharted, :25-30 (W = 11.2)
            const float white = 1.0f / uncharted2(11.2f);
            return gamma_encode(uncharted2(c * 2.0f) * white);
        }
        case 2: {  // toneMapHejlRichard, :34-37 (no gamma)
            c = fmaxf(0.0f, c - 0.004f);
            return (c * (6.2f * c + .5f)) / (c * (6.2f * c + 1.7f) + 0.06f);
        }
        case 3: {  // toneMapACES, :41-47
            const float A = 2.51f, B = 0.03f, C = 2.43f, D = 0.59f, E = 0.14f;
            float v = (c * (A * c + B)) / (c * (C * c + D) + E);
            return gamma_encode(fminf(fmaxf(v, 0.0f), 1.0f));
        }
        default:  // LINEARtoSRGB
            return gamma_encode(c);
    }
}

__device__ inline float smoothstep_f(float e0, float e1, float x) {
    float t = fminf(fmaxf((x - e0) / (e1 - e0), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

// Heatmap.glsl:4-33: ten colour stops blended with a 0.8-wide smoothstep.
__device__ inline float3 heatmap(float t) {
    constexpr uint8_t stops[10][3] = {{0, 2, 91},     {0, 108, 251}, {0, 221, 221}, {51, 221, 0},  {255, 252, 0},
                                      {255, 180, 0},  {255, 104, 0}, {226, 22, 0},  {191, 0, 83},  {145, 0, 65}};
    const float s = t * 10.0f;
    const int cur = int(s) <= 9 ? int(s) : 9;
    const int prv = cur >= 1 ? cur - 1 : 0;
    const int nxt = cur < 9 ? cur + 1 : 9;
    const float blur = 0.8f;
    const float wc = smoothstep_f(float(cur) - blur, float(cur) + blur, s) *
                     (1.0f - smoothstep_f(float(cur + 1) - blur, float(cur + 1) + blur, s));
    const float wp = 1.0f - smoothstep_f(float(cur) - blur, float(cur) + blur, s);
    const float wn = smoothstep_f(float(cur + 1) - blur, float(cur + 1) + blur, s);
    float r[3];
    for (int k = 0; k < 3; k++) {
        const float v = wc * (float(stops[cur][k]) / 255.0f) + wp * (float(stops[prv][k]) / 255.0f) +
                        wn * (float(stops[nxt][k]) / 255.0f);
        r[k] = fminf(fmaxf(v, 0.0f), 1.0f);
    }
    return mk3(r[0], r[1], r[2]);
}

// Float to 8-bit UNORM: clamp to [0, 1], scale by 255, round to nearest (even).
__device__ inline uint32_t unorm8(float c) {
    c = fminf(fmaxf(c, 0.0f), 1.0f);  // NaN -> 0
    return uint32_t(rintf(c * 255.0f));
}

// Writes the display pixel for accumulated colour `acc` (rgen:132-166).  `last_t`
// is the payload's t after the pixel's last trace (DISTANCE); `ticks` the clock
// ticks the pixel took (HEAT).
__device__ inline void write_output(const OutputArgs& o, uint32_t width, uint32_t x, uint32_t y, float4 acc,
                                    float last_t, uint64_t ticks) {
    const float total = float(o.total);
    float3 c = mk3(tone_map(acc.x / total, o.tone_mode), tone_map(acc.y / total, o.tone_mode),
                   tone_map(acc.z / total, o.tone_mode));
    if (o.mapping == RT_MAPPING_HEAT) {
        const float scale = 1000000.0f * o.heat_scale * o.heat_scale;
        c = heatmap(fminf(fmaxf(float(ticks) / scale, 0.0f), 1.0f));
    } else if (o.mapping == RT_MAPPING_DISTANCE) {
        const float v = fminf((o.heat_scale - fmaxf(last_t, 0.001f)) / o.heat_scale, 1.0f);
        c = mk3(v, v, v);
    }
    const uint32_t r = unorm8(c.x), g = unorm8(c.y), b = unorm8(c.z);
    o.out[size_t(y) * width + x] = o.bgra ? (b | (g << 8) | (r << 16) | (255u << 24)) : (r | (g << 8) | (b << 16) | (255u << 24));
}

#pragma clang fp contract(on)

}  // namespace rt
