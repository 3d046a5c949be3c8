// shade.h — device-side hit shading: the closest-hit, any-hit and miss shaders of
// the reference pipeline and their BRDF library, in fp32.
//
// Restates (paths under crates/examples/gltf_viewer/shaders/):
//   RayTracing.rchit:136-477      closest hit (material resolution, NEE, lobe choice)
//   RayTracing.rahit:44-104       alpha test (MASK cutoff, BLEND stochastic)
//   RayTracing.rmiss:14-45        miss (directional lights, 0.01 ambient)
//   lib/PBR.glsl                  GGX / Smith / Schlick / Frostbite, VNDF sampling
//   lib/Random.glsl               PCG4D counter RNG, TEA seeding, LCG
//   lib/RayTracingCommons.glsl    interpolation, offset_ray
//   lib/Material.glsl, lib/PunctualLight.glsl
#pragma once

#include "common.h"

namespace rt {

// ------------------------------------------------------------------- RNG
struct Rng4 {
    uint32_t x, y, z, w;
};

__device__ inline void pcg4d(uint32_t& x, uint32_t& y, uint32_t& z, uint32_t& w) {  // lib/Random.glsl:81-98
    x = x * 1664525u + 1013904223u;
    y = y * 1664525u + 1013904223u;
    z = z * 1664525u + 1013904223u;
    w = w * 1664525u + 1013904223u;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
    x ^= x >> 16;
    y ^= y >> 16;
    z ^= z >> 16;
    w ^= w >> 16;
    x += y * w;
    y += z * x;
    z += x * y;
    w += y * z;
}

__device__ inline float rng_next(Rng4& s) {  // rand(), lib/Random.glsl:102-105
    s.w += 1u;
    uint32_t x = s.x, y = s.y, z = s.z, w = s.w;
    pcg4d(x, y, z, w);
    return __uint_as_float(0x3f800000u | (x >> 9)) - 1.0f;
}

__host__ __device__ inline uint32_t tea16(uint32_t val0, uint32_t val1) {  // lib/Random.glsl:12-25
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}

__device__ inline float lcg_float(uint32_t& seed) {  // lib/Random.glsl:27-42
    seed = 1664525u * seed + 1013904223u;
    return float(seed & 0x00FFFFFFu) / float(0x01000000);
}

// ----------------------------------------------------------- frame inputs
struct SceneView {
    const rt_vertex* vertices;
    const uint32_t* indices;
    const rt_prim_info* prim_infos;
    const rt_material_raw* materials;
    const rt_light* dlights;
    const rt_light* plights;
    const rt_texture* textures;  // device copy of the descriptor array
    const float* srgb_lut;       // 256 entries
    uint32_t num_dlights, num_plights, num_textures;
    uint32_t num_vertices, num_indices, num_prim_infos, num_materials;
};

// --------------------------------------------------------------- textures
// texture() at LOD 0: magFilter, wrap on integer texel coordinates, sRGB decoded
// per texel before filtering.
__device__ inline int wrap_coord(int i, int size, uint32_t mode) {
    if (mode == RT_WRAP_CLAMP_TO_EDGE) return min(max(i, 0), size - 1);
    if (mode == RT_WRAP_MIRRORED_REPEAT) {
        int t = i % (2 * size);
        if (t < 0) t += 2 * size;
        return t < size ? t : 2 * size - 1 - t;
    }
    int t = i % size;
    return t < 0 ? t + size : t;
}

__device__ inline float4 texel(const rt_texture& t, const float* lut, int x, int y) {
    size_t idx = size_t(y) * t.width + size_t(x);
    if (t.format == RT_FMT_RGBA32F) return reinterpret_cast<const float4*>(t.data)[idx];
    uint32_t v = reinterpret_cast<const uint32_t*>(t.data)[idx];
    uint32_t r = v & 255u, g = (v >> 8) & 255u, b = (v >> 16) & 255u, a = v >> 24;
    if (t.format == RT_FMT_RGBA8_SRGB) return make_float4(lut[r], lut[g], lut[b], float(a) / 255.0f);
    return make_float4(float(r) / 255.0f, float(g) / 255.0f, float(b) / 255.0f, float(a) / 255.0f);
}

__device__ inline float4 sample_texture(const SceneView& sv, int32_t index, float u, float v) {
    if (index < 0 || uint32_t(index) >= sv.num_textures) return make_float4(1.f, 1.f, 1.f, 1.f);
    const rt_texture& t = sv.textures[index];
    int W = int(t.width), H = int(t.height);
    if (t.mag_filter == RT_FILTER_NEAREST) {
        int x = wrap_coord(int(floorf(u * float(W))), W, t.wrap_s);
        int y = wrap_coord(int(floorf(v * float(H))), H, t.wrap_t);
        return texel(t, sv.srgb_lut, x, y);
    }
    float fu = u * float(W) - 0.5f, fv = v * float(H) - 0.5f;
    float iu = floorf(fu), iv = floorf(fv);
    float a = fu - iu, b = fv - iv;
    int x0 = wrap_coord(int(iu), W, t.wrap_s), x1 = wrap_coord(int(iu) + 1, W, t.wrap_s);
    int y0 = wrap_coord(int(iv), H, t.wrap_t), y1 = wrap_coord(int(iv) + 1, H, t.wrap_t);
    float4 t00 = texel(t, sv.srgb_lut, x0, y0), t10 = texel(t, sv.srgb_lut, x1, y0);
    float4 t01 = texel(t, sv.srgb_lut, x0, y1), t11 = texel(t, sv.srgb_lut, x1, y1);
    float w00 = (1.f - a) * (1.f - b), w10 = a * (1.f - b), w01 = (1.f - a) * b, w11 = a * b;
    return make_float4(w00 * t00.x + w10 * t10.x + w01 * t01.x + w11 * t11.x,
                       w00 * t00.y + w10 * t10.y + w01 * t01.y + w11 * t11.y,
                       w00 * t00.z + w10 * t10.z + w01 * t01.z + w11 * t11.z,
                       w00 * t00.w + w10 * t10.w + w01 * t01.w + w11 * t11.w);
}

// getUV, lib/RayTracingCommons.glsl:65-72
__device__ inline float4 tex_uv(const SceneView& sv, rt_texture_info ti, float4 uv) {
    uint32_t c = uint32_t(ti.tex_coord);
    float u = 0.f, v = 0.f;
    if (c == 0u) { u = uv.x; v = uv.y; }
    else if (c == 1u) { u = uv.z; v = uv.w; }
    return sample_texture(sv, ti.index, u, v);
}

__device__ inline float4 mul4(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ inline float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }

// ------------------------------------------------------------------- BRDF
constexpr float kPi = 3.141592653589f;  // PBR.glsl:57
constexpr float kTwoPi = 2.0f * kPi;
constexpr float kOneOverPi = 1.0f / kPi;

__device__ inline float luminance(float3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ inline float pow5(float x) { return x * x * x * x * x; }
__device__ inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ inline float saturate(float x) { return clampf(x, 0.f, 1.f); }

struct MatBrdf {  // PBR.glsl:173-193 (fields the integrator reads)
    float3 base_color;
    float metallic, roughness, ior, transmission, specular_factor;
    float3 specular_color_factor, F0, F90, c_diff;
    bool front_face, volume;
    float3 attenuation_color;
    float attenuation_distance;
};

__device__ inline void mat_build(MatBrdf& m) {  // PBR.glsl:195-206
    float f = (m.ior - 1.f) / (m.ior + 1.f);
    float3 d = m.specular_color_factor * (f * f);
    float3 df0 = mk3(fminf(d.x, 1.f), fminf(d.y, 1.f), fminf(d.z, 1.f)) * m.specular_factor;
    float3 df90 = m.specular_color_factor;
    float mt = m.metallic;
    m.F0 = df0 * (1.f - mt) + m.base_color * mt;
    m.F90 = df90 * (1.f - mt) + mk3(1.f, 1.f, 1.f) * mt;
    m.c_diff = m.base_color * (1.f - mt);
}

__device__ inline float3 fresnel(float3 f0, float f90, float n_dot_s) {  // PBR.glsl:228-231
    float p = pow5(1.f - n_dot_s);
    return f0 + (mk3(f90, f90, f90) - f0) * p;
}
__device__ inline float shadowed_f90(float3 F90) { return fminf(1.f, luminance(F90)); }
__device__ inline float smith_g_a(float alpha, float nds) {  // PBR.glsl:248-250
    return nds / (fmaxf(0.00001f, alpha) * sqrtf(1.f - fminf(0.99999f, nds * nds)));
}
__device__ inline float smith_lambda(float a) { return (-1.f + sqrtf(1.f + 1.f / (a * a))) * 0.5f; }
__device__ inline float smith_g2(float alpha, float ndl, float ndv) {  // PBR.glsl:258-262
    return 1.f / (1.f + smith_lambda(smith_g_a(alpha, ndl)) + smith_lambda(smith_g_a(alpha, ndv)));
}
__device__ inline float ggx_d(float a2, float ndh) {  // PBR.glsl:286-289
    float b = (a2 - 1.f) * ndh * ndh + 1.f;
    return a2 / (kPi * b * b);
}
__device__ inline float smith_g1(float a2, float nds2) {  // PBR.glsl:446-448
    return 2.f / (sqrtf((a2 * (1.f - nds2) + nds2) / nds2) + 1.f);
}
__device__ inline float g2_over_g1(float a2, float ndl, float ndv) {  // PBR.glsl:470-474
    float g1v = smith_g1(a2, ndv * ndv), g1l = smith_g1(a2, ndl * ndl);
    return g1l / (g1v + g1l - g1v * g1l);
}
__device__ inline float frostbite(float r, float ndl, float ndv, float ldh) {  // PBR.glsl:486-496
    float bias = 0.5f * r;
    float factor = 1.f * (1.f - r) + (1.f / 1.51f) * r;
    float fd90m1 = bias + 2.f * ldh * ldh * r - 1.f;
    return (1.f + fd90m1 * pow5(1.f - ndl)) * (1.f + fd90m1 * pow5(1.f - ndv)) * factor;
}

struct BrdfTerms {
    float ndl, ndv, ldh, ndh, vdh;
    bool vback, lback;
};
__device__ inline BrdfTerms brdf_terms(float3 N, float3 L, float3 V) {  // prepareBRDFData, PBR.glsl:527-564
    BrdfTerms d;
    float3 H = normalize3(L + V);
    float ndl = dot3(N, L), ndv = dot3(N, V);
    d.vback = ndv <= 0.f;
    d.lback = ndl <= 0.f;
    d.ndl = fminf(fmaxf(0.00001f, ndl), 1.f);
    d.ndv = fminf(fmaxf(0.00001f, ndv), 1.f);
    d.ldh = saturate(dot3(L, H));
    d.ndh = saturate(dot3(N, H));
    d.vdh = saturate(dot3(V, H));
    return d;
}

__device__ inline float3 eval_combined_brdf(float3 N, float3 L, float3 V, const MatBrdf& m) {  // PBR.glsl:731-745
    BrdfTerms d = brdf_terms(N, L, V);
    if (d.vback || d.lback) return mk3(0.f, 0.f, 0.f);
    float alpha = m.roughness * m.roughness, a2 = alpha * alpha;
    float3 F = fresnel(m.F0, shadowed_f90(m.F90), d.vdh);
    float D = ggx_d(fmaxf(0.00001f, a2), d.ndh);
    float G2 = smith_g2(alpha, d.ndl, d.ndv);
    float3 spec = ((F * (G2 * D)) / (4.f * d.ndl * d.ndv)) * d.ndl;
    float3 diff = m.c_diff * (frostbite(m.roughness, d.ndl, d.ndv, d.ldh) * kOneOverPi * d.ndl);
    return diff + spec;
}

__device__ inline float3 brdf_probability(const MatBrdf& m, float3 V, float3 N) {  // PBR.glsl:324-358
    float sf0 = luminance(m.F0);
    float dr = luminance(m.c_diff);
    float fr = saturate(luminance(fresnel(mk3(sf0, sf0, sf0), shadowed_f90(m.F90), fmaxf(0.f, dot3(V, N)))));
    float specular = fr * m.specular_factor;
    float pen = dr * (1.f - m.specular_factor * fr);
    float diffuse = pen * (1.f - m.transmission);
    float trans = pen * m.transmission;
    float sum = fmaxf(0.0001f, specular + diffuse + trans);
    float p = clampf(specular / sum, 0.001f, 0.9f);
    float dd = (1.f - p) * (1.f - m.transmission);
    float tt = (1.f - p) * m.transmission;
    sum = p + dd + tt;
    return mk3(p / sum, dd / sum, tt / sum);
}

__device__ inline float4 rotation_to_z(float3 v) {  // PBR.glsl:362-368
    if (v.z < -0.99999f) return make_float4(1.f, 0.f, 0.f, 0.f);
    float4 q = make_float4(v.y, -v.x, 0.f, 1.f + v.z);
    float l = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return make_float4(q.x / l, q.y / l, q.z / l, q.w / l);
}
__device__ inline float3 rotate_point(float4 q, float3 v) {  // PBR.glsl:373-376
    float3 a = mk3(q.x, q.y, q.z);
    return a * (2.f * dot3(a, v)) + v * (q.w * q.w - dot3(a, a)) + cross3(a, v) * (2.f * q.w);
}

__device__ inline float3 sample_vndf(float3 Ve, float alpha, float ux, float uy) {  // PBR.glsl:413-436
    float3 Vh = normalize3(mk3(alpha * Ve.x, alpha * Ve.y, Ve.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    float3 T1 = lensq > 0.f ? mk3(-Vh.y, Vh.x, 0.f) * (1.f / sqrtf(lensq)) : mk3(1.f, 0.f, 0.f);
    float3 T2 = cross3(Vh, T1);
    float r = sqrtf(ux);
    float phi = kTwoPi * uy;
    float sp, cp;
    sincosf(phi, &sp, &cp);
    float t1 = r * cp, t2 = r * sp;
    float s = 0.5f * (1.f + Vh.z);
    t2 = sqrtf(1.f - t1 * t1) * (1.f - s) + t2 * s;
    float3 Nh = T1 * t1 + T2 * t2 + Vh * sqrtf(fmaxf(0.f, 1.f - t1 * t1 - t2 * t2));
    return normalize3(mk3(alpha * Nh.x, alpha * Nh.y, fmaxf(0.f, Nh.z)));
}

enum : uint32_t { kDiffuse = 1, kSpecular = 2, kTransmission = 3 };

__device__ inline float3 refract3(float3 I, float3 N, float eta) {  // GLSL refract
    float c = dot3(N, I);
    float k = 1.f - eta * eta * (1.f - c * c);
    if (k < 0.f) return mk3(0.f, 0.f, 0.f);
    return I * eta - N * (eta * c + sqrtf(k));
}

// evalIndirectCombinedBRDF, PBR.glsl:649-729 (+ sampleSpecularMicrofacet :587-616)
__device__ inline bool sample_indirect(float ux, float uy, float3 sn, float3 gn, float3 V, const MatBrdf& m,
                                       uint32_t type, float3& dir, float3& weight, float& volume_dis) {
    if (dot3(gn, V) < 0.f) return false;
    float4 q = rotation_to_z(sn);
    float3 Vl = rotate_point(q, V);
    float3 dl;
    float alpha = m.roughness * m.roughness, a2 = alpha * alpha;
    if (type == kDiffuse) {
        float a = sqrtf(ux), sp, cp;
        sincosf(kTwoPi * uy, &sp, &cp);
        dl = mk3(a * cp, a * sp, sqrtf(1.f - ux));
        BrdfTerms d = brdf_terms(mk3(0.f, 0.f, 1.f), dl, Vl);
        float3 F = fresnel(m.F0, shadowed_f90(m.F90), d.vdh);
        float fb = frostbite(m.roughness, d.ndl, d.ndv, d.ldh);
        weight = (mk3(1.f, 1.f, 1.f) - F * m.specular_factor) * m.c_diff * fb;
        weight = weight * (1.f - m.transmission);
    } else if (type == kSpecular) {
        float3 Hl = alpha == 0.f ? mk3(0.f, 0.f, 1.f) : sample_vndf(Vl, alpha, ux, uy);
        float3 nv = -Vl;
        float3 Ll = nv - Hl * (2.f * dot3(Hl, nv));
        float hdl = fmaxf(0.00001f, fminf(1.f, dot3(Hl, Ll)));
        float ndl = fmaxf(0.00001f, fminf(1.f, Ll.z));
        float ndv = fmaxf(0.00001f, fminf(1.f, Vl.z));
        float3 F = fresnel(m.F0, shadowed_f90(m.F90), hdl);
        weight = F * g2_over_g1(a2, ndl, ndv);
        dl = Ll;
        weight = weight * m.specular_factor;
    } else {
        if (m.volume) {
            float eta = m.front_face ? 1.f / m.ior : m.ior;
            float3 rf = refract3(-Vl, mk3(0.f, 0.f, 1.f), eta);
            if (rf.x == 0.f && rf.y == 0.f && rf.z == 0.f) {
                weight = mk3(0.f, 0.f, 0.f);
                return false;
            }
            dl = rf;
        } else {
            dl = -Vl;
        }
        float3 w = m.c_diff * m.transmission;
        weight = mk3(fmaxf(0.f, w.x), fmaxf(0.f, w.y), fmaxf(0.f, w.z));
        if (!m.front_face && m.volume) {
            float dis = volume_dis;
            volume_dis = -1.f;
            float3 att = mk3(expf(logf(m.attenuation_color.x) / m.attenuation_distance * dis),
                             expf(logf(m.attenuation_color.y) / m.attenuation_distance * dis),
                             expf(logf(m.attenuation_color.z) / m.attenuation_distance * dis));
            weight = weight * mk3(fminf(att.x, 1.f), fminf(att.y, 1.f), fminf(att.z, 1.f));
        }
    }
    if (luminance(weight) == 0.f) return false;
    dir = normalize3(rotate_point(make_float4(-q.x, -q.y, -q.z, q.w), dl));
    return true;
}

__device__ inline float convert_metallic(float3 diffuse, float3 specular, float max_spec) {  // lib/Material.glsl:76-89
    const float cmin = 0.04f;
    float pd = sqrtf(0.299f * diffuse.x * diffuse.x + 0.587f * diffuse.y * diffuse.y + 0.114f * diffuse.z * diffuse.z);
    float ps = sqrtf(0.299f * specular.x * specular.x + 0.587f * specular.y * specular.y + 0.114f * specular.z * specular.z);
    if (ps < cmin) return 0.f;
    float b = pd * (1.f - max_spec) / (1.f - cmin) + ps - 2.f * cmin;
    float c = cmin - ps;
    float D = fmaxf(b * b - 4.f * cmin * c, 0.f);
    return clampf((-b + sqrtf(D)) / (2.f * cmin), 0.f, 1.f);
}

__device__ inline float3 light_intensity_at(const rt_light& l, float dist) {  // lib/PunctualLight.glsl:17-30
    float3 c = mk3(l.color[0], l.color[1], l.color[2]) * l.intensity;
    if (l.kind == RT_LIGHT_POINT) {
        const float r2 = 0.25f;
        float d2 = dist * dist;
        return c * (2.f / (d2 + r2 + dist * sqrtf(d2 + r2)));
    }
    return c;
}

__device__ inline float3 offset_ray(float3 p, float3 n) {  // lib/RayTracingCommons.glsl:106-122
    const float origin = 1.f / 32.f, fs = 1.f / 65536.f, is = 256.f;
    float pp[3] = {p.x, p.y, p.z}, nn[3] = {n.x, n.y, n.z}, out[3];
    for (int k = 0; k < 3; k++) {
        int of = int(nn[k] * is);
        int bits = __float_as_int(pp[k]) + ((pp[k] < 0.f) ? -of : of);
        out[k] = fabsf(pp[k]) < origin ? pp[k] + fs * nn[k] : __int_as_float(bits);
    }
    return mk3(out[0], out[1], out[2]);
}

// ----------------------------------------------------------------- any-hit
// RayTracing.rahit / RayTracing.shadow.rahit:44-104.  True = ignore the candidate.
__device__ inline bool alpha_test_ignores(const SceneView& sv, uint32_t custom_index, uint32_t prim, float bu, float bv,
                                          Rng4& rng) {
    if (custom_index >= sv.num_prim_infos) return false;
    const rt_prim_info pi = sv.prim_infos[custom_index];
    if (pi.material_id >= sv.num_materials) return false;
    const rt_material_raw& mat = sv.materials[pi.material_id];
    uint32_t mode = mat.alpha_mode;
    if (mode == RT_ALPHA_OPAQUE) return false;
    uint32_t io = pi.i_offset + 3u * prim;
    if (io + 2u >= sv.num_indices) return false;
    uint32_t i0 = pi.v_offset + sv.indices[io], i1 = pi.v_offset + sv.indices[io + 1], i2 = pi.v_offset + sv.indices[io + 2];
    if (i0 >= sv.num_vertices || i1 >= sv.num_vertices || i2 >= sv.num_vertices) return false;
    const float4* a = reinterpret_cast<const float4*>(sv.vertices + i0);
    const float4* b = reinterpret_cast<const float4*>(sv.vertices + i1);
    const float4* c = reinterpret_cast<const float4*>(sv.vertices + i2);
    float b0 = 1.f - bu - bv;
    float4 ua = a[6], ub = b[6], uc = c[6];
    float4 uv = make_float4(ua.x * b0 + ub.x * bu + uc.x * bv, ua.y * b0 + ub.y * bu + uc.y * bv,
                            ua.z * b0 + ub.z * bu + uc.z * bv, ua.w * b0 + ub.w * bu + uc.w * bv);
    float4 ca = a[3], cb = b[3], cc = c[3];
    float4 vcol = make_float4(ca.x * b0 + cb.x * bu + cc.x * bv, ca.y * b0 + cb.y * bu + cc.y * bv,
                              ca.z * b0 + cb.z * bu + cc.z * bv, ca.w * b0 + cb.w * bu + cc.w * bv);
    float4 col = mul4(vcol, ld4(mat.base_color));
    if (mat.base_color_texture.index >= 0) col = mul4(col, tex_uv(sv, mat.base_color_texture, uv));
    float opacity = col.w;
    if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) {
        float4 df = ld4(mat.sg_diffuse_factor);
        if (mat.sg_diffuse_texture.index >= 0) df = mul4(df, tex_uv(sv, mat.sg_diffuse_texture, uv));
        opacity = vcol.w * df.w;
    }
    if (mode == RT_ALPHA_MASK) return opacity < mat.alpha_cutoff;
    return opacity <= rng_next(rng);
}

}  // namespace rt
