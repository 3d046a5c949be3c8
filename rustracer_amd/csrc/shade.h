// shade.h — device-side hit shading: the closest-hit, any-hit and miss shaders of
// the reference pipeline and their BRDF library.
//
// Precision: shading runs in fp64.  With roughness down to 0.05 the GGX lobe has a
// relative sensitivity of ~7e5 to n.h, so fp32 rounding of normals, half vectors
// and dot products moves highlight values by several percent; the fp64 chain keeps
// the image within 1e-3 of an FP64 evaluation of the same algorithm.  Traversal
// stays fp32 (it only decides which triangle is hit).
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

// rand(), lib/Random.glsl:102-105: the value is k / 2^23, exact in fp32 and fp64.
__device__ inline float rng_next(Rng4& s) {
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
    uint32_t has_sky;            // ubo.HasSky
    rt_texture sky;              // six square faces +x, -x, +y, -y, +z, -z
};

struct D4 {
    double x, y, z, w;
};
__device__ inline D4 mul4(D4 a, D4 b) { return D4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
__device__ inline D4 ld4(const float* p) {
    float4 f = *reinterpret_cast<const float4*>(p);
    return D4{f.x, f.y, f.z, f.w};
}
__device__ inline D3 xyz(D4 a) { return D3{a.x, a.y, a.z}; }
__device__ inline D3 ld3(const float* p) { return D3{p[0], p[1], p[2]}; }

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

__device__ inline D4 texel_at(const rt_texture& t, const float* lut, size_t idx) {
    if (t.format == RT_FMT_RGBA32F) {
        float4 f = ldg_f4(reinterpret_cast<const float4*>(t.data) + idx);
        return D4{f.x, f.y, f.z, f.w};
    }
    uint32_t v = ldg_u32(reinterpret_cast<const uint32_t*>(t.data) + idx);
    uint32_t r = v & 255u, g = (v >> 8) & 255u, b = (v >> 16) & 255u, a = v >> 24;
    if (t.format == RT_FMT_RGBA8_SRGB) return D4{lut[r], lut[g], lut[b], a / 255.0};
    return D4{r / 255.0, g / 255.0, b / 255.0, a / 255.0};
}
__device__ inline D4 texel(const rt_texture& t, const float* lut, int x, int y) {
    return texel_at(t, lut, size_t(y) * t.width + size_t(x));
}

__device__ inline double bilerp(double t00, double t10, double t01, double t11, double a, double b) {
    double top = t00 + a * (t10 - t00), bot = t01 + a * (t11 - t01);
    return top + b * (bot - top);
}

__device__ inline D4 sample_texture(const SceneView& sv, int32_t index, double u, double v) {
    if (index < 0 || uint32_t(index) >= sv.num_textures) return D4{1.0, 1.0, 1.0, 1.0};
    const rt_texture& t = sv.textures[index];
    int W = int(t.width), H = int(t.height);
    if (t.mag_filter == RT_FILTER_NEAREST) {
        int x = wrap_coord(int(floor(u * W)), W, t.wrap_s);
        int y = wrap_coord(int(floor(v * H)), H, t.wrap_t);
        return texel(t, sv.srgb_lut, x, y);
    }
    double fu = u * W - 0.5, fv = v * H - 0.5;
    double iu = floor(fu), iv = floor(fv);
    double a = fu - iu, b = fv - iv;
    int x0 = wrap_coord(int(iu), W, t.wrap_s), x1 = wrap_coord(int(iu) + 1, W, t.wrap_s);
    int y0 = wrap_coord(int(iv), H, t.wrap_t), y1 = wrap_coord(int(iv) + 1, H, t.wrap_t);
    D4 t00 = texel(t, sv.srgb_lut, x0, y0), t10 = texel(t, sv.srgb_lut, x1, y0);
    D4 t01 = texel(t, sv.srgb_lut, x0, y1), t11 = texel(t, sv.srgb_lut, x1, y1);
    // two nested lerps x + w (y - x): a constant neighbourhood filters to exactly
    // that constant, as texture hardware (fixed-point weights) gives it, whatever
    // the rounding or contraction of the arithmetic
    return D4{bilerp(t00.x, t10.x, t01.x, t11.x, a, b), bilerp(t00.y, t10.y, t01.y, t11.y, a, b),
              bilerp(t00.z, t10.z, t01.z, t11.z, a, b), bilerp(t00.w, t10.w, t01.w, t11.w, a, b)};
}

// --------------------------------------------------------------- cube map
// texture(samplerCube, dir) at LOD 0 for the skybox (RayTracing.rmiss:27-33), by the
// Vulkan rules: the face is the direction's largest-magnitude axis, (s, t) come
// from the face table below, bilinear texels that fall off the face are taken from
// the neighbouring face (seamless edges) and a texel off both edges is the mean of
// the three texels that meet at that cube corner.  Equal magnitudes pick z, then y
// (the order AMD's cube-face instruction uses).
//   face  major  sc   tc
//   0     +x     -z   -y
//   1     -x     +z   -y
//   2     +y     +x   +z
//   3     -y     +x   -z
//   4     +z     +x   -y
//   5     -z     -x   -y
constexpr uint32_t kMaxCubeFace = 16384;  // face size limit: the edge projection's W (2W + 2) fits in int

struct CubeAxes {
    int ma, sc, tc;         // axis index (0 x, 1 y, 2 z)
    int ma_s, sc_s, tc_s;   // +1 / -1
};
__device__ inline CubeAxes cube_axes(int face) {
    switch (face) {
        case 0: return CubeAxes{0, 2, 1, 1, -1, -1};
        case 1: return CubeAxes{0, 2, 1, -1, 1, -1};
        case 2: return CubeAxes{1, 0, 2, 1, 1, 1};
        case 3: return CubeAxes{1, 0, 2, -1, 1, -1};
        case 4: return CubeAxes{2, 0, 1, 1, 1, -1};
        default: return CubeAxes{2, 0, 1, -1, -1, -1};
    }
}
template <class T> __device__ inline int cube_face_of(T x, T y, T z) {
    const T ax = x < T(0) ? -x : x, ay = y < T(0) ? -y : y, az = z < T(0) ? -z : z;
    if (az >= ax && az >= ay) return z < T(0) ? 5 : 4;
    if (ay >= ax) return y < T(0) ? 3 : 2;
    return x < T(0) ? 1 : 0;
}
template <class T> __device__ inline T pick3(T x, T y, T z, int axis) { return axis == 0 ? x : (axis == 1 ? y : z); }

// Linear texel index of (i, j) on `face` when exactly one of i, j is one texel off
// the face: the texel centre, as an integer point of the cube scaled by W (a face
// spans -W..W), has its largest component W + 1 on the neighbouring face's axis;
// project it there.
__device__ inline uint32_t cube_edge_index(int W, int face, int i, int j) {
    const CubeAxes f = cube_axes(face);
    const int a = f.ma_s * W, b = f.sc_s * (2 * i + 1 - W), c = f.tc_s * (2 * j + 1 - W);
    const int px = f.ma == 0 ? a : (f.sc == 0 ? b : c);
    const int py = f.ma == 1 ? a : (f.sc == 1 ? b : c);
    const int pz = f.ma == 2 ? a : (f.sc == 2 ? b : c);
    const int g = cube_face_of(px, py, pz);
    const CubeAxes h = cube_axes(g);
    const int m = W + 1;
    const int sc = h.sc_s * pick3(px, py, pz, h.sc), tc = h.tc_s * pick3(px, py, pz, h.tc);
    return (uint32_t(g) * uint32_t(W) + uint32_t((W * (tc + m)) / (2 * m))) * uint32_t(W) +
           uint32_t((W * (sc + m)) / (2 * m));
}

// Texel (i, j) of `face`, i and j at most one texel off it.  The index arithmetic
// runs first and the fetches share one loop, so the shading kernels that inline
// the miss shader keep their register budget.
__device__ inline D4 cube_texel(const rt_texture& t, const float* lut, int face, int i, int j) {
    const int W = int(t.width);
    const bool oi = i < 0 || i >= W, oj = j < 0 || j >= W;
    uint32_t idx0, idx1 = 0, idx2 = 0;
    int n = 1;
    if (!oi && !oj) {
        idx0 = (uint32_t(face) * uint32_t(W) + uint32_t(j)) * uint32_t(W) + uint32_t(i);
    } else if (oi != oj) {
        idx0 = cube_edge_index(W, face, i, j);
    } else {
        const int ic = min(max(i, 0), W - 1), jc = min(max(j, 0), W - 1);
        idx0 = (uint32_t(face) * uint32_t(W) + uint32_t(jc)) * uint32_t(W) + uint32_t(ic);
        idx1 = cube_edge_index(W, face, i, jc);
        idx2 = cube_edge_index(W, face, ic, j);
        n = 3;
    }
    D4 acc = D4{0.0, 0.0, 0.0, 0.0};
#pragma unroll 1
    for (int k = 0; k < n; k++) {
        const D4 v = texel_at(t, lut, k == 0 ? idx0 : (k == 1 ? idx1 : idx2));
        acc = D4{acc.x + v.x, acc.y + v.y, acc.z + v.z, acc.w + v.w};
    }
    if (n == 3) acc = D4{acc.x / 3.0, acc.y / 3.0, acc.z / 3.0, acc.w / 3.0};
    return acc;
}

__device__ inline D4 sample_cube(const rt_texture& t, const float* lut, D3 d) {
    const int face = cube_face_of(d.x, d.y, d.z);
    const CubeAxes f = cube_axes(face);
    const double ma = fabs(pick3(d.x, d.y, d.z, f.ma));
    const double s = 0.5 * (f.sc_s * pick3(d.x, d.y, d.z, f.sc) / ma + 1.0);
    const double u = 0.5 * (f.tc_s * pick3(d.x, d.y, d.z, f.tc) / ma + 1.0);
    const int W = int(t.width);
    if (t.mag_filter == RT_FILTER_NEAREST)
        return cube_texel(t, lut, face, min(int(floor(s * W)), W), min(int(floor(u * W)), W));
    const double fu = s * W - 0.5, fv = u * W - 0.5;
    const double iu = floor(fu), iv = floor(fv);
    const double a = fu - iu, b = fv - iv;
    const int i0 = int(iu), j0 = int(iv);
    // two rows of two texels, each row lerped in a, then the rows in b: the nested
    // lerps of sample_texture
    D4 top = D4{0.0, 0.0, 0.0, 0.0}, bot = top;
#pragma unroll 1
    for (int r = 0; r < 2; r++) {
        const D4 x0 = cube_texel(t, lut, face, i0, j0 + r), x1 = cube_texel(t, lut, face, i0 + 1, j0 + r);
        const D4 row = D4{x0.x + a * (x1.x - x0.x), x0.y + a * (x1.y - x0.y), x0.z + a * (x1.z - x0.z),
                          x0.w + a * (x1.w - x0.w)};
        if (r == 0) top = row;
        else bot = row;
    }
    return D4{top.x + b * (bot.x - top.x), top.y + b * (bot.y - top.y), top.z + b * (bot.z - top.z),
              top.w + b * (bot.w - top.w)};
}

// getUV, lib/RayTracingCommons.glsl:65-72
__device__ inline D4 tex_uv(const SceneView& sv, rt_texture_info ti, D4 uv) {
    uint32_t c = uint32_t(ti.tex_coord);
    double u = 0, v = 0;
    if (c == 0u) { u = uv.x; v = uv.y; }
    else if (c == 1u) { u = uv.z; v = uv.w; }
    return sample_texture(sv, ti.index, u, v);
}

// ------------------------------------------------------------------- BRDF
constexpr double kPi = 3.141592653589;  // PBR.glsl:57
constexpr double kTwoPi = 2.0 * kPi;
constexpr double kOneOverPi = 1.0 / kPi;

__device__ inline double luminance(D3 c) { return c.x * 0.2126 + c.y * 0.7152 + c.z * 0.0722; }
__device__ inline double pow5(double x) { return x * x * x * x * x; }
__device__ inline double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
__device__ inline double saturate(double x) { return clampd(x, 0.0, 1.0); }

struct MatBrdf {  // PBR.glsl:173-193 (fields the integrator reads)
    D3 base_color;
    double metallic, roughness, ior, transmission, specular_factor;
    D3 specular_color_factor, F0, F90, c_diff;
    bool front_face, volume;
    D3 attenuation_color;
    double attenuation_distance;
};

__device__ inline void mat_build(MatBrdf& m) {  // PBR.glsl:195-206
    double f = (m.ior - 1.0) / (m.ior + 1.0);
    D3 d = m.specular_color_factor * (f * f);
    D3 df0 = d3(fmin(d.x, 1.0), fmin(d.y, 1.0), fmin(d.z, 1.0)) * m.specular_factor;
    D3 df90 = m.specular_color_factor;
    double mt = m.metallic;
    m.F0 = df0 * (1.0 - mt) + m.base_color * mt;
    m.F90 = df90 * (1.0 - mt) + d3(1.0) * mt;
    m.c_diff = m.base_color * (1.0 - mt);
}

__device__ inline D3 fresnel(D3 f0, double f90, double n_dot_s) {  // PBR.glsl:228-231
    double p = pow5(1.0 - n_dot_s);
    return f0 + (d3(f90) - f0) * p;
}
__device__ inline double shadowed_f90(D3 F90) { return fmin(1.0, luminance(F90)); }
__device__ inline double smith_g_a(double alpha, double nds) {  // PBR.glsl:248-250
    return nds / (fmax(0.00001, alpha) * sqrt(1.0 - fmin(0.99999, nds * nds)));
}
__device__ inline double smith_lambda(double a) { return (-1.0 + sqrt(1.0 + 1.0 / (a * a))) * 0.5; }
__device__ inline double smith_g2(double alpha, double ndl, double ndv) {  // PBR.glsl:258-262
    return 1.0 / (1.0 + smith_lambda(smith_g_a(alpha, ndl)) + smith_lambda(smith_g_a(alpha, ndv)));
}
__device__ inline double ggx_d(double a2, double ndh) {  // PBR.glsl:286-289
    double b = (a2 - 1.0) * ndh * ndh + 1.0;
    return a2 / (kPi * b * b);
}
__device__ inline double smith_g1(double a2, double nds2) {  // PBR.glsl:446-448
    return 2.0 / (sqrt((a2 * (1.0 - nds2) + nds2) / nds2) + 1.0);
}
__device__ inline double g2_over_g1(double a2, double ndl, double ndv) {  // PBR.glsl:470-474
    double g1v = smith_g1(a2, ndv * ndv), g1l = smith_g1(a2, ndl * ndl);
    return g1l / (g1v + g1l - g1v * g1l);
}
__device__ inline double frostbite(double r, double ndl, double ndv, double ldh) {  // PBR.glsl:486-496
    double bias = 0.5 * r;
    double factor = 1.0 * (1.0 - r) + (1.0 / 1.51) * r;
    double fd90m1 = bias + 2.0 * ldh * ldh * r - 1.0;
    return (1.0 + fd90m1 * pow5(1.0 - ndl)) * (1.0 + fd90m1 * pow5(1.0 - ndv)) * factor;
}

struct BrdfTerms {
    double ndl, ndv, ldh, ndh, vdh;
    bool vback, lback;
};
__device__ inline BrdfTerms brdf_terms(D3 N, D3 L, D3 V) {  // prepareBRDFData, PBR.glsl:527-564
    BrdfTerms d;
    D3 H = dnormalize(L + V);
    double ndl = ddot(N, L), ndv = ddot(N, V);
    d.vback = ndv <= 0.0;
    d.lback = ndl <= 0.0;
    d.ndl = fmin(fmax(0.00001, ndl), 1.0);
    d.ndv = fmin(fmax(0.00001, ndv), 1.0);
    d.ldh = saturate(ddot(L, H));
    d.ndh = saturate(ddot(N, H));
    d.vdh = saturate(ddot(V, H));
    return d;
}

__device__ inline D3 eval_combined_brdf(D3 N, D3 L, D3 V, const MatBrdf& m) {  // PBR.glsl:731-745
    BrdfTerms d = brdf_terms(N, L, V);
    if (d.vback || d.lback) return d3(0.0);
    double alpha = m.roughness * m.roughness, a2 = alpha * alpha;
    D3 F = fresnel(m.F0, shadowed_f90(m.F90), d.vdh);
    double D = ggx_d(fmax(0.00001, a2), d.ndh);
    double G2 = smith_g2(alpha, d.ndl, d.ndv);
    D3 spec = ((F * (G2 * D)) / (4.0 * d.ndl * d.ndv)) * d.ndl;
    D3 diff = m.c_diff * (frostbite(m.roughness, d.ndl, d.ndv, d.ldh) * kOneOverPi * d.ndl);
    return diff + spec;
}

__device__ inline D3 brdf_probability(const MatBrdf& m, D3 V, D3 N) {  // PBR.glsl:324-358
    double sf0 = luminance(m.F0);
    double dr = luminance(m.c_diff);
    double fr = saturate(luminance(fresnel(d3(sf0), shadowed_f90(m.F90), fmax(0.0, ddot(V, N)))));
    double specular = fr * m.specular_factor;
    double pen = dr * (1.0 - m.specular_factor * fr);
    double diffuse = pen * (1.0 - m.transmission);
    double trans = pen * m.transmission;
    double sum = fmax(0.0001, specular + diffuse + trans);
    double p = clampd(specular / sum, 0.001, 0.9);
    double dd = (1.0 - p) * (1.0 - m.transmission);
    double tt = (1.0 - p) * m.transmission;
    sum = p + dd + tt;
    return d3(p / sum, dd / sum, tt / sum);
}

struct Quat {
    double x, y, z, w;
};
__device__ inline Quat rotation_to_z(D3 v) {  // PBR.glsl:362-368
    if (v.z < -0.99999) return Quat{1.0, 0.0, 0.0, 0.0};
    Quat q{v.y, -v.x, 0.0, 1.0 + v.z};
    double l = sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    return Quat{q.x / l, q.y / l, q.z / l, q.w / l};
}
__device__ inline D3 rotate_point(Quat q, D3 v) {  // PBR.glsl:373-376
    D3 a = d3(q.x, q.y, q.z);
    return a * (2.0 * ddot(a, v)) + v * (q.w * q.w - ddot(a, a)) + dcross(a, v) * (2.0 * q.w);
}

__device__ inline D3 sample_vndf(D3 Ve, double alpha, double ux, double uy) {  // PBR.glsl:413-436
    D3 Vh = dnormalize(d3(alpha * Ve.x, alpha * Ve.y, Ve.z));
    double lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    D3 T1 = lensq > 0.0 ? d3(-Vh.y, Vh.x, 0.0) * (1.0 / sqrt(lensq)) : d3(1.0, 0.0, 0.0);
    D3 T2 = dcross(Vh, T1);
    double r = sqrt(ux);
    double phi = kTwoPi * uy;
    double sp, cp;
    sincos(phi, &sp, &cp);
    double t1 = r * cp, t2 = r * sp;
    double s = 0.5 * (1.0 + Vh.z);
    t2 = sqrt(1.0 - t1 * t1) * (1.0 - s) + t2 * s;
    D3 Nh = T1 * t1 + T2 * t2 + Vh * sqrt(fmax(0.0, 1.0 - t1 * t1 - t2 * t2));
    return dnormalize(d3(alpha * Nh.x, alpha * Nh.y, fmax(0.0, Nh.z)));
}

enum : uint32_t { kDiffuse = 1, kSpecular = 2, kTransmission = 3 };

__device__ inline D3 refract3(D3 I, D3 N, double eta) {  // GLSL refract
    double c = ddot(N, I);
    double k = 1.0 - eta * eta * (1.0 - c * c);
    if (k < 0.0) return d3(0.0);
    return I * eta - N * (eta * c + sqrt(k));
}
__device__ inline D3 reflect3(D3 I, D3 N) { return I - N * (2.0 * ddot(N, I)); }  // GLSL reflect

// Schlick's reflectance, lib/PBR.glsl:48-53.
__device__ inline double schlick(double cosine, double ri) {
    double r0 = (1.0 - ri) / (1.0 + ri);
    r0 *= r0;
    return r0 + (1.0 - r0) * pow(1.0 - cosine, 5.0);
}

// RandomInUnitSphere, lib/Random.glsl:56-67: LCG triples 2u - 1 until one lies inside
// the unit ball.  The components are exact in fp32 and their squares sum exactly in
// fp64, so the acceptance test is exact.  The loop is cut at 256 tries (a rejection
// rate of 1 - pi/6 makes that unreachable) so that every lane ends.
__device__ inline D3 lcg_unit_sphere(uint32_t& seed) {
    for (int k = 0; k < 256; k++) {
        const double x = 2.0 * double(lcg_float(seed)) - 1.0;
        const double y = 2.0 * double(lcg_float(seed)) - 1.0;
        const double z = 2.0 * double(lcg_float(seed)) - 1.0;
        if (x * x + y * y + z * z < 1.0) return d3(x, y, z);
    }
    return d3(0.0);
}

// evalIndirectCombinedBRDF, PBR.glsl:649-729 (+ sampleSpecularMicrofacet :587-616)
__device__ inline bool sample_indirect(double ux, double uy, D3 sn, D3 gn, D3 V, const MatBrdf& m, uint32_t type,
                                       D3& dir, D3& weight, double& volume_dis) {
    if (ddot(gn, V) < 0.0) return false;
    Quat q = rotation_to_z(sn);
    D3 Vl = rotate_point(q, V);
    D3 dl;
    double alpha = m.roughness * m.roughness, a2 = alpha * alpha;
    if (type == kDiffuse) {
        double a = sqrt(ux), sp, cp;
        sincos(kTwoPi * uy, &sp, &cp);
        dl = d3(a * cp, a * sp, sqrt(1.0 - ux));
        BrdfTerms d = brdf_terms(d3(0.0, 0.0, 1.0), dl, Vl);
        D3 F = fresnel(m.F0, shadowed_f90(m.F90), d.vdh);
        double fb = frostbite(m.roughness, d.ndl, d.ndv, d.ldh);
        weight = (d3(1.0) - F * m.specular_factor) * m.c_diff * fb;
        weight = weight * (1.0 - m.transmission);
    } else if (type == kSpecular) {
        D3 Hl = alpha == 0.0 ? d3(0.0, 0.0, 1.0) : sample_vndf(Vl, alpha, ux, uy);
        D3 nv = -Vl;
        D3 Ll = nv - Hl * (2.0 * ddot(Hl, nv));
        double hdl = fmax(0.00001, fmin(1.0, ddot(Hl, Ll)));
        double ndl = fmax(0.00001, fmin(1.0, Ll.z));
        double ndv = fmax(0.00001, fmin(1.0, Vl.z));
        D3 F = fresnel(m.F0, shadowed_f90(m.F90), hdl);
        weight = F * g2_over_g1(a2, ndl, ndv);
        dl = Ll;
        weight = weight * m.specular_factor;
    } else {
        if (m.volume) {
            double eta = m.front_face ? 1.0 / m.ior : m.ior;
            D3 rf = refract3(-Vl, d3(0.0, 0.0, 1.0), eta);
            if (rf.x == 0.0 && rf.y == 0.0 && rf.z == 0.0) {
                weight = d3(0.0);
                return false;
            }
            dl = rf;
        } else {
            dl = -Vl;
        }
        D3 w = m.c_diff * m.transmission;
        weight = d3(fmax(0.0, w.x), fmax(0.0, w.y), fmax(0.0, w.z));
        if (!m.front_face && m.volume) {
            double dis = volume_dis;
            volume_dis = -1.0;
            D3 att = d3(exp(log(m.attenuation_color.x) / m.attenuation_distance * dis),
                        exp(log(m.attenuation_color.y) / m.attenuation_distance * dis),
                        exp(log(m.attenuation_color.z) / m.attenuation_distance * dis));
            weight = weight * d3(fmin(att.x, 1.0), fmin(att.y, 1.0), fmin(att.z, 1.0));
        }
    }
    if (luminance(weight) == 0.0) return false;
    dir = dnormalize(rotate_point(Quat{-q.x, -q.y, -q.z, q.w}, dl));
    return true;
}

__device__ inline double convert_metallic(D3 diffuse, D3 specular, double max_spec) {  // lib/Material.glsl:76-89
    const double cmin = 0.04;
    double pd = sqrt(0.299 * diffuse.x * diffuse.x + 0.587 * diffuse.y * diffuse.y + 0.114 * diffuse.z * diffuse.z);
    double ps = sqrt(0.299 * specular.x * specular.x + 0.587 * specular.y * specular.y + 0.114 * specular.z * specular.z);
    if (ps < cmin) return 0.0;
    double b = pd * (1.0 - max_spec) / (1.0 - cmin) + ps - 2.0 * cmin;
    double c = cmin - ps;
    double D = fmax(b * b - 4.0 * cmin * c, 0.0);
    return clampd((-b + sqrt(D)) / (2.0 * cmin), 0.0, 1.0);
}

__device__ inline D3 light_intensity_at(const rt_light& l, double dist) {  // lib/PunctualLight.glsl:17-30
    D3 c = ld3(l.color) * double(l.intensity);
    if (l.kind == RT_LIGHT_POINT) {
        const double r2 = 0.25;
        double d2 = dist * dist;
        return c * (2.0 / (d2 + r2 + dist * sqrt(d2 + r2)));
    }
    return c;
}

// offset_ray, lib/RayTracingCommons.glsl:106-122: integer-ulp steps of the fp32 point
__device__ inline D3 offset_ray(D3 pd, D3 nd) {
    const float origin = 1.f / 32.f, fs = 1.f / 65536.f, is = 256.f;
    float pp[3] = {float(pd.x), float(pd.y), float(pd.z)}, nn[3] = {float(nd.x), float(nd.y), float(nd.z)}, out[3];
    for (int k = 0; k < 3; k++) {
        int of = int(nn[k] * is);
        int bits = __float_as_int(pp[k]) + ((pp[k] < 0.f) ? -of : of);
        out[k] = fabsf(pp[k]) < origin ? pp[k] + fs * nn[k] : __int_as_float(bits);
    }
    return d3(out[0], out[1], out[2]);
}

// ----------------------------------------------------------------- any-hit
// RayTracing.rahit / RayTracing.shadow.rahit:44-104.  True = ignore the candidate.
__device__ inline bool alpha_test_ignores(const SceneView& sv, uint32_t custom_index, uint32_t prim, float fu, float fv,
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
    const float* a = reinterpret_cast<const float*>(sv.vertices + i0);
    const float* b = reinterpret_cast<const float*>(sv.vertices + i1);
    const float* c = reinterpret_cast<const float*>(sv.vertices + i2);
    double bu = fu, bv = fv, b0 = 1.0 - bu - bv;
    D4 ua = ld4(a + 24), ub = ld4(b + 24), uc = ld4(c + 24);
    D4 uv = D4{ua.x * b0 + ub.x * bu + uc.x * bv, ua.y * b0 + ub.y * bu + uc.y * bv, ua.z * b0 + ub.z * bu + uc.z * bv,
               ua.w * b0 + ub.w * bu + uc.w * bv};
    D4 ca = ld4(a + 12), cb = ld4(b + 12), cc = ld4(c + 12);
    D4 vcol = D4{ca.x * b0 + cb.x * bu + cc.x * bv, ca.y * b0 + cb.y * bu + cc.y * bv, ca.z * b0 + cb.z * bu + cc.z * bv,
                 ca.w * b0 + cb.w * bu + cc.w * bv};
    D4 col = mul4(vcol, ld4(mat.base_color));
    if (mat.base_color_texture.index >= 0) col = mul4(col, tex_uv(sv, mat.base_color_texture, uv));
    double opacity = col.w;
    if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) {
        D4 df = ld4(mat.sg_diffuse_factor);
        if (mat.sg_diffuse_texture.index >= 0) df = mul4(df, tex_uv(sv, mat.sg_diffuse_texture, uv));
        opacity = vcol.w * df.w;
    }
    if (mode == RT_ALPHA_MASK) return opacity < double(mat.alpha_cutoff);
    return opacity <= double(rng_next(rng));
}

}  // namespace rt
