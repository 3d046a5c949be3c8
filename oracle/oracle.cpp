// oracle.cpp — CPU restatement of the reference ray-tracing path (FP64 and FP32).
//
// TEST INFRASTRUCTURE ONLY: the CPU reference the HIP library is tested against.  Loaded by tests/,
// __graft_entry__.smoke() and bench.py's CPU baseline; never by the product path.
//
// Every function cites the reference file:line it restates.  Paths are relative
// to crates/examples/gltf_viewer/shaders/ unless stated otherwise.
//
// Deviation shared with the HIP library (the reference is non-deterministic here):
// stream B of the RNG and the LCG seed use clockARB() in RayTracing.rgen:33,40.
// Both are seeded from tea16(ubo.RandomSeed, ubo.frame_count) instead.

#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>
#include <unordered_map>
#include <omp.h>

namespace {

// ------------------------------------------------------------------ vectors
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R a) : x(a), y(a), z(a) {}
    R operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    R& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class R> V3<R> operator-(V3<R> a) { return {-a.x, -a.y, -a.z}; }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class R> V3<R> operator*(V3<R> a, R s) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator*(R s, V3<R> a) { return {a.x * s, a.y * s, a.z * s}; }
template <class R> V3<R> operator/(V3<R> a, R s) { return {a.x / s, a.y / s, a.z / s}; }
template <class R> V3<R> operator/(V3<R> a, V3<R> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class R> R length(V3<R> a) { return std::sqrt(dot(a, a)); }
template <class R> V3<R> normalize(V3<R> a) { return a / length(a); }
template <class R> V3<R> vmin(V3<R> a, V3<R> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <class R> V3<R> vmax(V3<R> a, V3<R> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
// GLSL mix(x, y, a) = x * (1 - a) + y * a
template <class R> R mixs(R x, R y, R a) { return x * (R(1) - a) + y * a; }
template <class R> V3<R> mixv(V3<R> x, V3<R> y, R a) { return x * (R(1) - a) + y * a; }
template <class R> R clampr(R x, R lo, R hi) { return std::min(std::max(x, lo), hi); }
template <class R> R saturate(R x) { return clampr(x, R(0), R(1)); }

template <class R> struct V4 {
    R x, y, z, w;
    V4() : x(0), y(0), z(0), w(0) {}
    V4(R a, R b, R c, R d) : x(a), y(b), z(c), w(d) {}
    V3<R> xyz() const { return {x, y, z}; }
};
template <class R> V4<R> operator+(V4<R> a, V4<R> b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
template <class R> V4<R> operator*(V4<R> a, V4<R> b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
template <class R> V4<R> operator*(V4<R> a, R s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
template <class R> V4<R> normalize4(V4<R> a) {
    R l = std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z + a.w * a.w);
    return {a.x / l, a.y / l, a.z / l, a.w / l};
}
template <class R> V3<R> load3(const float* p) { return {R(p[0]), R(p[1]), R(p[2])}; }
template <class R> V4<R> load4(const float* p) { return {R(p[0]), R(p[1]), R(p[2]), R(p[3])}; }

// ---------------------------------------------------------------------- RNG
// pcg4d: lib/Random.glsl:81-98 (Jarzynski & Olano, JCGT 2020).
struct U4 { uint32_t x, y, z, w; };
inline U4 pcg4d(U4 v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    v.x ^= v.x >> 16u;
    v.y ^= v.y >> 16u;
    v.z ^= v.z >> 16u;
    v.w ^= v.w >> 16u;
    v.x += v.y * v.w;
    v.y += v.z * v.x;
    v.z += v.x * v.y;
    v.w += v.y * v.z;
    return v;
}
// uintToFloat: lib/Random.glsl:70-72 (exact in float and double).
inline float uint_to_float(uint32_t x) {
    uint32_t b = 0x3f800000u | (x >> 9);
    float f;
    std::memcpy(&f, &b, 4);
    return f - 1.0f;
}
// rand: lib/Random.glsl:102-105 — pre-increments the sample counter.
inline float rand_next(U4& s) {
    s.w += 1u;
    return uint_to_float(pcg4d(s).x);
}
// InitRandomSeed (TEA, 16 rounds): lib/Random.glsl:12-25.
inline uint32_t tea16(uint32_t val0, uint32_t val1) {
    uint32_t v0 = val0, v1 = val1, s0 = 0;
    for (int n = 0; n < 16; n++) {
        s0 += 0x9e3779b9u;
        v0 += ((v1 << 4) + 0xa341316cu) ^ (v1 + s0) ^ ((v1 >> 5) + 0xc8013ea4u);
        v1 += ((v0 << 4) + 0xad90777du) ^ (v0 + s0) ^ ((v0 >> 5) + 0x7e95761eu);
    }
    return v0;
}
// RandomInt / RandomFloat (LCG, Numerical Recipes constants): lib/Random.glsl:27-42.
inline float lcg_float(uint32_t& seed) {
    seed = 1664525u * seed + 1013904223u;
    return float(seed & 0x00FFFFFFu) / float(0x01000000);
}
// RandomInUnitDisk: lib/Random.glsl:44-54 (x drawn before y).
inline void lcg_unit_disk(uint32_t& seed, float& px, float& py) {
    for (;;) {
        float a = lcg_float(seed);
        float b = lcg_float(seed);
        px = 2.0f * a - 1.0f;
        py = 2.0f * b - 1.0f;
        if (px * px + py * py < 1.0f) return;
    }
}
// RandomInUnitSphere, lib/Random.glsl:56-67.  The components 2u - 1 are exact in
// fp32 and their squares sum exactly in fp64, so the acceptance test is exact.  Cut
// at 256 tries (with a rejection rate of 1 - pi/6 never reached).
inline void lcg_unit_sphere(uint32_t& seed, double p[3]) {
    for (int k = 0; k < 256; k++) {
        double x = 2.0 * double(lcg_float(seed)) - 1.0;
        double y = 2.0 * double(lcg_float(seed)) - 1.0;
        double z = 2.0 * double(lcg_float(seed)) - 1.0;
        if (x * x + y * y + z * z < 1.0) {
            p[0] = x; p[1] = y; p[2] = z;
            return;
        }
    }
    p[0] = p[1] = p[2] = 0.0;
}
inline uint32_t stream_b_seed(uint32_t random_seed, uint32_t frame) { return tea16(random_seed, frame); }

// ------------------------------------------------------------ BRDF (PBR.glsl)
template <class R> struct Consts {
    static constexpr R PI = R(3.141592653589);  // PBR.glsl:57 (written as a float literal)
    static constexpr R TWO_PI = R(2) * PI;
    static constexpr R ONE_OVER_PI = R(1) / PI;
};

template <class R> R luminance(V3<R> c) {  // PBR.glsl:221-224
    return c.x * R(0.2126) + c.y * R(0.7152) + c.z * R(0.0722);
}

template <class R> struct MatBrdf {  // PBR.glsl:173-193
    V3<R> base_color;
    R metallic, roughness, ior, transmission;
    bool use_spec;
    R specular_factor;
    V3<R> specular_color_factor;
    V3<R> dielectric_f0, dielectric_f90, F0, F90, c_diff;
    bool front_face;
    V3<R> attenuation_color;
    R attenuation_distance;
    bool volume;
    R t_diff;
};

template <class R> void mat_build(MatBrdf<R>& m) {  // PBR.glsl:195-206
    R factor = (m.ior - R(1)) / (m.ior + R(1));
    V3<R> d = m.specular_color_factor * (factor * factor);
    m.dielectric_f0 = vmin(d, V3<R>(R(1))) * m.specular_factor;
    m.dielectric_f90 = m.specular_color_factor;
    m.F0 = mixv(m.dielectric_f0, m.base_color, m.metallic);
    m.F90 = mixv(m.dielectric_f90, V3<R>(R(1)), m.metallic);
    m.c_diff = mixv(m.base_color, V3<R>(R(0)), m.metallic);
}

template <class R> V3<R> fresnel_schlick(V3<R> f0, R f90, R n_dot_s) {  // PBR.glsl:228-231
    R p = std::pow(R(1) - n_dot_s, R(5));
    return f0 + (V3<R>(f90) - f0) * p;
}
template <class R> R shadowed_f90(V3<R> F90) { return std::min(R(1), luminance(F90)); }  // PBR.glsl:312-322

template <class R> R smith_g_a(R alpha, R n_dot_s) {  // PBR.glsl:248-250
    return n_dot_s / (std::max(R(0.00001), alpha) * std::sqrt(R(1) - std::min(R(0.99999), n_dot_s * n_dot_s)));
}
template <class R> R smith_lambda_ggx(R a) {  // PBR.glsl:253-255
    return (R(-1) + std::sqrt(R(1) + (R(1) / (a * a)))) * R(0.5);
}
template <class R> R smith_g2_hc(R alpha, R n_dot_l, R n_dot_v) {  // PBR.glsl:258-262
    R aL = smith_g_a(alpha, n_dot_l);
    R aV = smith_g_a(alpha, n_dot_v);
    return R(1) / (R(1) + smith_lambda_ggx(aL) + smith_lambda_ggx(aV));
}
template <class R> R ggx_d(R alpha_sq, R n_dot_h) {  // PBR.glsl:286-289
    R b = ((alpha_sq - R(1)) * n_dot_h * n_dot_h + R(1));
    return alpha_sq / (Consts<R>::PI * b * b);
}
template <class R> R smith_g1_ggx(R alpha, R n_dot_s, R alpha_sq, R n_dot_s_sq) {  // PBR.glsl:446-448
    (void)alpha; (void)n_dot_s;
    return R(2) / (std::sqrt(((alpha_sq * (R(1) - n_dot_s_sq)) + n_dot_s_sq) / n_dot_s_sq) + R(1));
}
template <class R> R smith_g2_over_g1(R alpha, R alpha_sq, R n_dot_l, R n_dot_v) {  // PBR.glsl:470-474
    R g1v = smith_g1_ggx(alpha, n_dot_v, alpha_sq, n_dot_v * n_dot_v);
    R g1l = smith_g1_ggx(alpha, n_dot_l, alpha_sq, n_dot_l * n_dot_l);
    return g1l / (g1v + g1l - g1v * g1l);
}

template <class R> struct BrdfData {  // PBR.glsl:13-45
    V3<R> F0, c_diff, F90;
    R roughness, alpha, alpha_sq;
    V3<R> F;
    R n_dot_l, n_dot_v, l_dot_h, n_dot_h, v_dot_h;
    bool v_back, l_back;
};
template <class R> BrdfData<R> prepare_brdf(V3<R> N, V3<R> L, V3<R> V, const MatBrdf<R>& m) {  // PBR.glsl:527-564
    BrdfData<R> d;
    V3<R> H = normalize(L + V);
    R ndl = dot(N, L), ndv = dot(N, V);
    d.v_back = ndv <= R(0);
    d.l_back = ndl <= R(0);
    d.n_dot_l = std::min(std::max(R(0.00001), ndl), R(1));
    d.n_dot_v = std::min(std::max(R(0.00001), ndv), R(1));
    d.l_dot_h = saturate(dot(L, H));
    d.n_dot_h = saturate(dot(N, H));
    d.v_dot_h = saturate(dot(V, H));
    d.F0 = m.F0;
    d.F90 = m.F90;
    d.c_diff = m.c_diff;
    d.roughness = m.roughness;
    d.alpha = m.roughness * m.roughness;
    d.alpha_sq = d.alpha * d.alpha;
    d.F = fresnel_schlick(d.F0, shadowed_f90(m.F90), d.v_dot_h);
    return d;
}
template <class R> R frostbite(R roughness, R n_dot_l, R n_dot_v, R l_dot_h) {  // PBR.glsl:486-496
    R energy_bias = R(0.5) * roughness;
    R energy_factor = mixs(R(1), R(1) / R(1.51), roughness);
    R fd90m1 = energy_bias + R(2) * l_dot_h * l_dot_h * roughness - R(1);
    R fdl = R(1) + (fd90m1 * std::pow(R(1) - n_dot_l, R(5)));
    R fdv = R(1) + (fd90m1 * std::pow(R(1) - n_dot_v, R(5)));
    return fdl * fdv * energy_factor;
}
template <class R> V3<R> eval_combined_brdf(V3<R> N, V3<R> L, V3<R> V, const MatBrdf<R>& m) {  // PBR.glsl:731-745
    BrdfData<R> d = prepare_brdf(N, L, V, m);
    if (d.v_back || d.l_back) return V3<R>(R(0));
    // evalMicrofacet, PBR.glsl:292-303
    R D = ggx_d(std::max(R(0.00001), d.alpha_sq), d.n_dot_h);
    R G2 = smith_g2_hc(d.alpha, d.n_dot_l, d.n_dot_v);
    V3<R> spec = ((d.F * (G2 * D)) / (R(4) * d.n_dot_l * d.n_dot_v)) * d.n_dot_l;
    // evalFrostbiteDisneyDiffuse, PBR.glsl:516-518
    V3<R> diff = d.c_diff * (frostbite(d.roughness, d.n_dot_l, d.n_dot_v, d.l_dot_h) * Consts<R>::ONE_OVER_PI * d.n_dot_l);
    return diff + spec;
}
template <class R> V3<R> brdf_probability(const MatBrdf<R>& m, V3<R> V, V3<R> N) {  // PBR.glsl:324-358
    R spec_f0 = luminance(m.F0);
    R diff_refl = luminance(m.c_diff);
    R fres = saturate(luminance(fresnel_schlick(V3<R>(spec_f0), shadowed_f90(m.F90), std::max(R(0), dot(V, N)))));
    R specular = fres * m.specular_factor;
    R penetration = diff_refl * (R(1) - m.specular_factor * fres);
    R diffuse = penetration * (R(1) - m.transmission);
    R transmission = penetration * m.transmission;
    R sum = std::max(R(0.0001), specular + diffuse + transmission);
    R p = specular / sum;
    p = clampr(p, R(0.001), R(0.9));
    R d = (R(1) - p) * (R(1) - m.transmission);
    R t = (R(1) - p) * m.transmission;
    sum = p + d + t;
    return {p / sum, d / sum, t / sum};
}
template <class R> V4<R> rotation_to_z(V3<R> v) {  // PBR.glsl:362-368
    if (v.z < R(-0.99999)) return {R(1), R(0), R(0), R(0)};
    return normalize4(V4<R>(v.y, -v.x, R(0), R(1) + v.z));
}
template <class R> V3<R> rotate_point(V4<R> q, V3<R> v) {  // PBR.glsl:373-376
    V3<R> qa(q.x, q.y, q.z);
    return qa * (R(2) * dot(qa, v)) + v * (q.w * q.w - dot(qa, qa)) + cross(qa, v) * (R(2) * q.w);
}
template <class R> V4<R> invert_rotation(V4<R> q) { return {-q.x, -q.y, -q.z, q.w}; }  // PBR.glsl:580-583

template <class R> V3<R> sample_hemisphere(R ux, R uy) {  // PBR.glsl:381-394
    R a = std::sqrt(ux);
    R b = Consts<R>::TWO_PI * uy;
    return {a * std::cos(b), a * std::sin(b), std::sqrt(R(1) - ux)};
}
template <class R> V3<R> sample_ggx_vndf(V3<R> Ve, R ax, R ay, R ux, R uy) {  // PBR.glsl:413-436 (Heitz 2018)
    V3<R> Vh = normalize(V3<R>(ax * Ve.x, ay * Ve.y, Ve.z));
    R lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3<R> T1 = lensq > R(0) ? V3<R>(-Vh.y, Vh.x, R(0)) * (R(1) / std::sqrt(lensq)) : V3<R>(R(1), R(0), R(0));
    V3<R> T2 = cross(Vh, T1);
    R r = std::sqrt(ux);
    R phi = Consts<R>::TWO_PI * uy;
    R t1 = r * std::cos(phi);
    R t2 = r * std::sin(phi);
    R s = R(0.5) * (R(1) + Vh.z);
    t2 = mixs(std::sqrt(R(1) - t1 * t1), t2, s);
    V3<R> Nh = T1 * t1 + T2 * t2 + Vh * std::sqrt(std::max(R(0), R(1) - t1 * t1 - t2 * t2));
    return normalize(V3<R>(ax * Nh.x, ay * Nh.y, std::max(R(0), Nh.z)));
}
template <class R> V3<R> reflect(V3<R> I, V3<R> N) { return I - N * (R(2) * dot(N, I)); }
template <class R> R schlick(R cosine, R ri) {  // PBR.glsl:48-53
    R r0 = (R(1) - ri) / (R(1) + ri);
    r0 = r0 * r0;
    return r0 + (R(1) - r0) * std::pow(R(1) - cosine, R(5));
}
template <class R> V3<R> refract(V3<R> I, V3<R> N, R eta) {  // GLSL refract
    R k = R(1) - eta * eta * (R(1) - dot(N, I) * dot(N, I));
    if (k < R(0)) return V3<R>(R(0));
    return I * eta - N * (eta * dot(N, I) + std::sqrt(k));
}

enum { DIFFUSE_TYPE = 1, SPECULAR_TYPE = 2, TRANSMISSION_TYPE = 3 };

// evalIndirectCombinedBRDF: PBR.glsl:649-729, with sampleSpecularMicrofacet :587-616.
template <class R> bool sample_indirect(R ux, R uy, V3<R> shading_n, V3<R> geo_n, V3<R> V, const MatBrdf<R>& m,
                                        uint32_t type, V3<R>& ray_dir, V3<R>& weight, R& volume_dis) {
    if (dot(geo_n, V) < R(0)) return false;
    V4<R> q = rotation_to_z(shading_n);
    V3<R> Vl = rotate_point(q, V);
    const V3<R> Nl(R(0), R(0), R(1));
    V3<R> dl(R(0));
    if (type == DIFFUSE_TYPE) {
        dl = sample_hemisphere(ux, uy);
        BrdfData<R> d = prepare_brdf(Nl, dl, Vl, m);
        R fb = frostbite(d.roughness, d.n_dot_l, d.n_dot_v, d.l_dot_h);
        weight = (V3<R>(R(1)) - d.F * m.specular_factor) * d.c_diff * fb;
        weight = weight * (R(1) - m.transmission);
    } else if (type == SPECULAR_TYPE) {
        BrdfData<R> d = prepare_brdf(Nl, V3<R>(R(0), R(0), R(1)), Vl, m);
        V3<R> Hl;
        if (d.alpha == R(0)) Hl = V3<R>(R(0), R(0), R(1));
        else Hl = sample_ggx_vndf(Vl, d.alpha, d.alpha, ux, uy);
        V3<R> Ll = reflect(-Vl, Hl);
        R h_dot_l = std::max(R(0.00001), std::min(R(1), dot(Hl, Ll)));
        R n_dot_l = std::max(R(0.00001), std::min(R(1), Ll.z));
        R n_dot_v = std::max(R(0.00001), std::min(R(1), Vl.z));
        V3<R> F = fresnel_schlick(d.F0, shadowed_f90(d.F90), h_dot_l);
        weight = F * smith_g2_over_g1(d.alpha, d.alpha_sq, n_dot_l, n_dot_v);
        dl = Ll;
        weight = weight * m.specular_factor;
    } else {
        if (m.volume) {
            R eta = m.front_face ? R(1) / m.ior : m.ior;
            V3<R> rf = refract(-Vl, Nl, eta);
            if (rf.x == R(0) && rf.y == R(0) && rf.z == R(0)) {
                weight = V3<R>(R(0));
                return false;
            }
            dl = rf;
        } else {
            dl = -Vl;
        }
        weight = vmax(V3<R>(R(0)), m.c_diff * m.transmission);
        if (!m.front_face && m.volume) {
            R dis = volume_dis;
            volume_dis = R(-1);
            V3<R> sigma(std::log(m.attenuation_color.x) / m.attenuation_distance,
                        std::log(m.attenuation_color.y) / m.attenuation_distance,
                        std::log(m.attenuation_color.z) / m.attenuation_distance);
            V3<R> att(std::exp(sigma.x * dis), std::exp(sigma.y * dis), std::exp(sigma.z * dis));
            weight = weight * vmin(att, V3<R>(R(1)));
        }
    }
    if (luminance(weight) == R(0)) return false;
    ray_dir = normalize(rotate_point(invert_rotation(q), dl));
    return true;
}

// convertMetallic: lib/Material.glsl:76-89.
template <class R> R convert_metallic(V3<R> diffuse, V3<R> specular, R max_spec) {
    const R c_min = R(0.04);
    R pd = std::sqrt(R(0.299) * diffuse.x * diffuse.x + R(0.587) * diffuse.y * diffuse.y + R(0.114) * diffuse.z * diffuse.z);
    R ps = std::sqrt(R(0.299) * specular.x * specular.x + R(0.587) * specular.y * specular.y + R(0.114) * specular.z * specular.z);
    if (ps < c_min) return R(0);
    R a = c_min;
    R b = pd * (R(1) - max_spec) / (R(1) - c_min) + ps - R(2) * c_min;
    R c = c_min - ps;
    R D = std::max(b * b - R(4) * a * c, R(0));
    return clampr((-b + std::sqrt(D)) / (R(2) * a), R(0), R(1));
}

// getLightIntensityAtPoint: lib/PunctualLight.glsl:17-30 (Yuksel attenuation, r = 0.5).
template <class R> V3<R> light_intensity_at(const rt_light& l, R distance) {
    V3<R> color = V3<R>(R(l.color[0]), R(l.color[1]), R(l.color[2])) * R(l.intensity);
    if (l.kind == RT_LIGHT_POINT) {
        const R radius = R(0.5);
        R r2 = radius * radius, d2 = distance * distance;
        R att = R(2) / (d2 + r2 + distance * std::sqrt(d2 + r2));
        return color * att;
    }
    return color;
}

// offset_ray: lib/RayTracingCommons.glsl:106-122 (integer-ULP offset, float bits).
inline void offset_ray_f(const float p[3], const float n[3], float out[3]) {
    const float origin = 1.0f / 32.0f, float_scale = 1.0f / 65536.0f, int_scale = 256.0f;
    for (int k = 0; k < 3; k++) {
        int of = int(n[k] * int_scale);
        int32_t bits;
        std::memcpy(&bits, &p[k], 4);
        bits += (p[k] < 0) ? -of : of;
        float pi;
        std::memcpy(&pi, &bits, 4);
        out[k] = std::fabs(p[k]) < origin ? p[k] + float_scale * n[k] : pi;
    }
}

// ------------------------------------------------------------- textures
// texture(sampler2D, uv) at LOD 0 (RT stages have no derivatives): magFilter,
// address modes on integer texel coordinates, sRGB decoded per texel before
// filtering (Vulkan spec, "Texel Filtering" / "Wrapping Operation").
inline int wrap_coord(int i, int size, uint32_t mode) {
    if (mode == RT_WRAP_CLAMP_TO_EDGE) return std::min(std::max(i, 0), size - 1);
    if (mode == RT_WRAP_MIRRORED_REPEAT) {
        int t = i % (2 * size);
        if (t < 0) t += 2 * size;
        return t < size ? t : 2 * size - 1 - t;
    }
    int t = i % size;
    return t < 0 ? t + size : t;
}
inline double srgb_decode(double c) { return c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4); }
inline void texel_fetch(const rt_texture& t, int x, int y, double out[4]) {
    size_t idx = size_t(y) * t.width + size_t(x);
    if (t.format == RT_FMT_RGBA32F) {
        const float* p = (const float*)t.data + idx * 4;
        for (int k = 0; k < 4; k++) out[k] = p[k];
        return;
    }
    const uint8_t* p = (const uint8_t*)t.data + idx * 4;
    for (int k = 0; k < 4; k++) out[k] = p[k] / 255.0;
    if (t.format == RT_FMT_RGBA8_SRGB)
        for (int k = 0; k < 3; k++) out[k] = srgb_decode(out[k]);
}
inline void texture_sample(const rt_texture& t, double u, double v, double out[4]) {
    int W = int(t.width), H = int(t.height);
    if (t.mag_filter == RT_FILTER_NEAREST) {
        int x = wrap_coord(int(std::floor(u * W)), W, t.wrap_s);
        int y = wrap_coord(int(std::floor(v * H)), H, t.wrap_t);
        texel_fetch(t, x, y, out);
        return;
    }
    double fu = u * W - 0.5, fv = v * H - 0.5;
    double iu = std::floor(fu), iv = std::floor(fv);
    double a = fu - iu, b = fv - iv;
    int x0 = wrap_coord(int(iu), W, t.wrap_s), x1 = wrap_coord(int(iu) + 1, W, t.wrap_s);
    int y0 = wrap_coord(int(iv), H, t.wrap_t), y1 = wrap_coord(int(iv) + 1, H, t.wrap_t);
    double t00[4], t10[4], t01[4], t11[4];
    texel_fetch(t, x0, y0, t00);
    texel_fetch(t, x1, y0, t10);
    texel_fetch(t, x0, y1, t01);
    texel_fetch(t, x1, y1, t11);
    // bilinear as two nested lerps x + w (y - x): a constant neighbourhood filters to
    // exactly that constant (as fixed-point texture hardware gives it)
    for (int k = 0; k < 4; k++) {
        double top = t00[k] + a * (t10[k] - t00[k]), bot = t01[k] + a * (t11[k] - t01[k]);
        out[k] = top + b * (bot - top);
    }
}

// ------------------------------------------------------------ cube map
// texture(samplerCube, dir) at LOD 0, RayTracing.rmiss:27-33 with the skybox of
// cubumap.rs / globals.rs:201-240 (six square layers +x, -x, +y, -y, +z, -z).
// Vulkan's rules: the face is the largest-magnitude axis (equal magnitudes pick z,
// then y); (s, t) = ((sc / |ma| + 1) / 2, (tc / |ma| + 1) / 2) with the face's
// basis below; a bilinear texel off the face is the neighbouring face's texel
// under the same texel centre continued over the edge (seamless cube edges); a
// texel off two edges is the mean of the three texels meeting at that corner.
struct CubeBasis {
    double ma[3], sc[3], tc[3];  // unit vectors: face normal, direction of growing s, of growing t
};
inline CubeBasis cube_basis(int face) {
    static const CubeBasis kFaces[6] = {
        {{1, 0, 0}, {0, 0, -1}, {0, -1, 0}},   // +x: sc = -z, tc = -y
        {{-1, 0, 0}, {0, 0, 1}, {0, -1, 0}},   // -x: sc = +z, tc = -y
        {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}},     // +y: sc = +x, tc = +z
        {{0, -1, 0}, {1, 0, 0}, {0, 0, -1}},   // -y: sc = +x, tc = -z
        {{0, 0, 1}, {1, 0, 0}, {0, -1, 0}},    // +z: sc = +x, tc = -y
        {{0, 0, -1}, {-1, 0, 0}, {0, -1, 0}},  // -z: sc = -x, tc = -y
    };
    return kFaces[face];
}
inline void cube_project(const double d[3], int& face, double& s, double& t) {
    double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    if (az >= ax && az >= ay) face = d[2] < 0 ? 5 : 4;
    else if (ay >= ax) face = d[1] < 0 ? 3 : 2;
    else face = d[0] < 0 ? 1 : 0;
    CubeBasis b = cube_basis(face);
    double ma = d[0] * b.ma[0] + d[1] * b.ma[1] + d[2] * b.ma[2];  // = |major component|
    s = 0.5 * ((d[0] * b.sc[0] + d[1] * b.sc[1] + d[2] * b.sc[2]) / ma + 1.0);
    t = 0.5 * ((d[0] * b.tc[0] + d[1] * b.tc[1] + d[2] * b.tc[2]) / ma + 1.0);
}
inline void cube_texel_fetch(const rt_texture& c, int face, int i, int j, double out[4]) {
    const int W = int(c.width);
    const bool oi = i < 0 || i >= W, oj = j < 0 || j >= W;
    if (oi && oj) {
        int ic = std::min(std::max(i, 0), W - 1), jc = std::min(std::max(j, 0), W - 1);
        double p[4], q[4], r[4];
        cube_texel_fetch(c, face, ic, jc, p);
        cube_texel_fetch(c, face, i, jc, q);
        cube_texel_fetch(c, face, ic, j, r);
        for (int k = 0; k < 4; k++) out[k] = (p[k] + q[k] + r[k]) / 3.0;
        return;
    }
    if (oi || oj) {
        // the texel centre in face units (the face spans -1..1), past the edge
        CubeBasis b = cube_basis(face);
        double u = (2.0 * i + 1.0) / W - 1.0, v = (2.0 * j + 1.0) / W - 1.0, p[3];
        for (int k = 0; k < 3; k++) p[k] = b.ma[k] + u * b.sc[k] + v * b.tc[k];
        int g;
        double s, t;
        cube_project(p, g, s, t);
        cube_texel_fetch(c, g, std::min(int(std::floor(s * W)), W - 1), std::min(int(std::floor(t * W)), W - 1), out);
        return;
    }
    rt_texture f = c;
    f.height = c.width;
    size_t texel_bytes = c.format == RT_FMT_RGBA32F ? 16 : 4;
    f.data = (const uint8_t*)c.data + size_t(face) * W * W * texel_bytes;
    texel_fetch(f, i, j, out);
}
inline void cube_sample(const rt_texture& c, const double d[3], double out[4]) {
    int face;
    double s, t;
    cube_project(d, face, s, t);
    const int W = int(c.width);
    if (c.mag_filter == RT_FILTER_NEAREST) {
        cube_texel_fetch(c, face, int(std::floor(s * W)), int(std::floor(t * W)), out);
        return;
    }
    double fu = s * W - 0.5, fv = t * W - 0.5;
    double iu = std::floor(fu), iv = std::floor(fv);
    double a = fu - iu, b = fv - iv;
    int i0 = int(iu), j0 = int(iv);
    double t00[4], t10[4], t01[4], t11[4];
    cube_texel_fetch(c, face, i0, j0, t00);
    cube_texel_fetch(c, face, i0 + 1, j0, t10);
    cube_texel_fetch(c, face, i0, j0 + 1, t01);
    cube_texel_fetch(c, face, i0 + 1, j0 + 1, t11);
    for (int k = 0; k < 4; k++) {
        double top = t00[k] + a * (t10[k] - t00[k]), bot = t01[k] + a * (t11[k] - t01[k]);
        out[k] = top + b * (bot - top);
    }
}

// ------------------------------------------------- ray/triangle (watertight)
// Woop, Benthin & Wald, "Watertight Ray/Triangle Intersection", JCGT 2(1) 2013.
// Returns the Vulkan barycentrics (u, v) = weights of vertices 1 and 2.
struct WoopRay {
    int kx, ky, kz;
    double Sx, Sy, Sz;
};
inline WoopRay woop_setup(const double d[3]) {
    WoopRay w;
    double ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    w.kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    w.kx = (w.kz + 1) % 3;
    w.ky = (w.kx + 1) % 3;
    if (d[w.kz] < 0) std::swap(w.kx, w.ky);
    w.Sx = d[w.kx] / d[w.kz];
    w.Sy = d[w.ky] / d[w.kz];
    w.Sz = 1.0 / d[w.kz];
    return w;
}
inline bool woop_intersect(const WoopRay& w, const double o[3], const double v0[3], const double v1[3],
                           const double v2[3], double tmin, double tmax, double& t, double& bu, double& bv) {
    double A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) { A[k] = v0[k] - o[k]; B[k] = v1[k] - o[k]; C[k] = v2[k] - o[k]; }
    double Ax = A[w.kx] - w.Sx * A[w.kz], Ay = A[w.ky] - w.Sy * A[w.kz];
    double Bx = B[w.kx] - w.Sx * B[w.kz], By = B[w.ky] - w.Sy * B[w.kz];
    double Cx = C[w.kx] - w.Sx * C[w.kz], Cy = C[w.ky] - w.Sy * C[w.kz];
    double U = Cx * By - Cy * Bx;
    double V = Ax * Cy - Ay * Cx;
    double W = Bx * Ay - By * Ax;
    if (U == 0.0 || V == 0.0 || W == 0.0) {  // recompute the edge functions in extended precision
        long double CxBy = (long double)Cx * By, CyBx = (long double)Cy * Bx;
        U = double(CxBy - CyBx);
        long double AxCy = (long double)Ax * Cy, AyCx = (long double)Ay * Cx;
        V = double(AxCy - AyCx);
        long double BxAy = (long double)Bx * Ay, ByAx = (long double)By * Ax;
        W = double(BxAy - ByAx);
    }
    if ((U < 0 || V < 0 || W < 0) && (U > 0 || V > 0 || W > 0)) return false;
    double det = U + V + W;
    if (det == 0.0) return false;
    double Az = w.Sz * A[w.kz], Bz = w.Sz * B[w.kz], Cz = w.Sz * C[w.kz];
    double T = U * Az + V * Bz + W * Cz;
    double inv = 1.0 / det;
    t = T * inv;
    if (!(t >= tmin && t <= tmax)) return false;
    bu = V * inv;
    bv = W * inv;
    return true;
}

// --------------------------------------------------------------- BVH (CPU)
// A plain binary BVH (median split), independent of the GPU builder.  Only its
// traversal result matters: the closest accepted hit, which it must share with
// brute force (tests/test_oracle_traversal.py).
struct Box {
    double lo[3], hi[3];
    void reset() { for (int k = 0; k < 3; k++) { lo[k] = 1e300; hi[k] = -1e300; } }
    void grow(const double p[3]) { for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], p[k]); hi[k] = std::max(hi[k], p[k]); } }
    void grow(const Box& b) { for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], b.lo[k]); hi[k] = std::max(hi[k], b.hi[k]); } }
};
struct BNode { Box box; int left, right, first, count; };

struct Bvh {
    std::vector<BNode> nodes;
    std::vector<int> order;
    template <class F> void build(int n, F&& prim_box) {
        nodes.clear();
        order.resize(n);
        std::vector<Box> boxes(n);
        std::vector<double> cent(3 * size_t(n));
        for (int i = 0; i < n; i++) {
            order[i] = i;
            boxes[i] = prim_box(i);
            for (int k = 0; k < 3; k++) cent[3 * size_t(i) + k] = 0.5 * (boxes[i].lo[k] + boxes[i].hi[k]);
        }
        if (n == 0) return;
        nodes.reserve(2 * size_t(n));
        rec(0, n, boxes, cent);
    }
    int rec(int first, int count, const std::vector<Box>& boxes, const std::vector<double>& cent) {
        int id = int(nodes.size());
        nodes.push_back(BNode());
        Box b, cb;
        b.reset();
        cb.reset();
        for (int i = first; i < first + count; i++) {
            b.grow(boxes[order[i]]);
            cb.grow(&cent[3 * size_t(order[i])]);
        }
        nodes[id].box = b;
        if (count <= 4) {
            nodes[id].left = nodes[id].right = -1;
            nodes[id].first = first;
            nodes[id].count = count;
            return id;
        }
        int axis = 0;
        double ext = -1;
        for (int k = 0; k < 3; k++) if (cb.hi[k] - cb.lo[k] > ext) { ext = cb.hi[k] - cb.lo[k]; axis = k; }
        int mid = first + count / 2;
        std::nth_element(order.begin() + first, order.begin() + mid, order.begin() + first + count,
                         [&](int a, int c) { return cent[3 * size_t(a) + axis] < cent[3 * size_t(c) + axis]; });
        int l = rec(first, mid - first, boxes, cent);
        int r = rec(mid, first + count - mid, boxes, cent);
        nodes[id].left = l;
        nodes[id].right = r;
        nodes[id].first = 0;
        nodes[id].count = 0;
        return id;
    }
};

// Slab test, widened so that it never rejects a box the exact ray touches.
inline bool ray_box(const Box& b, const double o[3], const double invd[3], double tmin, double tmax, double& tnear) {
    double t0 = tmin, t1 = tmax;
    for (int k = 0; k < 3; k++) {
        double a = (b.lo[k] - o[k]) * invd[k];
        double c = (b.hi[k] - o[k]) * invd[k];
        if (std::isnan(a) || std::isnan(c)) {  // ray parallel and on the slab plane
            if (o[k] < b.lo[k] || o[k] > b.hi[k]) return false;
            continue;
        }
        double lo = std::min(a, c), hi = std::max(a, c);
        hi *= 1.0 + 1e-12;
        hi += 1e-300;
        lo -= std::fabs(lo) * 1e-12;
        t0 = std::max(t0, lo);
        t1 = std::min(t1, hi);
        if (t0 > t1) return false;
    }
    tnear = t0;
    return true;
}

struct Instance {
    double o2w[12], w2o[12];  // row-major 3x4
    int blas;
    uint32_t id, custom_index, mask, flags;
};

struct HitRec {
    double t, u, v;
    int inst, prim;
    bool found;
};

struct Scene {
    orc_scene_desc d;
    // Detached evaluation (the adjoint's definition of the derivative): when not
    // empty, closest-hit evaluates the BRDF values (NEE and the sampled weights)
    // with these textures while lobe choices, directions and densities keep using
    // d.textures.
    std::vector<rt_texture> eval_tex;
    // Detached evaluation of light intensities: when not empty, the NEE term and the
    // miss shader's directional term use these intensities while the RIS choice and
    // weights keep using the scene's (d.plights / d.dlights).
    std::vector<float> eval_pint, eval_dint;
    // Texel watch list of orc_texel_footprints: (texture << 32 | texel) -> entry.
    std::unordered_map<uint64_t, uint32_t> watch;
    rt_texture sky{};  // copy of *d.skybox
    bool has_sky_tex = false;
    std::vector<std::vector<double>> tri_pos;  // per BLAS: 9 doubles per triangle
    std::vector<Bvh> blas_bvh;
    std::vector<Box> blas_box;
    std::vector<Instance> inst;
    Bvh tlas;
    std::vector<Box> inst_box;
};

void invert34(const double m[12], double r[12]) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double inv = 1.0 / det;
    double r3[9] = {A * inv, -(b * i - c * h) * inv, (b * f - c * e) * inv,
                    B * inv, (a * i - c * g) * inv, -(a * f - c * d) * inv,
                    C * inv, -(a * h - b * g) * inv, (a * e - b * d) * inv};
    for (int row = 0; row < 3; row++) {
        for (int col = 0; col < 3; col++) r[row * 4 + col] = r3[row * 3 + col];
        r[row * 4 + 3] = -(r3[row * 3 + 0] * m[3] + r3[row * 3 + 1] * m[7] + r3[row * 3 + 2] * m[11]);
    }
}
inline void xform_point(const double m[12], const double p[3], double out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[r * 4] * p[0] + m[r * 4 + 1] * p[1] + m[r * 4 + 2] * p[2] + m[r * 4 + 3];
}
inline void xform_dir(const double m[12], const double p[3], double out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[r * 4] * p[0] + m[r * 4 + 1] * p[1] + m[r * 4 + 2] * p[2];
}

Scene* build_scene(const orc_scene_desc* desc) {
    auto* s = new Scene();
    s->d = *desc;
    if (desc->skybox) {
        s->sky = *desc->skybox;
        s->has_sky_tex = true;
    }
    uint32_t nb = desc->num_blases;
    s->tri_pos.resize(nb);
    s->blas_bvh.resize(nb);
    s->blas_box.resize(nb);
    for (uint32_t bi = 0; bi < nb; bi++) {
        const orc_blas& b = desc->blases[bi];
        auto& tp = s->tri_pos[bi];
        tp.resize(9 * size_t(b.prim_count));
        for (uint32_t p = 0; p < b.prim_count; p++)
            for (int k = 0; k < 3; k++) {
                uint32_t vi = b.first_vertex + desc->indices[b.index_offset + 3 * p + k];
                for (int c = 0; c < 3; c++) tp[9 * size_t(p) + 3 * k + c] = desc->vertices[vi].pos[c];
            }
        s->blas_bvh[bi].build(int(b.prim_count), [&](int p) {
            Box bx;
            bx.reset();
            for (int k = 0; k < 3; k++) bx.grow(&tp[9 * size_t(p) + 3 * k]);
            return bx;
        });
        Box bb;
        bb.reset();
        for (uint32_t p = 0; p < b.prim_count; p++)
            for (int k = 0; k < 3; k++) bb.grow(&tp[9 * size_t(p) + 3 * k]);
        s->blas_box[bi] = bb;
    }
    s->inst.resize(desc->num_instances);
    s->inst_box.resize(desc->num_instances);
    for (uint32_t i = 0; i < desc->num_instances; i++) {
        const rt_instance& ri = desc->instances[i];
        Instance& in = s->inst[i];
        for (int k = 0; k < 12; k++) in.o2w[k] = ri.transform[k / 4][k % 4];
        invert34(in.o2w, in.w2o);
        in.blas = int(ri.acceleration_structure_reference);
        in.id = i;
        in.custom_index = ri.custom_index_and_mask & 0xFFFFFFu;
        in.mask = ri.custom_index_and_mask >> 24;
        in.flags = ri.sbt_offset_and_flags >> 24;
        Box wb;
        wb.reset();
        const Box& ob = s->blas_box[in.blas];
        if (desc->blases[in.blas].prim_count > 0)
            for (int c = 0; c < 8; c++) {
                double p[3] = {(c & 1) ? ob.hi[0] : ob.lo[0], (c & 2) ? ob.hi[1] : ob.lo[1], (c & 4) ? ob.hi[2] : ob.lo[2]};
                double w[3];
                xform_point(in.o2w, p, w);
                // pad for the rounding of the transform
                double pad[3];
                for (int k = 0; k < 3; k++) pad[k] = 1e-9 * (std::fabs(w[k]) + 1.0);
                double lo[3] = {w[0] - pad[0], w[1] - pad[1], w[2] - pad[2]};
                double hi[3] = {w[0] + pad[0], w[1] + pad[1], w[2] + pad[2]};
                wb.grow(lo);
                wb.grow(hi);
            }
        s->inst_box[i] = wb;
    }
    s->tlas.build(int(desc->num_instances), [&](int i) { return s->inst_box[i]; });
    return s;
}

// Any-hit decision for one candidate (RayTracing.rahit / RayTracing.shadow.rahit:44-104).
// Returns true when the candidate is ignored (transparent).
struct AnyHitCtx {
    const Scene* s;
    U4* rng;  // stream consumed by BLEND draws (payload rngState)
    uint64_t* calls;
};

template <class R> V4<R> fetch_tex(const Scene& s, int32_t index, V4<R> uv0and1, int32_t coord) {
    // getUV: lib/RayTracingCommons.glsl:65-72
    R u = 0, v = 0;
    uint32_t c = uint32_t(coord);
    if (c == 0) { u = uv0and1.x; v = uv0and1.y; }
    else if (c == 1) { u = uv0and1.z; v = uv0and1.w; }
    double out[4] = {1, 1, 1, 1};
    if (index >= 0 && uint32_t(index) < s.d.num_textures) texture_sample(s.d.textures[index], double(u), double(v), out);
    return {R(out[0]), R(out[1]), R(out[2]), R(out[3])};
}

// Texels a texture_sample at (u, v) reads (the same wrap and filter rules).
inline int texture_taps(const rt_texture& t, double u, double v, uint32_t idx[4]) {
    int W = int(t.width), H = int(t.height);
    if (t.mag_filter == RT_FILTER_NEAREST) {
        int x = wrap_coord(int(std::floor(u * W)), W, t.wrap_s);
        int y = wrap_coord(int(std::floor(v * H)), H, t.wrap_t);
        idx[0] = uint32_t(y) * t.width + uint32_t(x);
        return 1;
    }
    double iu = std::floor(u * W - 0.5), iv = std::floor(v * H - 0.5);
    int x0 = wrap_coord(int(iu), W, t.wrap_s), x1 = wrap_coord(int(iu) + 1, W, t.wrap_s);
    int y0 = wrap_coord(int(iv), H, t.wrap_t), y1 = wrap_coord(int(iv) + 1, H, t.wrap_t);
    idx[0] = uint32_t(y0) * t.width + uint32_t(x0);
    idx[1] = uint32_t(y0) * t.width + uint32_t(x1);
    idx[2] = uint32_t(y1) * t.width + uint32_t(x0);
    idx[3] = uint32_t(y1) * t.width + uint32_t(x1);
    return 4;
}

// Texture fetch of the detached evaluation; `touched` (or NULL) collects the
// watched texels the fetch reads.
template <class R> V4<R> fetch_tex_eval(const Scene& s, int32_t index, V4<R> uv0and1, int32_t coord,
                                        std::vector<uint32_t>* touched) {
    R u = 0, v = 0;
    uint32_t c = uint32_t(coord);
    if (c == 0) { u = uv0and1.x; v = uv0and1.y; }
    else if (c == 1) { u = uv0and1.z; v = uv0and1.w; }
    double out[4] = {1, 1, 1, 1};
    if (index >= 0 && size_t(index) < s.eval_tex.size()) {
        texture_sample(s.eval_tex[size_t(index)], double(u), double(v), out);
        if (touched && !s.watch.empty()) {
            uint32_t idx[4];
            int n = texture_taps(s.eval_tex[size_t(index)], double(u), double(v), idx);
            for (int k = 0; k < n; k++) {
                auto it = s.watch.find((uint64_t(uint32_t(index)) << 32) | idx[k]);
                if (it != s.watch.end()) touched->push_back(it->second);
            }
        }
    }
    return {R(out[0]), R(out[1]), R(out[2]), R(out[3])};
}

// The sampled weight of evalIndirectCombinedBRDF for a direction L drawn with the
// sampling material ms, with the BRDF values of me and the density of ms
// (PBR.glsl:649-729 with f / pdf split): specular F D G2 / (G1(v) D)|ms, diffuse
// (1 - F s) c_diff Fd (1 - t), transmission c_diff t.
template <class R> V3<R> detached_weight(uint32_t type, V3<R> N, V3<R> L, V3<R> V, const MatBrdf<R>& ms,
                                         const MatBrdf<R>& me, bool delta) {
    if (type == SPECULAR_TYPE) {
        V3<R> H = normalize(L + V);
        R hdl = std::max(R(0.00001), std::min(R(1), dot(H, L)));
        R ndl = std::max(R(0.00001), std::min(R(1), dot(N, L)));
        R ndv = std::max(R(0.00001), std::min(R(1), dot(N, V)));
        R ndh = saturate(dot(N, H));
        V3<R> F = fresnel_schlick(me.F0, shadowed_f90(me.F90), hdl);
        R ae = me.roughness * me.roughness, as = ms.roughness * ms.roughness;
        if (delta || as == R(0)) return F * me.specular_factor;
        R g1v_e = smith_g1_ggx(ae, ndv, ae * ae, ndv * ndv), g1v_s = smith_g1_ggx(as, ndv, as * as, ndv * ndv);
        R g2_e = smith_g2_over_g1(ae, ae * ae, ndl, ndv) * g1v_e;
        R d_e = ggx_d(ae * ae, ndh), d_s = ggx_d(as * as, ndh);
        return F * (g2_e * d_e / (g1v_s * d_s) * me.specular_factor);
    }
    if (type == DIFFUSE_TYPE) {
        BrdfData<R> d = prepare_brdf(N, L, V, me);
        R fb = frostbite(d.roughness, d.n_dot_l, d.n_dot_v, d.l_dot_h);
        return (V3<R>(R(1)) - d.F * me.specular_factor) * d.c_diff * fb * (R(1) - me.transmission);
    }
    return vmax(V3<R>(R(0)), me.c_diff * me.transmission);
}

template <class R> bool anyhit_ignores(const AnyHitCtx& ah, uint32_t custom_index, uint32_t prim, R bu, R bv) {
    const Scene& s = *ah.s;
    if (ah.calls) (*ah.calls)++;
    const rt_prim_info& pi = s.d.prim_infos[custom_index];
    const rt_material_raw& mat = s.d.materials[pi.material_id];
    if (mat.alpha_mode == RT_ALPHA_OPAQUE) return false;
    uint32_t io = pi.i_offset + 3 * prim;
    const rt_vertex& v0 = s.d.vertices[pi.v_offset + s.d.indices[io]];
    const rt_vertex& v1 = s.d.vertices[pi.v_offset + s.d.indices[io + 1]];
    const rt_vertex& v2 = s.d.vertices[pi.v_offset + s.d.indices[io + 2]];
    R b0 = R(1) - bu - bv;
    auto mix4 = [&](const float* a, const float* b, const float* c) {
        return load4<R>(a) * b0 + load4<R>(b) * bu + load4<R>(c) * bv;
    };
    V4<R> uv = mix4(v0.uv0, v1.uv0, v2.uv0);
    V4<R> vcol = mix4(v0.color, v1.color, v2.color);
    V4<R> color4 = vcol * load4<R>(mat.base_color);
    if (mat.base_color_texture.index >= 0)
        color4 = color4 * fetch_tex<R>(s, mat.base_color_texture.index, uv, mat.base_color_texture.tex_coord);
    R opacity = color4.w;
    if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) {
        V4<R> df = load4<R>(mat.sg_diffuse_factor);
        if (mat.sg_diffuse_texture.index >= 0)
            df = df * fetch_tex<R>(s, mat.sg_diffuse_texture.index, uv, mat.sg_diffuse_texture.tex_coord);
        color4 = vcol * df;
        opacity = color4.w;
    }
    if (mat.alpha_mode == RT_ALPHA_MASK) return opacity < R(mat.alpha_cutoff);
    R u = R(rand_next(*ah.rng));
    return opacity <= u;
}

// Traversal with traceRayEXT semantics: closest accepted hit in [tmin, tmax]
// (ties in t: smaller (instance, primitive)), any-hit on non-opaque geometry,
// optional termination on the first accepted hit.
template <class R> HitRec trace(const Scene& s, const double o[3], const double d[3], double tmin, double tmax,
                                uint32_t ray_flags, const AnyHitCtx* ah, bool brute, orc_stats* st) {
    HitRec best;
    best.found = false;
    best.t = tmax;
    best.inst = best.prim = -1;
    best.u = best.v = 0;
    bool terminate = (ray_flags & RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT) != 0;
    bool done = false;
    auto visit_instance = [&](int ii) {
        const Instance& in = s.inst[ii];
        if ((in.mask & 0xFFu) == 0) return;
        const orc_blas& bl = s.d.blases[in.blas];
        if (bl.prim_count == 0) return;
        double oo[3], od[3];
        xform_point(in.w2o, o, oo);
        xform_dir(in.w2o, d, od);
        WoopRay w = woop_setup(od);
        bool opaque = bl.opaque != 0;
        if (in.flags & RT_INSTANCE_FORCE_OPAQUE) opaque = true;
        if (in.flags & RT_INSTANCE_FORCE_NO_OPAQUE) opaque = false;
        if (ray_flags & RT_RAY_FLAG_OPAQUE) opaque = true;
        if (ray_flags & RT_RAY_FLAG_NO_OPAQUE) opaque = false;
        const std::vector<double>& tp = s.tri_pos[in.blas];
        auto test_tri = [&](int p) {
            if (st) st->tri_tests++;
            double t, bu, bv;
            if (!woop_intersect(w, oo, &tp[9 * size_t(p)], &tp[9 * size_t(p) + 3], &tp[9 * size_t(p) + 6], tmin, best.t, t, bu, bv))
                return;
            if (best.found && t == best.t && (ii > best.inst || (ii == best.inst && p > best.prim))) return;
            if (!opaque && ah && anyhit_ignores<R>(*ah, in.custom_index, uint32_t(p), R(bu), R(bv))) return;
            best.found = true;
            best.t = t;
            best.u = bu;
            best.v = bv;
            best.inst = ii;
            best.prim = p;
            if (terminate) done = true;
        };
        if (brute) {
            for (uint32_t p = 0; p < bl.prim_count && !done; p++) test_tri(int(p));
            return;
        }
        const Bvh& bvh = s.blas_bvh[in.blas];
        double invd[3] = {1.0 / od[0], 1.0 / od[1], 1.0 / od[2]};
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0 && !done) {
            const BNode& n = bvh.nodes[stack[--sp]];
            if (st) st->node_visits++;
            double tn;
            if (!ray_box(n.box, oo, invd, tmin, best.t, tn)) continue;
            if (n.count > 0) {
                for (int k = 0; k < n.count && !done; k++) test_tri(bvh.order[n.first + k]);
                continue;
            }
            double tl = 0, tr = 0;
            bool hl = ray_box(bvh.nodes[n.left].box, oo, invd, tmin, best.t, tl);
            bool hr = ray_box(bvh.nodes[n.right].box, oo, invd, tmin, best.t, tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = n.right; stack[sp++] = n.left; }
                else { stack[sp++] = n.left; stack[sp++] = n.right; }
            } else if (hl) stack[sp++] = n.left;
            else if (hr) stack[sp++] = n.right;
        }
    };
    if (brute || s.inst.size() < 2) {
        for (size_t i = 0; i < s.inst.size() && !done; i++) visit_instance(int(i));
    } else {
        double invd[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
        int stack[128];
        int sp = 0;
        stack[sp++] = 0;
        while (sp > 0 && !done) {
            const BNode& n = s.tlas.nodes[stack[--sp]];
            double tn;
            if (!ray_box(n.box, o, invd, tmin, best.t, tn)) continue;
            if (n.count > 0) {
                for (int k = 0; k < n.count && !done; k++) visit_instance(s.tlas.order[n.first + k]);
                continue;
            }
            stack[sp++] = n.right;
            stack[sp++] = n.left;
        }
    }
    return best;
}

// --------------------------------------------------------------- integrator
template <class R> struct Payload {  // RayPayload, lib/RayTracingCommons.glsl:11-24
    V3<R> hit_value, hit_point;
    V3<R> hit_value_s;  // detached evaluation: the weight of the sampling material
    R t;
    V3<R> scatter_direction;
    bool need_scatter;
    uint32_t random_seed;
    V3<R> emittance;
    U4 rng;
    R volume_dis;
};

template <class R> V3<R> m34_point(const double m[12], V3<R> p) {
    return {R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z + R(m[3]),
            R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z + R(m[7]),
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z + R(m[11])};
}
template <class R> V3<R> m34_dir(const double m[12], V3<R> p) {
    return {R(m[0]) * p.x + R(m[1]) * p.y + R(m[2]) * p.z,
            R(m[4]) * p.x + R(m[5]) * p.y + R(m[6]) * p.z,
            R(m[8]) * p.x + R(m[9]) * p.y + R(m[10]) * p.z};
}

template <class R> struct Tracer {
    const Scene& s;
    const rt_ubo& ubo;
    orc_stats* st;
    double last_t = 0.0;  // Ray.t after the pixel's last trace (read by the DISTANCE mapping, rgen:160)
    std::vector<uint32_t>* touched = nullptr;  // watched texels read (orc_texel_footprints)

    // intensity of a point / directional light in the NEE and miss terms
    R plight_intensity(const rt_light* l) const {
        size_t i = size_t(l - s.d.plights);
        return i < s.eval_pint.size() ? R(s.eval_pint[i]) : R(l->intensity);
    }
    R dlight_intensity(uint32_t i) const {
        return i < s.eval_dint.size() ? R(s.eval_dint[i]) : R(s.d.dlights[i].intensity);
    }

    HitRec cast(V3<R> o, V3<R> d, R tmin, R tmax, uint32_t flags, U4* rng) {
        double od[3] = {double(o.x), double(o.y), double(o.z)};
        double dd[3] = {double(d.x), double(d.y), double(d.z)};
        AnyHitCtx ah{&s, rng, st ? &st->anyhit_calls : nullptr};
        return trace<R>(s, od, dd, double(tmin), double(tmax), flags, &ah, false, st);
    }

    // castShadowRay: RayTracing.rchit:35-59.  The shadow payload gets a copy of
    // Ray.rngState (the value on entry to closest-hit); the copy's draws are
    // written back to Ray.rngState and then overwritten at rchit:476.
    bool shadow_unoccluded(V3<R> p, V3<R> L, R dist, U4 rng_copy) {
        if (st) st->shadow_rays++;
        uint32_t flags = RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RT_RAY_FLAG_SKIP_CLOSEST_HIT;
        if (ubo.fully_opaque) flags |= RT_RAY_FLAG_OPAQUE;
        HitRec h = cast(p, L, R(0.1), dist, flags, &rng_copy);
        return !h.found;
    }

    // sampleLightRIS + sampleLightUniform: RayTracing.rchit:62-122.
    bool sample_light_ris(U4& rng, V3<R> pos, V3<R> n, const rt_light*& sel, R& weight) {
        uint32_t num = s.d.num_plights;
        if (num == 0) return false;
        R total = 0, sel_pdf = 0;
        uint32_t cand = std::min(num, 3u);
        for (uint32_t i = 0; i < cand; i++) {
            const rt_light& li = s.d.plights[i];
            V3<R> c = V3<R>(R(li.color[0]), R(li.color[1]), R(li.color[2])) * R(li.intensity);
            if (luminance(c) < R(0.1)) continue;
            float u = rand_next(rng);
            uint32_t idx = std::min(num - 1, uint32_t(R(u) * R(num)));
            const rt_light& candl = s.d.plights[idx];
            R cw = R(num);
            V3<R> lv = V3<R>(R(candl.transform[0]), R(candl.transform[1]), R(candl.transform[2])) - pos;
            R ld = length(lv);
            V3<R> L = normalize(lv);
            if (dot(n, L) < R(0.00001)) continue;
            R pdf = luminance(light_intensity_at<R>(candl, length(lv)));
            R ris = pdf * cw;
            total += ris;
            (void)ld;
            if (R(rand_next(rng)) < ris / total) {
                sel = &candl;
                sel_pdf = pdf;
            }
        }
        if (total == R(0)) return false;
        weight = (total / R(3)) / sel_pdf;
        return true;
    }

    // RayTracing.rmiss:14-45
    void miss(Payload<R>& ray, V3<R> dir) {
        V3<R> acc(R(0));
        V3<R> rd = normalize(dir);
        if (ray.t != R(0)) {
            for (uint32_t i = 0; i < s.d.num_dlights; i++) {
                const rt_light& li = s.d.dlights[i];
                R c = dot(normalize(load3<R>(li.transform)), rd);
                if (c < R(0)) acc = acc + load3<R>(li.color) * (-c * dlight_intensity(i));
            }
        }
        if (ubo.has_sky) {  // rmiss:27-33: light_acc += skyColor + light_acc
            double dd[3] = {double(rd.x), double(rd.y), double(rd.z)}, c[4];
            cube_sample(s.sky, dd, c);
            acc = acc + (V3<R>(R(c[0]), R(c[1]), R(c[2])) + acc);
        } else {
            acc = acc + V3<R>(R(0.01));
        }
        ray.hit_value = V3<R>(R(0));
        ray.hit_value_s = ray.hit_value;
        ray.need_scatter = false;
        ray.emittance = acc;
        if (s.d.num_dlights == 0) ray.emittance = V3<R>(R(0));
        ray.t = R(-1);
    }

    // RayTracing.rchit main :136-477 (mapping RENDER and the debug mappings).
    void closest_hit(Payload<R>& ray, const HitRec& h, V3<R> world_dir) {
        if (st) st->shade_events++;
        const Instance& in = s.inst[h.inst];
        const rt_prim_info& pi = s.d.prim_infos[in.custom_index];
        const rt_material_raw& mat = s.d.materials[pi.material_id];
        V3<R> last_hit = ray.hit_point;
        ray.t = R(h.t);
        uint32_t io = pi.i_offset + 3 * uint32_t(h.prim);
        const rt_vertex& v0 = s.d.vertices[pi.v_offset + s.d.indices[io]];
        const rt_vertex& v1 = s.d.vertices[pi.v_offset + s.d.indices[io + 1]];
        const rt_vertex& v2 = s.d.vertices[pi.v_offset + s.d.indices[io + 2]];
        // getMixVertexAndGeoNormal: lib/RayTracingCommons.glsl:86-101
        R bu = R(h.u), bv = R(h.v), b0 = R(1) - bu - bv;
        auto mix3 = [&](V3<R> a, V3<R> b, V3<R> c) { return a * b0 + b * bu + c * bv; };
        auto mix4 = [&](V4<R> a, V4<R> b, V4<R> c) { return a * b0 + b * bu + c * bv; };
        V4<R> uv = mix4(load4<R>(v0.uv0), load4<R>(v1.uv0), load4<R>(v2.uv0));
        V3<R> p0 = load3<R>(v0.pos), p1 = load3<R>(v1.pos), p2 = load3<R>(v2.pos);
        V3<R> pos = mix3(p0, p1, p2);
        V4<R> vcol = mix4(load4<R>(v0.color), load4<R>(v1.color), load4<R>(v2.color));
        V3<R> normal = normalize(mix3(normalize(load3<R>(v0.normal)), normalize(load3<R>(v1.normal)), normalize(load3<R>(v2.normal))));
        V3<R> geo = cross(p1 - p0, p2 - p0);  // calculate_geo_normal, :79-84
        V3<R> geo_normal = normalize(m34_dir(in.o2w, geo));  // normal_transform, rchit:28-30
        V3<R> origin = m34_point(in.o2w, pos);
        V4<R> color4 = vcol * load4<R>(mat.base_color);
        if (mat.base_color_texture.index >= 0) {
            color4 = color4 * fetch_tex<R>(s, mat.base_color_texture.index, uv, mat.base_color_texture.tex_coord);
        }
        V3<R> color = color4.xyz();
        ray.need_scatter = false;
        ray.hit_point = pos;
        if (mat.normal_texture.index >= 0) {  // rchit:178-185, getNormal :124-129
            V4<R> tx = fetch_tex<R>(s, mat.normal_texture.index, uv, mat.normal_texture.tex_coord);
            V3<R> nt = normalize(V3<R>(tx.x * R(2) - R(1), tx.y * R(2) - R(1), tx.z * R(2) - R(1)));
            V4<R> tan4 = normalize4(mix4(load4<R>(v0.tangent), load4<R>(v1.tangent), load4<R>(v2.tangent)));
            V3<R> tm = tan4.xyz();
            V3<R> tangent = normalize(tm - normal * dot(tm, normal));
            V3<R> bit = normalize(cross(normal, tangent) * tan4.w);
            normal = tangent * nt.x + bit * nt.y + normal * nt.z;
        }
        normal = normalize(m34_dir(in.o2w, normal));
        V3<R> V = -normalize(world_dir);
        R cosv = dot(V, geo_normal);
        bool front = cosv >= R(0);
        geo_normal = front ? geo_normal : -geo_normal;
        V3<R> outward = dot(geo_normal, normal) < R(0) ? -normal : normal;
        V3<R> emit = load3<R>(mat.emissive_factor);
        if (mat.emissive_texture.index >= 0) {
            V4<R> tx = fetch_tex<R>(s, mat.emissive_texture.index, uv, mat.emissive_texture.tex_coord);
            emit = emit * tx.xyz();
        }
        R metallic = R(mat.metallic_factor), roughness = R(mat.roughness_factor);
        if (mat.metallic_roughness_texture.index >= 0) {
            V4<R> tx = fetch_tex<R>(s, mat.metallic_roughness_texture.index, uv, mat.metallic_roughness_texture.tex_coord);
            roughness = roughness * tx.y;
            metallic = metallic * tx.z;
        }
        V3<R> spec_wf(R(1));
        if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) {  // rchit:215-239
            V4<R> df = load4<R>(mat.sg_diffuse_factor);
            V4<R> sgf = load4<R>(mat.sg_specular_glossiness_factor);
            if (mat.sg_diffuse_texture.index >= 0)
                df = df * fetch_tex<R>(s, mat.sg_diffuse_texture.index, uv, mat.sg_diffuse_texture.tex_coord);
            if (mat.sg_specular_glossiness_texture.index >= 0)
                sgf = sgf * fetch_tex<R>(s, mat.sg_specular_glossiness_texture.index, uv, mat.sg_specular_glossiness_texture.tex_coord);
            spec_wf = sgf.xyz();
            roughness = R(1) - sgf.w;
            R max_spec = std::max(std::max(spec_wf.x, spec_wf.y), spec_wf.z);
            color = vcol.xyz() * df.xyz();
            metallic = convert_metallic(color, spec_wf, max_spec);
        }
        R trans = R(0);
        if (mat.transmission_exist) {
            trans = R(mat.transmission_factor);
            if (mat.transmission_texture.index >= 0)
                trans = trans * fetch_tex<R>(s, mat.transmission_texture.index, uv, mat.transmission_texture.tex_coord).x;
        }
        uint32_t mapping = ubo.mapping;
        if (mat.unlit) mapping = RT_MAPPING_ALBEDO;
        switch (mapping) {  // rchit:258-286
            case RT_MAPPING_ALBEDO: ray.emittance = color; return;
            case RT_MAPPING_TRIANGLE: ray.emittance = V3<R>(R(1) - bu - bv, bu, bv); return;
            case RT_MAPPING_METALLIC: ray.emittance = V3<R>(metallic); return;
            case RT_MAPPING_ROUGHNESS: ray.emittance = V3<R>(roughness); return;
            case RT_MAPPING_NORMAL: ray.emittance = (outward + V3<R>(R(1))) / R(2); return;
            case RT_MAPPING_TRANSMISSION: ray.emittance = V3<R>(trans); return;
            case RT_MAPPING_INSTANCE:
            case RT_MAPPING_GEO_ID: {
                uint32_t x = mapping == RT_MAPPING_INSTANCE ? uint32_t(h.inst) : in.custom_index;
                x += x << 10; x ^= x >> 6; x += x << 3; x ^= x >> 11; x += x << 15;  // jenkinsHash
                ray.emittance = V3<R>(R((x >> 0) & 0xFFu) / R(255), R((x >> 8) & 0xFFu) / R(255), R((x >> 16) & 0xFFu) / R(255));
                return;
            }
            case RT_MAPPING_TANGENT: {
                V4<R> tan4 = normalize4(mix4(load4<R>(v0.tangent), load4<R>(v1.tangent), load4<R>(v2.tangent)));
                ray.emittance = (normalize(m34_dir(in.o2w, tan4.xyz())) + V3<R>(R(1))) / R(2);
                return;
            }
            default: break;
        }
        R spec_factor = R(mat.specular_factor);
        V3<R> spec_color = load3<R>(mat.specular_color_factor);
        if (mat.specular_texture.index >= 0)
            spec_factor = spec_factor * fetch_tex<R>(s, mat.specular_texture.index, uv, mat.specular_texture.tex_coord).w;
        if (mat.specular_color_texture.index >= 0)
            spec_color = spec_color * fetch_tex<R>(s, mat.specular_color_texture.index, uv, mat.specular_color_texture.tex_coord).xyz();
        ray.hit_point = origin;
        U4 rng = ray.rng;
        ray.emittance = emit * R(ubo.exposure);
        ray.need_scatter = false;
        MatBrdf<R> m;
        m.base_color = color;
        m.metallic = metallic;
        m.roughness = roughness;
        m.ior = mat.volume_exists ? R(mat.ior) : R(1);
        m.transmission = trans;
        m.specular_factor = spec_factor;
        m.specular_color_factor = spec_color;
        m.use_spec = mat.specular_exist != 0;
        m.front_face = front;
        mat_build(m);
        if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) {
            R max_spec = std::max(std::max(spec_wf.x, spec_wf.y), spec_wf.z);
            m.c_diff = color * (R(1) - max_spec);
            m.F0 = spec_wf;
        }
        m.attenuation_color = load3<R>(mat.attenuation_color);
        m.attenuation_distance = R(mat.attenuation_distance);
        m.volume = mat.volume_exists != 0;
        const bool detached = !s.eval_tex.empty();
        MatBrdf<R> me = m;  // BRDF values (detached evaluation), else m itself
        if (detached && mat.workflow != RT_WORKFLOW_SPECULAR_GLOSSINESS) {
            V4<R> c4 = vcol * load4<R>(mat.base_color);
            if (mat.base_color_texture.index >= 0)
                c4 = c4 * fetch_tex_eval<R>(s, mat.base_color_texture.index, uv, mat.base_color_texture.tex_coord,
                                            touched);
            R me_metal = R(mat.metallic_factor), me_rough = R(mat.roughness_factor);
            if (mat.metallic_roughness_texture.index >= 0) {
                V4<R> tx = fetch_tex_eval<R>(s, mat.metallic_roughness_texture.index, uv,
                                             mat.metallic_roughness_texture.tex_coord, touched);
                me_rough = me_rough * tx.y;
                me_metal = me_metal * tx.z;
            }
            me.base_color = c4.xyz();
            me.metallic = me_metal;
            me.roughness = me_rough;
            mat_build(me);
        }
        R displacement = length(origin - last_hit);
        m.t_diff = displacement;
        const rt_light* light = nullptr;
        R lw = 0;
        if (sample_light_ris(rng, origin, geo_normal, light, lw)) {  // rchit:338-350
            V3<R> lv = load3<R>(light->transform) - origin;
            R ldist = length(lv);
            lv = normalize(lv);
            if (shadow_unoccluded(origin, lv, ldist, ray.rng)) {
                V3<R> f = eval_combined_brdf(outward, lv, V, detached ? me : m);
                ray.emittance = ray.emittance + f * (lw * plight_intensity(light)) * load3<R>(light->color);
            }
        }
        uint32_t type;
        V3<R> thr(R(1));
        if (metallic == R(1) && roughness == R(0)) {  // rchit:352-395
            type = SPECULAR_TYPE;
        } else {
            V3<R> p = brdf_probability(m, V, outward);
            R r = R(rand_next(rng));
            if (r < p.x) {
                type = SPECULAR_TYPE;
                thr = thr / p.x;
                if (ray.volume_dis >= R(0)) ray.volume_dis += displacement;
            } else if (r >= p.x && r <= p.x + p.y) {
                type = DIFFUSE_TYPE;
                thr = thr / p.y;
                if (ray.volume_dis >= R(0)) ray.volume_dis += displacement;
            } else {
                type = TRANSMISSION_TYPE;
                if (m.volume) {
                    if (ray.volume_dis >= R(0)) ray.volume_dis += displacement;
                    else if (front) ray.volume_dis = R(0);
                }
                thr = thr / p.z;
            }
        }
        if (type == TRANSMISSION_TYPE) {  // rchit:397-399
            float pf[3] = {float(origin.x), float(origin.y), float(origin.z)};
            float nf[3] = {float(-geo_normal.x), float(-geo_normal.y), float(-geo_normal.z)};
            float of[3];
            offset_ray_f(pf, nf, of);
            origin = V3<R>(R(of[0]), R(of[1]), R(of[2]));
        }
        R ux = R(rand_next(rng));
        R uy = R(rand_next(rng));
        V3<R> dir(R(0)), w(R(0));
        ray.need_scatter = sample_indirect(ux, uy, outward, geo_normal, V, m, type, dir, w, ray.volume_dis);
        ray.hit_value_s = thr * w;
        if (detached && ray.need_scatter)
            w = detached_weight(type, outward, dir, V, m, me, metallic == R(1) && roughness == R(0));
        thr = thr * w;
        ray.hit_point = origin;
        ray.scatter_direction = dir;
        ray.hit_value = thr;
        ray.rng = rng;
        if (ubo.debug == 1) {  // rchit:437-475, from the unnormalised gl_WorldRayDirectionEXT
            double u3[3];
            if (trans > R(0)) {
                R ratio = front ? R(1) / R(mat.ior) : R(mat.ior);
                R cos_theta = std::fabs(cosv);
                V3<R> refracted = refract(world_dir, outward, ratio);
                R reflect_prob = (refracted.x != R(0) || refracted.y != R(0) || refracted.z != R(0))
                                     ? schlick(cos_theta, ratio) : R(1);
                ray.hit_value = color;
                ray.need_scatter = true;
                if (R(lcg_float(ray.random_seed)) < reflect_prob) ray.scatter_direction = reflect(world_dir, normal);
                else ray.scatter_direction = refracted;
            } else if (length(emit) < R(0.01) && roughness == R(1)) {
                bool scattered = dot(world_dir, geo_normal) < R(0);
                lcg_unit_sphere(ray.random_seed, u3);
                ray.need_scatter = scattered;
                ray.scatter_direction = normalize(outward + V3<R>(R(u3[0]), R(u3[1]), R(u3[2])));
                ray.hit_value = scattered ? color : V3<R>(R(0));
            } else if (metallic > R(0)) {
                V3<R> reflected = reflect(world_dir, outward);
                bool scattered = dot(reflected, geo_normal) > R(0);
                ray.need_scatter = scattered;
                ray.hit_value = scattered ? color : V3<R>(R(0));
                lcg_unit_sphere(ray.random_seed, u3);
                ray.scatter_direction = reflected + V3<R>(R(u3[0]), R(u3[1]), R(u3[2])) * R(0.08);
            }
        }
    }

    // RayTracing.rgen main :25-167, for one pixel; returns the summed radiance.
    V3<R> pixel(uint32_t x, uint32_t y, uint32_t W, uint32_t H) {
        uint32_t seed_b = stream_b_seed(ubo.random_seed, ubo.frame_count);
        Payload<R> ray;
        ray.hit_value = ray.hit_value_s = V3<R>(R(0));
        ray.random_seed = tea16(tea16(x, y), seed_b);
        U4 rng_a{x, y, ubo.frame_count, 0};
        ray.rng = U4{x, y, seed_b, 0};
        ray.hit_point = V3<R>(R(0));
        ray.hit_value = V3<R>(R(0));
        ray.hit_value_s = ray.hit_value;
        ray.emittance = V3<R>(R(0));
        ray.need_scatter = false;
        ray.t = R(0);
        ray.volume_dis = R(-1);
        V3<R> radiance(R(0));
        const float* mvi = ubo.model_view_inverse;
        const float* pin = ubo.projection_inverse;
        auto mat4_vec4 = [](const float* m, R a, R b, R c, R d) {  // column-major mat4 * vec4
            return V4<R>(R(m[0]) * a + R(m[4]) * b + R(m[8]) * c + R(m[12]) * d,
                         R(m[1]) * a + R(m[5]) * b + R(m[9]) * c + R(m[13]) * d,
                         R(m[2]) * a + R(m[6]) * b + R(m[10]) * c + R(m[14]) * d,
                         R(m[3]) * a + R(m[7]) * b + R(m[11]) * c + R(m[15]) * d);
        };
        for (uint32_t sidx = 0; sidx < ubo.number_of_samples; sidx++) {
            R px = R(x) + R(0.5), py = R(y) + R(0.5);
            if (ubo.antialiasing) {
                R jx = R(rand_next(rng_a));
                R jy = R(rand_next(rng_a));
                px += jx - R(0.5);
                py += jy - R(0.5);
            }
            R uvx = (px / R(W)) * R(2) - R(1);
            R uvy = (py / R(H)) * R(2) - R(1);
            float dx, dy;
            lcg_unit_disk(ray.random_seed, dx, dy);
            R offx = R(ubo.aperture) / R(2) * R(dx), offy = R(ubo.aperture) / R(2) * R(dy);
            V4<R> origin = mat4_vec4(mvi, offx, offy, R(0), R(1));
            V4<R> target = mat4_vec4(pin, uvx, uvy, R(1), R(1));
            V3<R> tdir = normalize(target.xyz() * R(ubo.focus_distance) - V3<R>(offx, offy, R(0)));
            V4<R> direction = mat4_vec4(mvi, tdir.x, tdir.y, tdir.z, R(0));
            R tfar = R(10000.0);
            if (ubo.orthographic_fov_dis > 0.f) {  // rgen:67-75
                R k = R(1) + R(ubo.orthographic_fov_dis);
                origin = mat4_vec4(mvi, k * uvx, -(k * uvy), R(0), R(1));
                direction = mat4_vec4(mvi, R(0), R(0), R(-1), R(0));
                tfar = R(10) * R(10000.0);
            }
            V3<R> thr(R(1)), thr_s(R(1));
            ray.t = R(0);
            ray.volume_dis = R(-1);
            uint32_t flags = ubo.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
            V3<R> o = origin.xyz(), d = direction.xyz();
            for (uint32_t b = 0; b < ubo.number_of_bounces; b++) {
                if (st) { if (b == 0) st->primary_rays++; else st->extension_rays++; }
                HitRec h = cast(o, d, R(0.001), tfar, flags, &ray.rng);
                if (h.found) closest_hit(ray, h, d);
                else miss(ray, d);
                V3<R> hit_color = ray.hit_value;
                R t = ray.t;
                bool scattered = ray.need_scatter;
                radiance = radiance + thr * ray.emittance;
                ray.emittance = V3<R>(R(0));
                if (b + 1 == ubo.number_of_bounces) break;
                if (b > 3) {  // Russian roulette, rgen:109-118 (MIN_BOUNCES = 3)
                    R rr = clampr(luminance(s.eval_tex.empty() ? thr : thr_s), R(0.01), R(0.95));
                    R prop = R(rand_next(rng_a));
                    if (rr < prop) break;
                    thr = thr / rr;
                    thr_s = thr_s / rr;
                }
                thr = thr * hit_color;
                if (!s.eval_tex.empty()) thr_s = thr_s * ray.hit_value_s;
                if (!scattered || t < R(0)) break;
                o = ray.hit_point;
                d = ray.scatter_direction;
            }
        }
        last_t = double(ray.t);
        return radiance;
    }
};

}  // namespace

// ------------------------------------------------------------------ C API
extern "C" {

void orc_pcg4d(const uint32_t in[4], uint32_t out[4]) {
    U4 r = pcg4d(U4{in[0], in[1], in[2], in[3]});
    out[0] = r.x; out[1] = r.y; out[2] = r.z; out[3] = r.w;
}
uint32_t orc_rand_sequence(const uint32_t state[4], uint32_t n, float* out) {
    U4 s{state[0], state[1], state[2], state[3]};
    for (uint32_t i = 0; i < n; i++) out[i] = rand_next(s);
    return s.w;
}
uint32_t orc_tea16(uint32_t v0, uint32_t v1) { return tea16(v0, v1); }
uint32_t orc_lcg_sequence(uint32_t seed, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; i++) out[i] = lcg_float(seed);
    return seed;
}
uint32_t orc_stream_b_seed(uint32_t random_seed, uint32_t frame_count) { return stream_b_seed(random_seed, frame_count); }

static MatBrdf<double> to_mat(const orc_matbrdf* m) {
    MatBrdf<double> r;
    r.base_color = V3<double>(m->base_color[0], m->base_color[1], m->base_color[2]);
    r.metallic = m->metallic;
    r.roughness = m->roughness;
    r.ior = m->ior;
    r.transmission = m->transmission;
    r.use_spec = m->use_spec != 0;
    r.specular_factor = m->specular_factor;
    r.specular_color_factor = V3<double>(m->specular_color_factor[0], m->specular_color_factor[1], m->specular_color_factor[2]);
    r.front_face = m->front_face != 0;
    r.volume = m->volume != 0;
    r.attenuation_color = V3<double>(m->attenuation_color[0], m->attenuation_color[1], m->attenuation_color[2]);
    r.attenuation_distance = m->attenuation_distance;
    r.t_diff = 0;
    mat_build(r);
    if (m->workflow_sg) {
        V3<double> sp(m->sg_specular[0], m->sg_specular[1], m->sg_specular[2]);
        double mx = std::max(std::max(sp.x, sp.y), sp.z);
        r.c_diff = r.base_color * (1.0 - mx);
        r.F0 = sp;
    }
    return r;
}
void orc_mat_build(orc_matbrdf* m) {
    MatBrdf<double> r = to_mat(m);
    for (int k = 0; k < 3; k++) { m->F0[k] = r.F0[k]; m->F90[k] = r.F90[k]; m->c_diff[k] = r.c_diff[k]; }
}
static V3<double> v3d(const double* p) { return V3<double>(p[0], p[1], p[2]); }
static void put3(double* o, V3<double> v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
void orc_eval_combined_brdf(const double N[3], const double L[3], const double V[3], const orc_matbrdf* m, double out[3]) {
    put3(out, eval_combined_brdf(v3d(N), v3d(L), v3d(V), to_mat(m)));
}
void orc_brdf_probability(const orc_matbrdf* m, const double V[3], const double N[3], double out[3]) {
    put3(out, brdf_probability(to_mat(m), v3d(V), v3d(N)));
}
int orc_sample_indirect_vd(const double u[2], const double sn[3], const double gn[3], const double V[3],
                           const orc_matbrdf* m, uint32_t type, double* volume_dis, double dir[3], double weight[3]) {
    V3<double> d(0.0), w(0.0);
    bool r = sample_indirect(u[0], u[1], v3d(sn), v3d(gn), v3d(V), to_mat(m), type, d, w, *volume_dis);
    put3(dir, d);
    put3(weight, w);
    return r ? 1 : 0;
}
int orc_sample_indirect(const double u[2], const double sn[3], const double gn[3], const double V[3], const orc_matbrdf* m,
                        uint32_t type, double dir[3], double weight[3]) {
    double vd = -1;
    return orc_sample_indirect_vd(u, sn, gn, V, m, type, &vd, dir, weight);
}
double orc_ggx_d(double a2, double ndh) { return ggx_d(a2, ndh); }
double orc_smith_g2_height_correlated(double a, double ndl, double ndv) { return smith_g2_hc(a, ndl, ndv); }
double orc_smith_g1_ggx(double a, double nds, double a2, double nds2) { return smith_g1_ggx(a, nds, a2, nds2); }
void orc_sample_ggx_vndf(const double Ve[3], double alpha, const double u[2], double out[3]) {
    put3(out, sample_ggx_vndf(v3d(Ve), alpha, alpha, u[0], u[1]));
}
void orc_sample_hemisphere(const double u[2], double out[3]) { put3(out, sample_hemisphere(u[0], u[1])); }
void orc_rotation_to_z(const double v[3], double q[4]) {
    V4<double> r = rotation_to_z(v3d(v));
    q[0] = r.x; q[1] = r.y; q[2] = r.z; q[3] = r.w;
}
void orc_rotate_point(const double q[4], const double v[3], double out[3]) {
    put3(out, rotate_point(V4<double>(q[0], q[1], q[2], q[3]), v3d(v)));
}
double orc_convert_metallic(const double d[3], const double s[3], double mx) { return convert_metallic(v3d(d), v3d(s), mx); }
double orc_frostbite_diffuse(double r, double ndl, double ndv, double ldh) { return frostbite(r, ndl, ndv, ldh); }
double orc_light_intensity_luminance(const rt_light* l, double distance) { return luminance(light_intensity_at<double>(*l, distance)); }
void orc_offset_ray(const float p[3], const float n[3], float out[3]) { offset_ray_f(p, n, out); }
void orc_texture_sample(const rt_texture* tex, double u, double v, double out[4]) { texture_sample(*tex, u, v, out); }
void orc_cube_sample(const rt_texture* cube, const double dir[3], double out[4]) { cube_sample(*cube, dir, out); }
int orc_intersect_triangle(const double o[3], const double d[3], const double v0[3], const double v1[3],
                           const double v2[3], double tmin, double tmax, double out[3]) {
    WoopRay w = woop_setup(d);
    double t, u, v;
    if (!woop_intersect(w, o, v0, v1, v2, tmin, tmax, t, u, v)) return 0;
    out[0] = t; out[1] = u; out[2] = v;
    return 1;
}

void* orc_scene_create(const orc_scene_desc* desc) {
    if (!desc) return nullptr;
    for (uint32_t i = 0; i < desc->num_instances; i++)
        if (desc->instances[i].acceleration_structure_reference >= desc->num_blases) return nullptr;
    return build_scene(desc);
}
void orc_scene_destroy(void* s) { delete (Scene*)s; }
void orc_scene_set_eval_textures(void* s, const rt_texture* tex, uint32_t n) {
    Scene* sc = (Scene*)s;
    sc->eval_tex.assign(tex, tex + n);
}
void orc_scene_set_eval_lights(void* s, const float* plight_intensity, uint32_t np, const float* dlight_intensity,
                               uint32_t nd) {
    Scene* sc = (Scene*)s;
    sc->eval_pint.assign(plight_intensity, plight_intensity + (plight_intensity ? np : 0));
    sc->eval_dint.assign(dlight_intensity, dlight_intensity + (dlight_intensity ? nd : 0));
}

static void fill_hit(const HitRec& h, const Scene& s, rt_hit* out, double* tuv) {
    std::memset(out, 0, sizeof(*out));
    if (!h.found) {
        out->primitive_id = 0xFFFFFFFFu;
        out->instance_id = 0xFFFFFFFFu;
        out->custom_index = 0xFFFFFFFFu;
        out->t = -1.0f;
        if (tuv) { tuv[0] = -1; tuv[1] = tuv[2] = 0; }
        return;
    }
    out->t = float(h.t);
    out->u = float(h.u);
    out->v = float(h.v);
    out->instance_id = uint32_t(h.inst);
    out->primitive_id = uint32_t(h.prim);
    out->custom_index = s.inst[h.inst].custom_index;
    if (tuv) { tuv[0] = h.t; tuv[1] = h.u; tuv[2] = h.v; }
}

int orc_trace_rays(void* sp, const float* rays, uint32_t n, uint32_t flags, uint32_t seed, rt_hit* hits, double* tuv,
                   int brute, int threads) {
    const Scene& s = *(Scene*)sp;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < int64_t(n); i++) {
        const float* r = rays + 8 * size_t(i);
        double o[3] = {r[0], r[1], r[2]}, d[3] = {r[4], r[5], r[6]};
        U4 rng{uint32_t(i), 0u, seed, 0u};  // stream of a ray query: (index, 0, seed)
        AnyHitCtx ah{&s, &rng, nullptr};
        HitRec h = trace<double>(s, o, d, r[3], r[7], flags, &ah, brute != 0, nullptr);
        fill_hit(h, s, hits + i, tuv ? tuv + 3 * size_t(i) : nullptr);
    }
    return 0;
}

int orc_render_pixels(void* sp, const rt_ubo* ubo, uint32_t W, uint32_t H, const uint32_t* xy, uint32_t n,
                      int precision, int threads, double* out, orc_stats* stats) {
    const Scene& s = *(Scene*)sp;
    if (ubo->has_sky && !s.has_sky_tex) return -2;
    if (threads > 0) omp_set_num_threads(threads);
    for (uint32_t i = 0; i < n; i++)
        if (xy[2 * i] >= W || xy[2 * i + 1] >= H) return -1;
    orc_stats total{};
#pragma omp parallel
    {
        orc_stats local{};
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < int64_t(n); i++) {
            uint32_t x = xy[2 * i], y = xy[2 * i + 1];
            double* o = out + 4 * size_t(i);
            if (precision == 32) {
                Tracer<float> tr{s, *ubo, stats ? &local : nullptr};
                V3<float> r = tr.pixel(x, y, W, H);
                o[0] = r.x; o[1] = r.y; o[2] = r.z;
                o[3] = tr.last_t;
            } else {
                Tracer<double> tr{s, *ubo, stats ? &local : nullptr};
                V3<double> r = tr.pixel(x, y, W, H);
                o[0] = r.x; o[1] = r.y; o[2] = r.z;
                o[3] = tr.last_t;
            }
        }
        if (stats) {
#pragma omp critical
            {
                total.primary_rays += local.primary_rays;
                total.extension_rays += local.extension_rays;
                total.shadow_rays += local.shadow_rays;
                total.node_visits += local.node_visits;
                total.tri_tests += local.tri_tests;
                total.shade_events += local.shade_events;
                total.anyhit_calls += local.anyhit_calls;
            }
        }
    }
    if (stats) *stats = total;
    return 0;
}

int orc_texel_footprints(void* sp, const rt_ubo* ubo, uint32_t W, uint32_t H, const uint32_t* xy, uint32_t n,
                         const uint32_t* watch_tex, const uint32_t* watch_texel, uint32_t n_watch, uint32_t* out_pairs,
                         uint64_t capacity, uint64_t* n_out, int threads) {
    Scene& s = *(Scene*)sp;
    if (ubo->has_sky && !s.has_sky_tex) return -2;
    if (s.eval_tex.empty()) return -4;
    for (uint32_t i = 0; i < n; i++)
        if (xy[2 * i] >= W || xy[2 * i + 1] >= H) return -1;
    if (threads > 0) omp_set_num_threads(threads);
    s.watch.clear();
    for (uint32_t k = 0; k < n_watch; k++) s.watch[(uint64_t(watch_tex[k]) << 32) | watch_texel[k]] = k;
    std::vector<std::vector<uint32_t>> pairs;
#pragma omp parallel
    {
        std::vector<uint32_t> local, touched;
#pragma omp for schedule(dynamic, 4)
        for (int64_t i = 0; i < int64_t(n); i++) {
            touched.clear();
            Tracer<double> tr{s, *ubo, nullptr};
            tr.touched = &touched;
            (void)tr.pixel(xy[2 * i], xy[2 * i + 1], W, H);
            std::sort(touched.begin(), touched.end());
            touched.erase(std::unique(touched.begin(), touched.end()), touched.end());
            for (uint32_t w : touched) {
                local.push_back(w);
                local.push_back(uint32_t(i));
            }
        }
#pragma omp critical
        pairs.push_back(std::move(local));
    }
    s.watch.clear();
    uint64_t total = 0;
    for (auto& v : pairs) total += v.size() / 2;
    *n_out = total;
    if (total > capacity) return -3;
    uint64_t o = 0;
    for (auto& v : pairs)
        for (size_t k = 0; k < v.size(); k++) out_pairs[o++] = v[k];
    return 0;
}

int orc_trace_primary(void* sp, const rt_ubo* ubo, uint32_t W, uint32_t H, rt_hit* hits, double* tuv, int threads) {
    const Scene& s = *(Scene*)sp;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < int64_t(W) * H; i++) {
        uint32_t x = uint32_t(i % W), y = uint32_t(i / W);
        // camera ray of sample 0, RayTracing.rgen:49-75
        U4 rng_a{x, y, ubo->frame_count, 0};
        double px = x + 0.5, py = y + 0.5;
        if (ubo->antialiasing) {
            double jx = rand_next(rng_a), jy = rand_next(rng_a);
            px += jx - 0.5;
            py += jy - 0.5;
        }
        double uvx = (px / W) * 2 - 1, uvy = (py / H) * 2 - 1;
        const float* mvi = ubo->model_view_inverse;
        const float* pin = ubo->projection_inverse;
        double o[3], d[3], tg[3];
        for (int r = 0; r < 3; r++) {
            o[r] = mvi[12 + r];
            tg[r] = pin[r] * uvx + pin[4 + r] * uvy + pin[8 + r] + pin[12 + r];
        }
        double l = std::sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
        for (int r = 0; r < 3; r++) tg[r] /= l;
        for (int r = 0; r < 3; r++) d[r] = mvi[r] * tg[0] + mvi[4 + r] * tg[1] + mvi[8 + r] * tg[2];
        uint32_t flags = ubo->fully_opaque ? RT_RAY_FLAG_OPAQUE : 0u;
        uint32_t seed_b = stream_b_seed(ubo->random_seed, ubo->frame_count);
        U4 rng_b{x, y, seed_b, 0};
        AnyHitCtx ah{&s, &rng_b, nullptr};
        HitRec h = trace<double>(s, o, d, 0.001, 10000.0, flags, &ah, false, nullptr);
        fill_hit(h, s, hits + i, tuv ? tuv + 3 * size_t(i) : nullptr);
    }
    return 0;
}

int orc_render(void* sp, const rt_ubo* ubo, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1,
               int precision, int threads, double* accum, orc_stats* stats) {
    const Scene& s = *(Scene*)sp;
    if (x1 > W || y1 > H || x0 > x1 || y0 > y1) return -1;
    if (ubo->has_sky && !s.has_sky_tex) return -2;
    if (threads > 0) omp_set_num_threads(threads);
    bool reset = ubo->number_of_samples == ubo->total_number_of_samples;
    uint32_t w = x1 - x0;
    int64_t npx = int64_t(w) * (y1 - y0);
    orc_stats total{};
#pragma omp parallel
    {
        orc_stats local{};
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < npx; i++) {
            uint32_t x = x0 + uint32_t(i % w), y = y0 + uint32_t(i / w);
            double rgb[3];
            if (precision == 32) {
                Tracer<float> tr{s, *ubo, stats ? &local : nullptr};
                V3<float> r = tr.pixel(x, y, W, H);
                rgb[0] = r.x; rgb[1] = r.y; rgb[2] = r.z;
            } else {
                Tracer<double> tr{s, *ubo, stats ? &local : nullptr};
                V3<double> r = tr.pixel(x, y, W, H);
                rgb[0] = r.x; rgb[1] = r.y; rgb[2] = r.z;
            }
            double* a = accum + 4 * (size_t(y) * W + x);
            for (int k = 0; k < 3; k++) a[k] = (reset ? 0.0 : a[k]) + rgb[k];
            a[3] = 0.0;
        }
        if (stats) {
#pragma omp critical
            {
                total.primary_rays += local.primary_rays;
                total.extension_rays += local.extension_rays;
                total.shadow_rays += local.shadow_rays;
                total.node_visits += local.node_visits;
                total.tri_tests += local.tri_tests;
                total.shade_events += local.shade_events;
                total.anyhit_calls += local.anyhit_calls;
            }
        }
    }
    if (stats) *stats = total;
    return 0;
}

// ------------------------------------------------------------ display image
// RayTracing.rgen:132-166 after accumulation: tone mapping (shaders/lib/Tonemapping.glsl),
// the HEAT / DISTANCE mappings (shaders/lib/Heatmap.glsl, rgen:151-163) and the
// 8-bit UNORM store.  fp32 like the shaders; this file builds with contraction off.
static float gamma22(float c) { return std::pow(c, 1.0f / 2.2f); }  // LINEARtoSRGB, Tonemapping.glsl:9-11

static float uncharted2_impl(float x) {  // Tonemapping.glsl:15-23
    const float A = 0.15f, B = 0.50f, C = 0.10f, D = 0.20f, E = 0.02f, F = 0.30f;
    float num = x * (A * x + C * B) + D * E;
    float den = x * (A * x + B) + D * F;
    return num / den - E / F;
}

float orc_tone_map(float c, uint32_t mode) {
    if (mode == 0) return gamma22(c / (c + 1.0f));  // defaultToneMap, :49-52
    if (mode == 1) {                                // toneMapUncharted, :25-30
        float white = 1.0f / uncharted2_impl(11.2f);
        return gamma22(uncharted2_impl(c * 2.0f) * white);
    }
    if (mode == 2) {  // toneMapHejlRichard, :34-37
        float x = std::max(0.0f, c - 0.004f);
        return (x * (6.2f * x + .5f)) / (x * (6.2f * x + 1.7f) + 0.06f);
    }
    if (mode == 3) {  // toneMapACES, :41-47
        float v = (c * (2.51f * c + 0.03f)) / (c * (2.43f * c + 0.59f) + 0.14f);
        return gamma22(std::min(std::max(v, 0.0f), 1.0f));
    }
    return gamma22(c);
}

static float smooth_step(float a, float b, float x) {  // GLSL smoothstep
    float t = std::min(std::max((x - a) / (b - a), 0.0f), 1.0f);
    return t * t * (3.0f - 2.0f * t);
}

void orc_heatmap(float t, float out[3]) {  // Heatmap.glsl:4-33
    static const float c[10][3] = {{0, 2, 91},    {0, 108, 251}, {0, 221, 221}, {51, 221, 0}, {255, 252, 0},
                                   {255, 180, 0}, {255, 104, 0}, {226, 22, 0},  {191, 0, 83}, {145, 0, 65}};
    float s = t * 10.0f;
    int cur = int(s) <= 9 ? int(s) : 9;
    int prv = cur >= 1 ? cur - 1 : 0;
    int nxt = cur < 9 ? cur + 1 : 9;
    const float blur = 0.8f;
    float wc = smooth_step(float(cur) - blur, float(cur) + blur, s) *
               (1.0f - smooth_step(float(cur + 1) - blur, float(cur + 1) + blur, s));
    float wp = 1.0f - smooth_step(float(cur) - blur, float(cur) + blur, s);
    float wn = smooth_step(float(cur + 1) - blur, float(cur + 1) + blur, s);
    for (int k = 0; k < 3; k++) {
        float r = wc * (c[cur][k] / 255.0f) + wp * (c[prv][k] / 255.0f) + wn * (c[nxt][k] / 255.0f);
        out[k] = std::min(std::max(r, 0.0f), 1.0f);
    }
}

static uint32_t to_unorm8(float v) {  // clamp, scale by 255, round to nearest even
    if (!(v > 0.0f)) return 0u;
    if (v > 1.0f) v = 1.0f;
    return uint32_t(std::nearbyint(v * 255.0f));
}

void orc_output_image(const float* accum, uint32_t n, const float* last_t, const uint64_t* ticks, const rt_ubo* ubo,
                      int bgra, uint32_t* out) {
    const float total = float(ubo->total_number_of_samples);
    for (uint32_t i = 0; i < n; i++) {
        float col[3];
        for (int k = 0; k < 3; k++) col[k] = orc_tone_map(accum[4 * size_t(i) + k] / total, ubo->tone_mapping_mode);
        if (ubo->mapping == RT_MAPPING_HEAT) {
            float scale = 1000000.0f * ubo->heatmap_scale * ubo->heatmap_scale;
            float t = ticks ? float(ticks[i]) / scale : 0.0f;
            orc_heatmap(std::min(std::max(t, 0.0f), 1.0f), col);
        } else if (ubo->mapping == RT_MAPPING_DISTANCE) {
            float t = last_t ? last_t[i] : 0.0f;
            float g = std::min((ubo->heatmap_scale - std::max(t, 0.001f)) / ubo->heatmap_scale, 1.0f);
            col[0] = col[1] = col[2] = g;
        }
        uint32_t r = to_unorm8(col[0]), g = to_unorm8(col[1]), b = to_unorm8(col[2]);
        out[i] = bgra ? (b | g << 8 | r << 16 | 255u << 24) : (r | g << 8 | b << 16 | 255u << 24);
    }
}

// ------------------------------------------------------------ skinning
// AnimationCompute.comp:14-38 in fp64: transform = sum_k weights[k] * bones[joints[k]]
// (column-major mat4), pos = T (pos, 1), normal = normalize(T (normal, 0)).xyz,
// tangent = normalize(T (tangent.xyz, 0)) with tangent.w kept.  Vertices with
// skin_index < 0 are copied; returns the number of vertices whose skin index or
// joints are out of range (copied unchanged).
uint32_t orc_skin_vertices(const rt_vertex* in, rt_vertex* out, uint32_t n, const float* skins, uint32_t num_skins) {
    uint32_t bad = 0;
    for (uint32_t i = 0; i < n; i++) {
        rt_vertex v = in[i];
        if (v.skin_index >= 0) {
            bool ok = uint32_t(v.skin_index) < num_skins;
            for (int k = 0; k < 4; k++) ok = ok && v.joints[k] < 256u;
            if (!ok) {
                bad++;
            } else {
                double T[16] = {0};
                for (int k = 0; k < 4; k++) {
                    const float* B = skins + (size_t(v.skin_index) * 256 + v.joints[k]) * 16;
                    for (int e = 0; e < 16; e++) T[e] += double(v.weights[k]) * double(B[e]);
                }
                auto apply = [&](const float* x, double w, double r[4]) {
                    for (int row = 0; row < 4; row++)
                        r[row] = T[row] * x[0] + T[4 + row] * x[1] + T[8 + row] * x[2] + T[12 + row] * w;
                };
                double p[4], nn[4], t[4];
                apply(v.pos, 1.0, p);
                apply(v.normal, 0.0, nn);
                apply(v.tangent, 0.0, t);
                const double nl = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
                const double tl = std::sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2] + t[3] * t[3]);
                for (int k = 0; k < 3; k++) {
                    v.pos[k] = float(p[k]);
                    v.normal[k] = float(nn[k] / nl);
                    v.tangent[k] = float(t[k] / tl);
                }
            }
        }
        out[i] = v;
    }
    return bad;
}

}  // extern "C"
