// brdf_dual.h — forward-mode derivatives of the closest-hit BRDF terms with respect
// to the material parameters a parameter texture drives: base colour (r, g, b),
// roughness and metallic.
//
// The values follow the reference shading (crates/examples/gltf_viewer/shaders/
// lib/PBR.glsl: matBuild :195-206, evalCombinedBRDF :731-745, the sampled weights
// of evalIndirectCombinedBRDF :649-729) for the metallic-roughness workflow.  Each
// scalar carries its derivative along the five parameters (dual numbers), so a
// vertex's Jacobian comes out of one evaluation.
//
// The sampled weights are differentiated with the sampling density held at its
// value ("detached"): weight(p) = f(p) cos / pdf(p0).  The derivative of the
// path estimate is then an unbiased estimate of the derivative of the expected
// image with the paths' lobe choices and directions fixed.
//
// Plain C++ (host and device): the CPU tests compile this header with g++.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define RT_HD __host__ __device__
#else
#define RT_HD
#endif

namespace rt {
namespace ad {

constexpr int kParams = 5;  // d/d base.r, base.g, base.b, roughness, metallic
enum : int { kBaseR = 0, kBaseG = 1, kBaseB = 2, kRough = 3, kMetal = 4 };

struct Dn {
    double v;
    double d[kParams];
};

RT_HD inline Dn cst(double v) {
    Dn r;
    r.v = v;
    for (int k = 0; k < kParams; k++) r.d[k] = 0.0;
    return r;
}
RT_HD inline Dn var(double v, int k) {
    Dn r = cst(v);
    r.d[k] = 1.0;
    return r;
}
RT_HD inline Dn operator+(Dn a, Dn b) {
    Dn r;
    r.v = a.v + b.v;
    for (int k = 0; k < kParams; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
RT_HD inline Dn operator-(Dn a, Dn b) {
    Dn r;
    r.v = a.v - b.v;
    for (int k = 0; k < kParams; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
RT_HD inline Dn operator-(Dn a) {
    Dn r;
    r.v = -a.v;
    for (int k = 0; k < kParams; k++) r.d[k] = -a.d[k];
    return r;
}
RT_HD inline Dn operator*(Dn a, Dn b) {
    Dn r;
    r.v = a.v * b.v;
    for (int k = 0; k < kParams; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
RT_HD inline Dn operator*(Dn a, double s) {
    Dn r;
    r.v = a.v * s;
    for (int k = 0; k < kParams; k++) r.d[k] = a.d[k] * s;
    return r;
}
RT_HD inline Dn operator*(double s, Dn a) { return a * s; }
RT_HD inline Dn operator+(Dn a, double s) {
    a.v += s;
    return a;
}
RT_HD inline Dn operator+(double s, Dn a) { return a + s; }
RT_HD inline Dn operator-(Dn a, double s) {
    a.v -= s;
    return a;
}
RT_HD inline Dn operator-(double s, Dn a) { return (-a) + s; }
RT_HD inline Dn operator/(Dn a, Dn b) {
    Dn r;
    r.v = a.v / b.v;
    const double ib2 = 1.0 / (b.v * b.v);
    for (int k = 0; k < kParams; k++) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) * ib2;
    return r;
}
RT_HD inline Dn operator/(Dn a, double s) { return a * (1.0 / s); }
RT_HD inline Dn operator/(double s, Dn b) { return cst(s) / b; }
RT_HD inline Dn dsqrt(Dn a) {
    Dn r;
    r.v = sqrt(a.v);
    const double h = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < kParams; k++) r.d[k] = a.d[k] * h;
    return r;
}
// min / max against a constant: the derivative of the selected operand
RT_HD inline Dn dmin(Dn a, double c) { return a.v < c ? a : cst(c); }
RT_HD inline Dn dmax(Dn a, double c) { return a.v > c ? a : cst(c); }
RT_HD inline Dn dclamp(Dn a, double lo, double hi) { return dmin(dmax(a, lo), hi); }
RT_HD inline Dn dpow5(Dn a) { return a * a * a * a * a; }

struct V3 {
    double x, y, z;
};
RT_HD inline V3 v3(double x, double y, double z) {
    V3 r{x, y, z};
    return r;
}
RT_HD inline double dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
RT_HD inline V3 norm3(V3 a) {
    double l = sqrt(dot3(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}

struct Dv3 {
    Dn x, y, z;
};
RT_HD inline Dv3 dv3(Dn x, Dn y, Dn z) {
    Dv3 r{x, y, z};
    return r;
}
RT_HD inline Dv3 dv3c(double x, double y, double z) { return dv3(cst(x), cst(y), cst(z)); }
RT_HD inline Dv3 operator+(Dv3 a, Dv3 b) { return dv3(a.x + b.x, a.y + b.y, a.z + b.z); }
RT_HD inline Dv3 operator-(Dv3 a, Dv3 b) { return dv3(a.x - b.x, a.y - b.y, a.z - b.z); }
RT_HD inline Dv3 operator*(Dv3 a, Dn s) { return dv3(a.x * s, a.y * s, a.z * s); }
RT_HD inline Dv3 operator*(Dv3 a, double s) { return dv3(a.x * s, a.y * s, a.z * s); }
RT_HD inline Dv3 mulv(Dv3 a, Dv3 b) { return dv3(a.x * b.x, a.y * b.y, a.z * b.z); }
RT_HD inline Dn lum(Dv3 c) { return c.x * 0.2126 + c.y * 0.7152 + c.z * 0.0722; }

// The material inputs of matBuild at one vertex (values of the forward pass).
struct MatIn {
    double base[3];
    double metallic, roughness;
    double ior;  // already 1 unless the volume extension is present
    double transmission, specular_factor;
    double spec_color[3];
};

// MaterialBrdf fields as functions of the five parameters.
struct MatD {
    Dv3 F0, c_diff;
    Dn f90s;  // shadowedF90(F90) = min(1, luminance(F90))
    Dn rough, alpha;
    double transmission, specular_factor;
};

// matBuild (PBR.glsl:195-206) with base colour, roughness and metallic as variables.
RT_HD inline MatD mat_dual(const MatIn& m) {
    MatD r;
    Dv3 base = dv3(var(m.base[0], kBaseR), var(m.base[1], kBaseG), var(m.base[2], kBaseB));
    Dn mt = var(m.metallic, kMetal);
    r.rough = var(m.roughness, kRough);
    r.alpha = r.rough * r.rough;
    const double f = (m.ior - 1.0) / (m.ior + 1.0);
    double df0[3], df90[3];
    for (int c = 0; c < 3; c++) {
        double d = m.spec_color[c] * (f * f);
        df0[c] = (d < 1.0 ? d : 1.0) * m.specular_factor;
        df90[c] = m.spec_color[c];
    }
    Dn one_m = 1.0 - mt;
    r.F0 = dv3(one_m * df0[0] + base.x * mt, one_m * df0[1] + base.y * mt, one_m * df0[2] + base.z * mt);
    Dv3 F90 = dv3(one_m * df90[0] + mt, one_m * df90[1] + mt, one_m * df90[2] + mt);
    r.f90s = dmin(lum(F90), 1.0);
    r.c_diff = base * one_m;
    r.transmission = m.transmission;
    r.specular_factor = m.specular_factor;
    return r;
}

// evalFresnelSchlick (PBR.glsl:228-231)
RT_HD inline Dv3 fresnel_d(Dv3 f0, Dn f90, double n_dot_s) {
    double p = 1.0 - n_dot_s;
    p = p * p * p * p * p;
    return dv3(f0.x + (f90 - f0.x) * p, f0.y + (f90 - f0.y) * p, f0.z + (f90 - f0.z) * p);
}
// Smith_G_a, Smith_G_Lambda_GGX, Smith_G2_Height_Correlated (PBR.glsl:248-262)
RT_HD inline Dn smith_lambda_d(Dn alpha, double nds) {
    Dn a = nds / (dmax(alpha, 0.00001) * sqrt(1.0 - fmin(0.99999, nds * nds)));
    return (dsqrt(1.0 + 1.0 / (a * a)) - 1.0) * 0.5;
}
RT_HD inline Dn smith_g2_d(Dn alpha, double ndl, double ndv) {
    return 1.0 / (1.0 + smith_lambda_d(alpha, ndl) + smith_lambda_d(alpha, ndv));
}
// GGX_D (PBR.glsl:286-289), pi as the shaders define it (PBR.glsl:57)
RT_HD inline Dn ggx_d_d(Dn a2, double ndh) {
    Dn b = (a2 - 1.0) * (ndh * ndh) + 1.0;
    return a2 / (b * b * 3.141592653589);
}
// Smith_G1_GGX (PBR.glsl:446-448)
RT_HD inline Dn smith_g1_d(Dn a2, double nds2) {
    return 2.0 / (dsqrt((a2 * (1.0 - nds2) + nds2) / nds2) + 1.0);
}
// frostbiteDisneyDiffuse (PBR.glsl:486-496)
RT_HD inline Dn frostbite_d(Dn r, double ndl, double ndv, double ldh) {
    Dn bias = r * 0.5;
    Dn factor = (1.0 - r) * 1.0 + r * (1.0 / 1.51);
    Dn fd90m1 = bias + r * (2.0 * ldh * ldh) - 1.0;
    double pl = 1.0 - ndl, pv = 1.0 - ndv;
    pl = pl * pl * pl * pl * pl;
    pv = pv * pv * pv * pv * pv;
    return (1.0 + fd90m1 * pl) * (1.0 + fd90m1 * pv) * factor;
}

// prepareBRDFData's dot products (PBR.glsl:527-564)
struct Dots {
    double ndl, ndv, ldh, ndh, vdh;
    bool back;
};
RT_HD inline Dots dots(V3 N, V3 L, V3 V) {
    Dots d;
    V3 H = norm3(v3(L.x + V.x, L.y + V.y, L.z + V.z));
    double ndl = dot3(N, L), ndv = dot3(N, V);
    d.back = ndv <= 0.0 || ndl <= 0.0;
    d.ndl = fmin(fmax(0.00001, ndl), 1.0);
    d.ndv = fmin(fmax(0.00001, ndv), 1.0);
    d.ldh = fmin(fmax(dot3(L, H), 0.0), 1.0);
    d.ndh = fmin(fmax(dot3(N, H), 0.0), 1.0);
    d.vdh = fmin(fmax(dot3(V, H), 0.0), 1.0);
    return d;
}

// evalCombinedBRDF (PBR.glsl:731-745): the NEE factor.
RT_HD inline Dv3 eval_brdf_d(V3 N, V3 L, V3 V, const MatD& m) {
    Dots d = dots(N, L, V);
    if (d.back) return dv3c(0.0, 0.0, 0.0);
    Dv3 F = fresnel_d(m.F0, m.f90s, d.vdh);
    Dn D = ggx_d_d(dmax(m.alpha * m.alpha, 0.00001), d.ndh);
    Dn G2 = smith_g2_d(m.alpha, d.ndl, d.ndv);
    Dn s = G2 * D * (d.ndl / (4.0 * d.ndl * d.ndv));
    Dv3 spec = F * s;
    Dn fb = frostbite_d(m.rough, d.ndl, d.ndv, d.ldh) * (d.ndl / 3.141592653589);
    return m.c_diff * fb + spec;
}

enum : int { kLobeDiffuse = 1, kLobeSpecular = 2, kLobeTransmission = 3 };

// The sampled throughput factor of evalIndirectCombinedBRDF for a direction L
// already drawn (lobe `type`, chosen with probability p_lobe), with the sampling
// density held at its value: specular F D G2 / (G1 D)|p0 / p_lobe, diffuse
// (1 - F s) c_diff Fd (1 - t) / p_lobe, transmission c_diff t / p_lobe.
// `delta` marks the metallic 1 / roughness 0 mirror (no lobe draw: p_lobe = 1).
RT_HD inline Dv3 sample_weight_d(int type, V3 N, V3 L, V3 V, const MatD& m, double p_lobe, bool delta) {
    if (type == kLobeSpecular) {
        V3 H = norm3(v3(L.x + V.x, L.y + V.y, L.z + V.z));
        double hdl = fmax(0.00001, fmin(1.0, dot3(H, L)));
        double ndl = fmax(0.00001, fmin(1.0, dot3(N, L)));
        double ndv = fmax(0.00001, fmin(1.0, dot3(N, V)));
        Dv3 F = fresnel_d(m.F0, m.f90s, hdl);
        if (delta || m.alpha.v == 0.0) return F * (m.specular_factor / p_lobe);
        // f cos / pdf = F D G2 / (G1(v) D) for the VNDF density G1(v) D / (4 n.v);
        // the density's G1(v) D is detached.  G2 = G1(l) G1(v) / (G1(l) + G1(v) -
        // G1(l) G1(v)), the form whose ratio to G1(v) the shader evaluates.
        Dn a2 = m.alpha * m.alpha;
        Dn g1v = smith_g1_d(a2, ndv * ndv), g1l = smith_g1_d(a2, ndl * ndl);
        Dn g2 = g1l * g1v / (g1v + g1l - g1v * g1l);
        Dn D = ggx_d_d(a2, fmin(fmax(dot3(N, H), 0.0), 1.0));
        Dn w = g2 * D * (m.specular_factor / (g1v.v * D.v * p_lobe));
        return F * w;
    }
    if (type == kLobeDiffuse) {
        Dots d = dots(N, L, V);
        Dv3 F = fresnel_d(m.F0, m.f90s, d.vdh);
        Dn fb = frostbite_d(m.rough, d.ndl, d.ndv, d.ldh);
        Dv3 one_f = dv3(1.0 - F.x * m.specular_factor, 1.0 - F.y * m.specular_factor, 1.0 - F.z * m.specular_factor);
        return mulv(one_f, m.c_diff) * (fb * ((1.0 - m.transmission) / p_lobe));
    }
    Dv3 w = m.c_diff * (m.transmission / p_lobe);
    return dv3(dmax(w.x, 0.0), dmax(w.y, 0.0), dmax(w.z, 0.0));
}

}  // namespace ad
}  // namespace rt
