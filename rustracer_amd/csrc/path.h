// path.h — the per-pixel integrator of the reference raygen shader and the hit
// shaders it triggers, as device functions shared by the megakernel and the
// wavefront kernels.
//
// Restates crates/examples/gltf_viewer/shaders/RayTracing.rgen:25-167 (sample and
// bounce loops, Russian roulette, accumulation), RayTracing.rchit:136-477
// (closest hit), RayTracing.rmiss:14-45 (miss) and RayTracing.rchit:35-59,
// RayTracing.shadow.rahit, RayTracing.shadow.rmiss (shadow ray).
//
// Precision: the path state (ray origin and direction, throughput, radiance) and
// all shading are fp64; traversal receives the rays rounded to fp32 and returns the
// closest triangle, whose t and barycentrics are then recomputed in fp64 from the
// fp64 ray (refine_hit).
#pragma once

#include "render.h"

namespace rt {

constexpr double kTMin = 0.001;     // lib/Camera.glsl:2
constexpr double kTMax = 10000.0;   // lib/Camera.glsl:3
constexpr double kShadowTMin = 0.1; // RayTracing.rchit:43
constexpr uint32_t kMinBounces = 3; // RayTracing.rgen:15

struct PathCounters {
    uint32_t primary = 0, extension = 0, shadow = 0, shade = 0, tex = 0, lights = 0, bad = 0;
    TraceStats tr;   // closest-hit traversal
    TraceStats trs;  // shadow traversal
};

__device__ inline void flush_counters(unsigned long long* c, const PathCounters& pc) {
    uint32_t v[kNumCounters] = {pc.primary, pc.extension, pc.shadow, pc.tr.nodes + pc.trs.nodes,
                                pc.tr.tris + pc.trs.tris, pc.tr.instances + pc.trs.instances, pc.shade,
                                pc.tr.anyhits + pc.trs.anyhits, pc.tex, pc.lights,
                                pc.tr.overflows + pc.trs.overflows, pc.bad, pc.trs.nodes, pc.trs.tris,
                                pc.trs.instances, 0, 0};
#pragma unroll
    for (int k = 0; k < kNumCounters; k++) {
        unsigned long long s = wave_sum<unsigned long long>(v[k]);
        if (__lane_id() == 0 && s) atomicAdd(&c[k], s);
    }
}

// Any-hit functor for the traversal: the alpha test on the candidate's material.
struct AlphaAnyHit {
    const SceneView* sv;
    const TlasView* tl;
    Rng4* rng;
    __device__ bool ignore(uint32_t inst, uint32_t prim, float u, float v) {
        return alpha_test_ignores(*sv, tl->inst[inst].custom_index, prim, u, v, *rng);
    }
};

// column-major mat4 * (a, b, c, w), xyz part
__device__ inline D3 mat4_mul(const float* m, double a, double b, double c, double w) {
    return d3(m[0] * a + m[4] * b + m[8] * c + m[12] * w, m[1] * a + m[5] * b + m[9] * c + m[13] * w,
              m[2] * a + m[6] * b + m[10] * c + m[14] * w);
}

// Camera ray of one sample, RayTracing.rgen:52-75.
__device__ inline void camera_ray(const FrameArgs& fa, uint32_t x, uint32_t y, Rng4& rng_a, uint32_t& lcg, D3& o,
                                  D3& d, double& tfar) {
    double px = double(x) + 0.5, py = double(y) + 0.5;
    if (fa.antialiasing) {
        double jx = rng_next(rng_a);
        double jy = rng_next(rng_a);
        px += jx - 0.5;
        py += jy - 0.5;
    }
    double uvx = (px / double(fa.width)) * 2.0 - 1.0;
    double uvy = (py / double(fa.height)) * 2.0 - 1.0;
    double offx = 0.0, offy = 0.0;
    // RandomInUnitDisk (lib/Random.glsl:44-54).  With aperture 0 the offset is 0 and the
    // draws are skipped, unless debug == 1 shading reads the LCG after them.
    if (fa.aperture != 0.f || fa.debug) {
        float dx, dy;
        for (;;) {
            float a = lcg_float(lcg), b = lcg_float(lcg);
            dx = 2.f * a - 1.f;
            dy = 2.f * b - 1.f;
            if (dx * dx + dy * dy < 1.f) break;
        }
        offx = double(fa.aperture) / 2.0 * dx;
        offy = double(fa.aperture) / 2.0 * dy;
    }
    o = mat4_mul(fa.mvi, offx, offy, 0.0, 1.0);
    D3 tg = mat4_mul(fa.pin, uvx, uvy, 1.0, 1.0);
    D3 td = dnormalize(tg * double(fa.focus_distance) - d3(offx, offy, 0.0));
    d = mat4_mul(fa.mvi, td.x, td.y, td.z, 0.0);
    tfar = kTMax;
    if (fa.ortho > 0.f) {  // rgen:67-75
        double k = 1.0 + fa.ortho;
        o = mat4_mul(fa.mvi, k * uvx, -(k * uvy), 0.0, 1.0);
        d = mat4_mul(fa.mvi, 0.0, 0.0, -1.0, 0.0);
        tfar = 10.0 * kTMax;
    }
}

// t and barycentrics of the hit triangle recomputed in fp64 from the fp64 ray
// (Woop, Benthin & Wald 2013 without the inside test: the triangle is known).
__device__ inline void refine_hit(const TlasView& tl, const HitInfo& h, D3 o, D3 d, double& t, double& u, double& v) {
    const InstanceRecord& ir = tl.inst[h.inst];
    const double* m = ir.w2o_d;
    D3 oo = d3(m[0] * o.x + m[1] * o.y + m[2] * o.z + m[3], m[4] * o.x + m[5] * o.y + m[6] * o.z + m[7],
               m[8] * o.x + m[9] * o.y + m[10] * o.z + m[11]);
    D3 od = d3(m[0] * d.x + m[1] * d.y + m[2] * d.z, m[4] * d.x + m[5] * d.y + m[6] * d.z,
               m[8] * d.x + m[9] * d.y + m[10] * d.z);
    const float4* tp = ir.tris + 3 * size_t(h.rec);
    float4 fa = ldg_f4(tp), fb = ldg_f4(tp + 1), fc = ldg_f4(tp + 2);
    double ax = fabs(od.x), ay = fabs(od.y), az = fabs(od.z);
    int kz = (ax > ay) ? (ax > az ? 0 : 2) : (ay > az ? 1 : 2);
    int kx = kz == 2 ? 0 : kz + 1;
    int ky = kx == 2 ? 0 : kx + 1;
    if (dcomp(od, kz) < 0.0) { int tt = kx; kx = ky; ky = tt; }
    double dz = dcomp(od, kz);
    double Sx = dcomp(od, kx) / dz, Sy = dcomp(od, ky) / dz, Sz = 1.0 / dz;
    D3 A = d3(fa.x, fa.y, fa.z) - oo, B = d3(fb.x, fb.y, fb.z) - oo, C = d3(fc.x, fc.y, fc.z) - oo;
    double Ax = dcomp(A, kx) - Sx * dcomp(A, kz), Ay = dcomp(A, ky) - Sy * dcomp(A, kz);
    double Bx = dcomp(B, kx) - Sx * dcomp(B, kz), By = dcomp(B, ky) - Sy * dcomp(B, kz);
    double Cx = dcomp(C, kx) - Sx * dcomp(C, kz), Cy = dcomp(C, ky) - Sy * dcomp(C, kz);
    double U = Cx * By - Cy * Bx, V = Ax * Cy - Ay * Cx, W = Bx * Ay - By * Ax;
    double det = U + V + W;
    if (det == 0.0) {
        t = h.t;
        u = h.u;
        v = h.v;
        return;
    }
    double T = U * (Sz * dcomp(A, kz)) + V * (Sz * dcomp(B, kz)) + W * (Sz * dcomp(C, kz));
    double inv = 1.0 / det;
    t = T * inv;
    u = V * inv;
    v = W * inv;
}

// RayPayload, lib/RayTracingCommons.glsl:11-24
struct Payload {
    D3 hit_value, hit_point, scatter_dir, emittance;
    double t;
    bool need_scatter;
    Rng4 rng;
    double volume_dis;
    uint32_t lcg;  // Ray.RandomSeed (the LCG of RandomInUnitDisk / RandomInUnitSphere)
};

// RayTracing.rmiss:14-45
__device__ inline void miss_shader(const SceneView& sv, Payload& ray, D3 dir) {
    D3 acc = d3(0.0);
    D3 rd = dnormalize(dir);
    if (ray.t != 0.0) {
        for (uint32_t i = 0; i < sv.num_dlights; i++) {
            const rt_light& li = sv.dlights[i];
            double c = ddot(dnormalize(ld3(li.transform)), rd);
            if (c < 0.0) acc = acc + ld3(li.color) * (-c * double(li.intensity));
        }
    }
    if (sv.has_sky) {
        // rmiss:27-33: light_acc += skyColor + light_acc
        const D4 sky = sample_cube(sv.sky, sv.srgb_lut, rd);
        acc = acc + (xyz(sky) + acc);
    } else {
        acc = acc + d3(0.01);
    }
    ray.hit_value = d3(0.0);
    ray.need_scatter = false;
    ray.emittance = sv.num_dlights == 0 ? d3(0.0) : acc;
    ray.t = -1.0;
}

__device__ inline D3 o2w_dir(const InstanceRecord& ir, D3 v) {
    float4 r0 = ir.o2w[0], r1 = ir.o2w[1], r2 = ir.o2w[2];
    return d3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
              r2.x * v.x + r2.y * v.y + r2.z * v.z);
}
__device__ inline D3 o2w_point(const InstanceRecord& ir, D3 v) {
    float4 r0 = ir.o2w[0], r1 = ir.o2w[1], r2 = ir.o2w[2];
    return d3(r0.x * v.x + r0.y * v.y + r0.z * v.z + r0.w, r1.x * v.x + r1.y * v.y + r1.z * v.z + r1.w,
              r2.x * v.x + r2.y * v.y + r2.z * v.z + r2.w);
}

// Result of the closest-hit shader up to the point where it casts its shadow ray.
struct ShadeState {
    D3 origin;        // world hit point
    D3 geo_normal;    // geometric normal facing V
    D3 outward;       // shading normal on V's side of the geometric normal
    D3 V;
    MatBrdf m;
    bool front;
    bool has_light;
    D3 light_dir;
    double light_dist;
    D3 light_scale;   // light_weight * intensity * color
    double light_w;   // light_weight (the RIS weight, rchit:119)
    uint32_t light_index;  // the chosen point light
    Rng4 rng;         // local stream after the RIS draws
    double displacement;
    // read by the adjoint only (dead in the forward kernels)
    const rt_material_raw* mat;
    D4 uv, vcol;
    uint32_t lobe;    // kDiffuse / kSpecular / kTransmission chosen by shade_end
    double p_lobe;    // its probability (1 for the metallic-1 roughness-0 mirror)
    bool delta;       // the metallic-1 roughness-0 mirror (no lobe draw)
    // inputs of the debug == 1 scatter (rchit:437-475) beyond the material
    bool normal_flipped;  // outward = -normal
    bool emit_small;      // length(emittance) < 0.01 (before exposure)
    float ior_raw;        // MaterialRaw.ior
};

__device__ inline D4 mix4(D4 a, D4 b, D4 c, double b0, double bu, double bv) {
    return D4{a.x * b0 + b.x * bu + c.x * bv, a.y * b0 + b.y * bu + c.y * bv, a.z * b0 + b.z * bu + c.z * bv,
              a.w * b0 + b.w * bu + c.w * bv};
}

// First half of RayTracing.rchit main (:136-340): material resolution and the RIS
// light choice.  Returns false when a debug mapping or an invalid reference ended
// the shader (ray already updated).
template <bool kStats>
__device__ inline bool shade_begin(const FrameArgs& fa, Payload& ray, const HitInfo& h, double ht, double bu,
                                   double bv, D3 world_dir, ShadeState& s, PathCounters& pc) {
    const SceneView& sv = fa.sv;
    const InstanceRecord& ir = fa.tlas.inst[h.inst];
    if (kStats) pc.shade++;
    uint32_t ci = ir.custom_index;
    ray.t = ht;
    ray.need_scatter = false;
    if (ci >= sv.num_prim_infos) { pc.bad++; ray.emittance = d3(0.0); return false; }
    const rt_prim_info pi = sv.prim_infos[ci];
    if (pi.material_id >= sv.num_materials) { pc.bad++; ray.emittance = d3(0.0); return false; }
    const rt_material_raw& mat = sv.materials[pi.material_id];
    D3 last_hit = ray.hit_point;
    uint32_t io = pi.i_offset + 3u * h.prim;
    if (io + 2u >= sv.num_indices) { pc.bad++; ray.emittance = d3(0.0); return false; }
    uint32_t i0 = pi.v_offset + sv.indices[io], i1 = pi.v_offset + sv.indices[io + 1], i2 = pi.v_offset + sv.indices[io + 2];
    if (i0 >= sv.num_vertices || i1 >= sv.num_vertices || i2 >= sv.num_vertices) {
        pc.bad++;
        ray.emittance = d3(0.0);
        return false;
    }
    const float* va = reinterpret_cast<const float*>(sv.vertices + i0);
    const float* vb = reinterpret_cast<const float*>(sv.vertices + i1);
    const float* vc = reinterpret_cast<const float*>(sv.vertices + i2);
    // getMixVertexAndGeoNormal, lib/RayTracingCommons.glsl:86-101
    const double b0 = 1.0 - bu - bv;
    D3 p0 = ld3(va), p1 = ld3(vb), p2 = ld3(vc);
    D4 uv = mix4(ld4(va + 24), ld4(vb + 24), ld4(vc + 24), b0, bu, bv);
    D3 pos = p0 * b0 + p1 * bu + p2 * bv;
    D4 vcol = mix4(ld4(va + 12), ld4(vb + 12), ld4(vc + 12), b0, bu, bv);
    D3 normal = dnormalize(dnormalize(ld3(va + 4)) * b0 + dnormalize(ld3(vb + 4)) * bu + dnormalize(ld3(vc + 4)) * bv);
    D3 geo_normal = dnormalize(o2w_dir(ir, dcross(p1 - p0, p2 - p0)));
    D3 origin = o2w_point(ir, pos);
    D4 color4 = mul4(vcol, ld4(mat.base_color));
    if (mat.base_color_texture.index >= 0) {
        color4 = mul4(color4, tex_uv(sv, mat.base_color_texture, uv));
        if (kStats) pc.tex++;
    }
    D3 color = xyz(color4);
    ray.hit_point = pos;
    if (mat.normal_texture.index >= 0) {  // rchit:178-185, getNormal :124-129
        D4 tx = tex_uv(sv, mat.normal_texture, uv);
        if (kStats) pc.tex++;
        D3 nt = dnormalize(d3(tx.x * 2.0 - 1.0, tx.y * 2.0 - 1.0, tx.z * 2.0 - 1.0));
        D4 t4 = mix4(ld4(va + 8), ld4(vb + 8), ld4(vc + 8), b0, bu, bv);
        double tl = sqrt(t4.x * t4.x + t4.y * t4.y + t4.z * t4.z + t4.w * t4.w);
        t4 = D4{t4.x / tl, t4.y / tl, t4.z / tl, t4.w / tl};
        D3 tm = xyz(t4);
        D3 tangent = dnormalize(tm - normal * ddot(tm, normal));
        D3 bit = dnormalize(dcross(normal, tangent) * t4.w);
        normal = tangent * nt.x + bit * nt.y + normal * nt.z;
    }
    normal = dnormalize(o2w_dir(ir, normal));
    D3 V = -dnormalize(world_dir);
    bool front = ddot(V, geo_normal) >= 0.0;
    geo_normal = front ? geo_normal : -geo_normal;
    D3 outward = ddot(geo_normal, normal) < 0.0 ? -normal : normal;
    D3 emit = ld3(mat.emissive_factor);
    if (mat.emissive_texture.index >= 0) {
        emit = emit * xyz(tex_uv(sv, mat.emissive_texture, uv));
        if (kStats) pc.tex++;
    }
    double metallic = mat.metallic_factor, roughness = mat.roughness_factor;
    if (mat.metallic_roughness_texture.index >= 0) {
        D4 tx = tex_uv(sv, mat.metallic_roughness_texture, uv);
        if (kStats) pc.tex++;
        roughness *= tx.y;
        metallic *= tx.z;
    }
    D3 spec_wf = d3(1.0);
    const bool sg = mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS;
    if (sg) {  // rchit:215-239
        D4 df = ld4(mat.sg_diffuse_factor);
        D4 sgf = ld4(mat.sg_specular_glossiness_factor);
        if (mat.sg_diffuse_texture.index >= 0) { df = mul4(df, tex_uv(sv, mat.sg_diffuse_texture, uv)); if (kStats) pc.tex++; }
        if (mat.sg_specular_glossiness_texture.index >= 0) {
            sgf = mul4(sgf, tex_uv(sv, mat.sg_specular_glossiness_texture, uv));
            if (kStats) pc.tex++;
        }
        spec_wf = xyz(sgf);
        roughness = 1.0 - sgf.w;
        double mx = fmax(fmax(spec_wf.x, spec_wf.y), spec_wf.z);
        color = xyz(vcol) * xyz(df);
        metallic = convert_metallic(color, spec_wf, mx);
    }
    double trans = 0.0;
    if (mat.transmission_exist) {
        trans = mat.transmission_factor;
        if (mat.transmission_texture.index >= 0) {
            trans *= tex_uv(sv, mat.transmission_texture, uv).x;
            if (kStats) pc.tex++;
        }
    }
    uint32_t mapping = mat.unlit ? uint32_t(RT_MAPPING_ALBEDO) : fa.mapping;
    if (mapping != RT_MAPPING_RENDER && mapping != RT_MAPPING_HEAT && mapping != RT_MAPPING_DISTANCE) {
        // debug mappings, rchit:258-286 (hitValue keeps its previous value, needScatter is false)
        D3 e = d3(0.0);
        switch (mapping) {
            case RT_MAPPING_ALBEDO: e = color; break;
            case RT_MAPPING_TRIANGLE: e = d3(b0, bu, bv); break;
            case RT_MAPPING_METALLIC: e = d3(metallic); break;
            case RT_MAPPING_ROUGHNESS: e = d3(roughness); break;
            case RT_MAPPING_NORMAL: e = (outward + d3(1.0)) / 2.0; break;
            case RT_MAPPING_TRANSMISSION: e = d3(trans); break;
            case RT_MAPPING_INSTANCE:
            case RT_MAPPING_GEO_ID: {
                uint32_t k = mapping == RT_MAPPING_INSTANCE ? ir.instance_id : ci;
                k += k << 10; k ^= k >> 6; k += k << 3; k ^= k >> 11; k += k << 15;  // jenkinsHash
                e = d3((k & 255u) / 255.0, ((k >> 8) & 255u) / 255.0, ((k >> 16) & 255u) / 255.0);
                break;
            }
            case RT_MAPPING_TANGENT: {
                D4 t4 = mix4(ld4(va + 8), ld4(vb + 8), ld4(vc + 8), b0, bu, bv);
                double tl = sqrt(t4.x * t4.x + t4.y * t4.y + t4.z * t4.z + t4.w * t4.w);
                e = (dnormalize(o2w_dir(ir, xyz(t4) / tl)) + d3(1.0)) / 2.0;
                break;
            }
            default: break;
        }
        ray.emittance = e;
        return false;
    }
    double spec_factor = mat.specular_factor;
    D3 spec_color = ld3(mat.specular_color_factor);
    if (mat.specular_texture.index >= 0) { spec_factor *= tex_uv(sv, mat.specular_texture, uv).w; if (kStats) pc.tex++; }
    if (mat.specular_color_texture.index >= 0) {
        spec_color = spec_color * xyz(tex_uv(sv, mat.specular_color_texture, uv));
        if (kStats) pc.tex++;
    }
    ray.hit_point = origin;
    ray.emittance = emit * double(fa.exposure);
    MatBrdf& m = s.m;
    m.base_color = color;
    m.metallic = metallic;
    m.roughness = roughness;
    m.ior = mat.volume_exists ? double(mat.ior) : 1.0;
    m.transmission = trans;
    m.specular_factor = spec_factor;
    m.specular_color_factor = spec_color;
    m.front_face = front;
    mat_build(m);
    if (sg) {
        double mx = fmax(fmax(spec_wf.x, spec_wf.y), spec_wf.z);
        m.c_diff = color * (1.0 - mx);
        m.F0 = spec_wf;
    }
    m.attenuation_color = ld3(mat.attenuation_color);
    m.attenuation_distance = mat.attenuation_distance;
    m.volume = mat.volume_exists != 0;
    s.displacement = dlength(origin - last_hit);
    s.mat = &mat;
    s.uv = uv;
    s.vcol = vcol;
    s.origin = origin;
    s.geo_normal = geo_normal;
    s.outward = outward;
    s.V = V;
    s.front = front;
    s.normal_flipped = ddot(geo_normal, normal) < 0.0;
    s.emit_small = dlength(emit) < 0.01;
    s.ior_raw = mat.ior;
    // sampleLightRIS, rchit:77-122 (uniform candidate index, Yuksel-attenuated weights)
    Rng4 rng = ray.rng;
    s.has_light = false;
    uint32_t num = sv.num_plights;
    if (num > 0) {
        double total = 0.0, sel_pdf = 0.0;
        uint32_t sel = 0;
        uint32_t cand = num < 3u ? num : 3u;
        for (uint32_t i = 0; i < cand; i++) {
            const rt_light& li = sv.plights[i];
            if (kStats) pc.lights++;
            if (luminance(ld3(li.color) * double(li.intensity)) < 0.1) continue;
            uint32_t idx = min(num - 1u, uint32_t(double(rng_next(rng)) * double(num)));
            const rt_light& cl = sv.plights[idx];
            if (kStats) pc.lights++;
            D3 lv = ld3(cl.transform) - origin;
            double ld = dlength(lv);
            D3 L = dnormalize(lv);
            if (ddot(geo_normal, L) < 0.00001) continue;
            double pdf = luminance(light_intensity_at(cl, ld));
            double ris = pdf * double(num);
            total += ris;
            if (double(rng_next(rng)) < ris / total) {
                sel = idx;
                sel_pdf = pdf;
            }
        }
        if (total != 0.0) {
            const rt_light& l = sv.plights[sel];
            double w = (total / 3.0) / sel_pdf;
            D3 lv = ld3(l.transform) - origin;
            s.light_dist = dlength(lv);
            s.light_dir = dnormalize(lv);
            s.light_scale = ld3(l.color) * (w * double(l.intensity));
            s.light_w = w;
            s.light_index = sel;
            s.has_light = true;
        }
    }
    s.rng = rng;
    return true;
}

// Second half of RayTracing.rchit main (:352-477): lobe choice and the next
// direction.  The NEE term (rchit:346-349) is evaluated by shade_bounce, because
// whether it is added depends on the shadow ray, which the wavefront form traces
// in a separate kernel.
__device__ inline void shade_end(Payload& ray, ShadeState& s) {
    const MatBrdf& m = s.m;
    Rng4 rng = s.rng;
    uint32_t type;
    D3 thr = d3(1.0);
    s.delta = m.metallic == 1.0 && m.roughness == 0.0;
    s.p_lobe = 1.0;
    if (s.delta) {
        type = kSpecular;
    } else {
        D3 p = brdf_probability(m, s.V, s.outward);
        double r = rng_next(rng);
        if (r < p.x) {
            type = kSpecular;
            thr = thr / p.x;
            s.p_lobe = p.x;
            if (ray.volume_dis >= 0.0) ray.volume_dis += s.displacement;
        } else if (r >= p.x && r <= p.x + p.y) {
            type = kDiffuse;
            thr = thr / p.y;
            s.p_lobe = p.y;
            if (ray.volume_dis >= 0.0) ray.volume_dis += s.displacement;
        } else {
            type = kTransmission;
            if (m.volume) {
                if (ray.volume_dis >= 0.0) ray.volume_dis += s.displacement;
                else if (s.front) ray.volume_dis = 0.0;
            }
            thr = thr / p.z;
            s.p_lobe = p.z;
        }
    }
    s.lobe = type;
    D3 origin = s.origin;
    if (type == kTransmission) origin = offset_ray(origin, -s.geo_normal);
    double ux = rng_next(rng);
    double uy = rng_next(rng);
    D3 dir = d3(0.0), w = d3(0.0);
    ray.need_scatter = sample_indirect(ux, uy, s.outward, s.geo_normal, s.V, m, type, dir, w, ray.volume_dis);
    ray.hit_point = origin;
    ray.scatter_dir = dir;
    ray.hit_value = thr * w;
    ray.rng = rng;
}

// ubo.debug == 1, rchit:437-475: the legacy materials replace the scatter (hitValue =
// colour when the ray scatters), drawing from the LCG; `dir` is the unnormalised
// gl_WorldRayDirectionEXT.
__device__ inline void debug_scatter(Payload& ray, const ShadeState& s, D3 dir) {
    const MatBrdf& m = s.m;
    const D3 normal = s.normal_flipped ? -s.outward : s.outward;
    bool scatter;
    D3 next;
    if (m.transmission > 0.0) {  // dielectric: Schlick reflection or refraction
        const double ior = s.ior_raw;
        const double ratio = s.front ? 1.0 / ior : ior;
        const double cos_theta = fabs(ddot(s.V, s.geo_normal));
        const D3 refracted = refract3(dir, s.outward, ratio);
        const bool tir = refracted.x == 0.0 && refracted.y == 0.0 && refracted.z == 0.0;
        const double reflect_prob = tir ? 1.0 : schlick(cos_theta, ratio);
        scatter = true;
        next = double(lcg_float(ray.lcg)) < reflect_prob ? reflect3(dir, normal) : refracted;
    } else if (s.emit_small && m.roughness == 1.0) {  // Lambertian
        scatter = ddot(dir, s.geo_normal) < 0.0;
        next = dnormalize(s.outward + lcg_unit_sphere(ray.lcg));
    } else if (m.metallic > 0.0) {  // fuzzy metal
        const D3 reflected = reflect3(dir, s.outward);
        scatter = ddot(reflected, s.geo_normal) > 0.0;
        next = reflected + lcg_unit_sphere(ray.lcg) * 0.08;
    } else {
        return;
    }
    ray.need_scatter = scatter;
    ray.hit_value = scatter ? m.base_color : d3(0.0);
    ray.scatter_dir = next;
}

// Integrator state of one pixel between bounces (registers in the megakernel,
// structure-of-arrays in HBM in the wavefront form).  Streams A and B keep only
// their counters: their other words are (x, y, frame_count) and (x, y, seed_b).
struct PathState {
    D3 o, d;            // ray of the current bounce
    D3 thr, rad;        // throughput, radiance summed over the pixel's samples
    D3 hit_point;       // Ray.hitPoint (persists across samples, rchit:142)
    double volume_dis;  // Ray.volume_dis
    double tfar;
    uint32_t wa, wb, lcg, sample, bounce;
    float last_t;       // payload t after the last trace (Ray.t, read by the DISTANCE mapping)
};

__device__ inline void init_path(const FrameArgs& fa, uint32_t x, uint32_t y, PathState& p) {
    p.wa = 0u;
    p.wb = 0u;
    p.lcg = tea16(tea16(x, y), fa.seed_b);  // Ray.RandomSeed, rgen:33
    p.rad = d3(0.0);
    p.hit_point = d3(0.0);
    p.volume_dis = -1.0;
    p.sample = 0u;
    p.last_t = 0.f;
}

// Start of one sample, rgen:52-80.
__device__ inline void begin_sample(const FrameArgs& fa, uint32_t x, uint32_t y, PathState& p) {
    Rng4 ra{x, y, fa.frame_count, p.wa};
    camera_ray(fa, x, y, ra, p.lcg, p.o, p.d, p.tfar);
    p.wa = ra.w;
    p.thr = d3(1.0);
    p.volume_dis = -1.0;
    p.bounce = 0u;
    p.last_t = 0.f;  // Ray.t = 0, rgen:79
}

// What one traced bounce contributes, and the shadow ray it needs.
struct BounceOut {
    D3 pend_with;     // throughput * (emittance + NEE): added when the shadow ray is unoccluded
    D3 pend_without;  // throughput * emittance
    D3 so, sd;        // shadow ray origin and direction
    double sdist;
    uint32_t shadow_wb;  // stream-B counter of the shadow payload's copy (rchit:45)
    bool shadow, cont;
};

// Closest-hit observer of shade_bounce: vertex(state, payload, brdf, nee) sees each
// shaded vertex after its scatter, with the NEE BRDF value and term (zero without a
// light).  The adjoint's replay differentiates there; the forward kernels pass this.
struct NoTap {
    __device__ void vertex(const ShadeState&, const Payload&, D3, D3) {}
};

// The shaders for the traced result of p's current ray (closest-hit with NEE set-up
// or miss), then the rest of the raygen bounce loop (rgen:98-128): Russian
// roulette, throughput update, next ray.
template <bool kStats, class Tap>
__device__ inline void shade_bounce(const FrameArgs& fa, uint32_t x, uint32_t y, PathState& p, bool found,
                                    const HitInfo& h, BounceOut& r, PathCounters& pc, Tap& tap) {
    Payload ray;
    ray.hit_point = p.hit_point;
    ray.t = p.bounce == 0u ? 0.0 : 1.0;  // only "Ray.t != 0" is read before it is set (rmiss:18)
    ray.rng = Rng4{x, y, fa.seed_b, p.wb};
    ray.lcg = p.lcg;
    ray.volume_dis = p.volume_dis;
    ray.emittance = d3(0.0);
    ray.need_scatter = false;
    ray.hit_value = d3(0.0);
    ray.scatter_dir = d3(0.0);
    D3 nee = d3(0.0), brdf = d3(0.0);
    r.shadow = false;
    if (found) {
        double ht, bu, bv;
        refine_hit(fa.tlas, h, p.o, p.d, ht, bu, bv);
        ShadeState ss;
        if (shade_begin<kStats>(fa, ray, h, ht, bu, bv, p.d, ss, pc)) {
            if (ss.has_light) {
                r.shadow = true;
                r.so = ss.origin;
                r.sd = ss.light_dir;
                r.sdist = ss.light_dist;
                r.shadow_wb = ray.rng.w;
                brdf = eval_combined_brdf(ss.outward, ss.light_dir, ss.V, ss.m);
                nee = brdf * ss.light_scale;
            }
            shade_end(ray, ss);
            // the counting kernels carry the debug == 1 materials (the host runs them
            // for such frames), so the timed kernels keep their registers
            if (kStats && fa.debug) debug_scatter(ray, ss, p.d);
            tap.vertex(ss, ray, brdf, nee);
        }
    } else {
        miss_shader(fa.sv, ray, p.d);
    }
    r.pend_without = p.thr * ray.emittance;
    r.pend_with = p.thr * (ray.emittance + nee);
    p.hit_point = ray.hit_point;
    p.volume_dis = ray.volume_dis;
    p.wb = ray.rng.w;
    p.lcg = ray.lcg;
    p.last_t = float(ray.t);
    r.cont = false;
    if (p.bounce + 1u >= fa.bounces) return;
    double rr = 1.0;
    if (p.bounce > kMinBounces) {  // Russian roulette, rgen:109-118
        Rng4 ra{x, y, fa.frame_count, p.wa};
        rr = clampd(luminance(p.thr), 0.01, 0.95);
        double prop = rng_next(ra);
        p.wa = ra.w;
        if (rr < prop) return;
        p.thr = p.thr / rr;
    }
    p.thr = p.thr * ray.hit_value;
    if (!ray.need_scatter || ray.t < 0.0) return;
    p.o = ray.hit_point;
    p.d = ray.scatter_dir;
    p.bounce++;
    r.cont = true;
}

template <bool kStats>
__device__ inline void shade_bounce(const FrameArgs& fa, uint32_t x, uint32_t y, PathState& p, bool found,
                                    const HitInfo& h, BounceOut& r, PathCounters& pc) {
    NoTap tap;
    shade_bounce<kStats>(fa, x, y, p, found, h, r, pc, tap);
}

constexpr uint32_t kShadowFlags = RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RT_RAY_FLAG_SKIP_CLOSEST_HIT;

// The shadow ray of castShadowRay (rchit:35-59): origin without offset, tMin 0.1,
// tMax the light distance; returns true when unoccluded.
template <bool kAnyHit, bool kStats>
__device__ inline bool shadow_unoccluded(const FrameArgs& fa, uint32_t x, uint32_t y, const BounceOut& r,
                                         uint32_t* stack, uint32_t stride, PathCounters& pc) {
    if (kStats) pc.shadow++;
    Rng4 shadow_rng{x, y, fa.seed_b, r.shadow_wb};
    AlphaAnyHit sah{&fa.sv, &fa.tlas, &shadow_rng};
    HitInfo sh;
    // the fp32 ray, as the wavefront shadow queue holds it
    const float3 so = to_f3(r.so), sd = to_f3(r.sd);
    return !trace_ray<kAnyHit, kStats>(fa.tlas, so, sd, RayFromF32{so, sd}, float(kShadowTMin), float(r.sdist),
                                       kShadowFlags | (fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : 0u), sah, stack, stride,
                                       sh, pc.trs);
}

// The sample loop of RayTracing.rgen main (:49-130) for one pixel from path state p:
// with `resume`, p's current sample has begun and its ray p.o, p.d at bounce
// p.bounce is the next to trace (the state the wavefront queues hold between
// iterations); otherwise the loop begins sample p.sample.  Runs to the last sample.
template <bool kAnyHit, bool kStats>
__device__ inline void run_samples(const FrameArgs& fa, uint32_t x, uint32_t y, PathState& p, bool resume,
                                   uint32_t* stack, uint32_t stride, PathCounters& pc) {
    const uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
    for (; p.sample < fa.samples; p.sample++) {
        if (!resume) begin_sample(fa, x, y, p);
        resume = false;
        if (fa.bounces == 0u) continue;
        for (;;) {
            if (kStats) { if (p.bounce == 0u) pc.primary++; else pc.extension++; }
            Rng4 rb{x, y, fa.seed_b, p.wb};
            AlphaAnyHit ah{&fa.sv, &fa.tlas, &rb};
            HitInfo h;
            bool found = trace_ray<kAnyHit, kStats>(fa.tlas, to_f3(p.o), to_f3(p.d), RayFromF64{p.o, p.d}, float(kTMin),
                                                    float(p.tfar), flags, ah, stack, stride, h, pc.tr);
            p.wb = rb.w;
            BounceOut r;
            shade_bounce<kStats>(fa, x, y, p, found, h, r, pc);
            bool unocc = r.shadow && shadow_unoccluded<kAnyHit, kStats>(fa, x, y, r, stack, stride, pc);
            p.rad = p.rad + (unocc ? r.pend_with : r.pend_without);
            if (!r.cont) break;
        }
    }
}

// RayTracing.rgen main for one pixel (megakernel form); returns the summed radiance
// and the payload's t after the last trace.
template <bool kAnyHit, bool kStats>
__device__ inline D3 render_pixel(const FrameArgs& fa, uint32_t x, uint32_t y, uint32_t* stack, uint32_t stride,
                                  PathCounters& pc, float& last_t) {
    PathState p;
    init_path(fa, x, y, p);
    run_samples<kAnyHit, kStats>(fa, x, y, p, false, stack, stride, pc);
    last_t = p.last_t;
    return p.rad;
}

// ------------------------------------------------------ pixels in stepped lanes
// A lane of a persistent kernel that runs pixels' raygen loops (run_samples: the
// same operations on the same values, in the same order) with its traversals
// stepped rather than run to completion.  The kernel's waves alternate: step the
// lanes that trace (path or shadow ray) while at least half of the lanes that hold
// a pixel do, then the others advance together (shading, next ray, next sample), so
// no lane waits each bounce for the slowest traversal of its wave.
enum : uint32_t { kLaneIdle = 0, kLaneExt, kLaneShadow, kLaneShade, kLaneNext, kLaneDone };
enum : uint32_t { kAfterExt = 0, kAfterDone, kAfterNext };  // what follows a shadow ray

// The ray of a lane's traversal: the path's fp64 ray, or the fp32 shadow ray.
struct LaneRay {
    const PathState& p;
    float3 so, sd;
    bool shadow;
    __device__ void get(D3& O, D3& Dd) const {
        if (shadow) {
            O = d3(so);
            Dd = d3(sd);
        } else {
            O = p.o;
            Dd = p.d;
        }
    }
};

template <bool kAnyHit, bool kStats>
struct PathLane {
    Traversal<kAnyHit, kStats, kStackDepth> tr;
    TraceStats ts;
    PathState p;
    BounceOut r;  // the last shaded bounce: its radiance terms while its shadow ray is traced
    HitInfo h;
    Rng4 rng{0u, 0u, 0u, 0u};
    float3 so{0.f, 0.f, 0.f}, sd{0.f, 0.f, 0.f};
    uint32_t x = 0, y = 0, state = kLaneIdle, after = kAfterNext;
    bool found = false;

    __device__ bool tracing() const { return state == kLaneExt || state == kLaneShadow; }

    // the path's ray (traceRayEXT of rgen:84-96)
    __device__ void start_ext(const FrameArgs& fa, uint32_t* stack, uint32_t stride, PathCounters& pc) {
        if (kStats) { if (p.bounce == 0u) pc.primary++; else pc.extension++; }
        rng = Rng4{x, y, fa.seed_b, p.wb};
        tr.begin(fa.tlas, to_f3(p.o), to_f3(p.d), float(kTMin), float(p.tfar),
                 fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE, stack, stride);
        state = kLaneExt;
    }
    // castShadowRay, rchit:35-59, as shadow_unoccluded traces it
    __device__ void start_shadow(const FrameArgs& fa, float3 o, float3 d, float tmax, uint32_t wb, uint32_t then,
                                 uint32_t* stack, uint32_t stride, PathCounters& pc) {
        if (kStats) pc.shadow++;
        so = o;
        sd = d;
        rng = Rng4{x, y, fa.seed_b, wb};
        tr.begin(fa.tlas, o, d, float(kShadowTMin), tmax, kShadowFlags | (fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : 0u),
                 stack, stride);
        after = then;
        state = kLaneShadow;
    }
    // A new pixel from its first sample (render_pixel); false when it has nothing to
    // trace (state kLaneDone).
    __device__ bool begin_pixel(const FrameArgs& fa, uint32_t px, uint32_t py, uint32_t* stack, uint32_t stride,
                                PathCounters& pc) {
        x = px;
        y = py;
        init_path(fa, x, y, p);
        r.cont = false;
        for (; p.sample < fa.samples; p.sample++) {
            begin_sample(fa, x, y, p);
            if (fa.bounces) {
                start_ext(fa, stack, stride, pc);
                return true;
            }
        }
        state = kLaneDone;
        return false;
    }
    // One traversal step of a tracing lane; at the traversal's end its result.
    template <class AnyHitFn>
    __device__ void step(const FrameArgs& fa, AnyHitFn& ah, uint32_t* stack, uint32_t stride, PathCounters& pc) {
        const bool shadow = state == kLaneShadow;
        if (!tr.step(fa.tlas, LaneRay{p, so, sd, shadow}, ah, stack, stride, ts)) return;
        if (kStats) {
            TraceStats& to = shadow ? pc.trs : pc.tr;
            to.nodes += ts.nodes;
            to.tris += ts.tris;
            to.instances += ts.instances;
            to.anyhits += ts.anyhits;
            to.overflows += ts.overflows;
            ts = TraceStats{};
        }
        if (shadow) {
            p.rad = p.rad + (!tr.found ? r.pend_with : r.pend_without);
            if (after == kAfterExt) r.cont = true;  // the path's current ray comes next
            state = after == kAfterDone ? kLaneDone : kLaneNext;
        } else {
            p.wb = rng.w;
            found = tr.found;
            h = tr.hit;
            state = kLaneShade;
        }
    }
    // A lane that holds a pixel and does not trace: the shaders and the rest of the
    // bounce, then the next ray or sample, until it traces again or its pixel is
    // complete (state kLaneDone).
    __device__ void advance(const FrameArgs& fa, uint32_t* stack, uint32_t stride, PathCounters& pc) {
        if (state == kLaneShade) {
            shade_bounce<kStats>(fa, x, y, p, found, h, r, pc);
            if (r.shadow) {
                start_shadow(fa, to_f3(r.so), to_f3(r.sd), float(r.sdist), r.shadow_wb, kAfterNext, stack, stride, pc);
                return;
            }
            p.rad = p.rad + r.pend_without;
            state = kLaneNext;
        }
        if (state == kLaneNext) {
            if (r.cont) {
                start_ext(fa, stack, stride, pc);
                return;
            }
            state = kLaneDone;
            while (++p.sample < fa.samples) {
                begin_sample(fa, x, y, p);
                if (fa.bounces) {
                    start_ext(fa, stack, stride, pc);
                    return;
                }
            }
        }
    }
};

}  // namespace rt
