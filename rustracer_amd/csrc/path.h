// path.h — the per-pixel integrator of the reference raygen shader and the hit
// shaders it triggers, as device functions shared by the megakernel and the
// wavefront kernels.
//
// Restates crates/examples/gltf_viewer/shaders/RayTracing.rgen:25-167 (sample and
// bounce loops, Russian roulette, accumulation), RayTracing.rchit:136-477
// (closest hit), RayTracing.rmiss:14-45 (miss) and RayTracing.rchit:35-59,
// RayTracing.shadow.rahit, RayTracing.shadow.rmiss (shadow ray).
#pragma once

#include "render.h"

namespace rt {

constexpr float kTMin = 0.001f;     // lib/Camera.glsl:2
constexpr float kTMax = 10000.0f;   // lib/Camera.glsl:3
constexpr float kShadowTMin = 0.1f; // RayTracing.rchit:43
constexpr uint32_t kMinBounces = 3; // RayTracing.rgen:15

struct PathCounters {
    uint32_t primary = 0, extension = 0, shadow = 0, shade = 0, tex = 0, lights = 0, bad = 0;
    TraceStats tr;
};

__device__ inline void flush_counters(unsigned long long* c, const PathCounters& pc) {
    uint32_t v[kNumCounters] = {pc.primary, pc.extension, pc.shadow, pc.tr.nodes, pc.tr.tris, pc.tr.instances,
                                pc.shade, pc.tr.anyhits, pc.tex, pc.lights, pc.tr.overflows, pc.bad, 0, 0, 0, 0};
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

// Camera ray of one sample, RayTracing.rgen:52-75.
__device__ inline void camera_ray(const FrameArgs& fa, uint32_t x, uint32_t y, Rng4& rng_a, uint32_t& lcg,
                                  float3& o, float3& d, float& tfar) {
    float px = float(x) + 0.5f, py = float(y) + 0.5f;
    if (fa.antialiasing) {
        float jx = rng_next(rng_a);
        float jy = rng_next(rng_a);
        px += jx - 0.5f;
        py += jy - 0.5f;
    }
    float uvx = (px / float(fa.width)) * 2.f - 1.f;
    float uvy = (py / float(fa.height)) * 2.f - 1.f;
    float offx = 0.f, offy = 0.f;
    if (fa.aperture != 0.f) {  // RandomInUnitDisk (lib/Random.glsl:44-54); with aperture 0 the offset is 0
        float dx, dy;
        for (;;) {
            float a = lcg_float(lcg), b = lcg_float(lcg);
            dx = 2.f * a - 1.f;
            dy = 2.f * b - 1.f;
            if (dx * dx + dy * dy < 1.f) break;
        }
        offx = fa.aperture / 2.f * dx;
        offy = fa.aperture / 2.f * dy;
    }
    const float* M = fa.mvi;
    const float* P = fa.pin;
    o = mk3(M[0] * offx + M[4] * offy + M[12], M[1] * offx + M[5] * offy + M[13], M[2] * offx + M[6] * offy + M[14]);
    float3 tg = mk3(P[0] * uvx + P[4] * uvy + P[8] + P[12], P[1] * uvx + P[5] * uvy + P[9] + P[13],
                    P[2] * uvx + P[6] * uvy + P[10] + P[14]);
    float3 td = normalize3(tg * fa.focus_distance - mk3(offx, offy, 0.f));
    d = mk3(M[0] * td.x + M[4] * td.y + M[8] * td.z, M[1] * td.x + M[5] * td.y + M[9] * td.z,
            M[2] * td.x + M[6] * td.y + M[10] * td.z);
    tfar = kTMax;
    if (fa.ortho > 0.f) {  // rgen:67-75
        float k = 1.f + fa.ortho;
        float ax = k * uvx, ay = -(k * uvy);
        o = mk3(M[0] * ax + M[4] * ay + M[12], M[1] * ax + M[5] * ay + M[13], M[2] * ax + M[6] * ay + M[14]);
        d = mk3(-M[8], -M[9], -M[10]);
        tfar = 10.f * kTMax;
    }
}

// RayPayload, lib/RayTracingCommons.glsl:11-24
struct Payload {
    float3 hit_value, hit_point, scatter_dir, emittance;
    float t;
    bool need_scatter;
    Rng4 rng;
    float volume_dis;
};

// RayTracing.rmiss:14-45 (skybox off: HasSky is not supported)
__device__ inline void miss_shader(const SceneView& sv, Payload& ray, float3 dir) {
    float3 acc = mk3(0.f, 0.f, 0.f);
    float3 rd = normalize3(dir);
    if (ray.t != 0.f) {
        for (uint32_t i = 0; i < sv.num_dlights; i++) {
            const rt_light& li = sv.dlights[i];
            float c = dot3(normalize3(mk3(li.transform[0], li.transform[1], li.transform[2])), rd);
            if (c < 0.f) acc = acc + mk3(li.color[0], li.color[1], li.color[2]) * (-c * li.intensity);
        }
    }
    acc = acc + mk3(0.01f, 0.01f, 0.01f);
    ray.hit_value = mk3(0.f, 0.f, 0.f);
    ray.need_scatter = false;
    ray.emittance = sv.num_dlights == 0 ? mk3(0.f, 0.f, 0.f) : acc;
    ray.t = -1.f;
}

__device__ inline float3 o2w_dir(const InstanceRecord& ir, float3 v) {
    float4 r0 = ir.o2w[0], r1 = ir.o2w[1], r2 = ir.o2w[2];
    return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z, r1.x * v.x + r1.y * v.y + r1.z * v.z,
               r2.x * v.x + r2.y * v.y + r2.z * v.z);
}
__device__ inline float3 o2w_point(const InstanceRecord& ir, float3 v) {
    float4 r0 = ir.o2w[0], r1 = ir.o2w[1], r2 = ir.o2w[2];
    return mk3(r0.x * v.x + r0.y * v.y + r0.z * v.z + r0.w, r1.x * v.x + r1.y * v.y + r1.z * v.z + r1.w,
               r2.x * v.x + r2.y * v.y + r2.z * v.z + r2.w);
}

// Result of the closest-hit shader up to the point where it casts its shadow ray.
struct ShadeState {
    float3 origin;        // world hit point
    float3 geo_normal;    // geometric normal facing V
    float3 outward;       // shading normal on V's side of the geometric normal
    float3 V;
    MatBrdf m;
    bool front;
    bool has_light;
    float3 light_dir;
    float light_dist;
    float3 light_scale;   // light_weight * intensity * color
    Rng4 rng;             // local stream after the RIS draws
    float displacement;   // |origin - previous hit point|, for volume bookkeeping
};

// First half of RayTracing.rchit main (:136-340): material resolution and the RIS
// light choice.  Returns false when a debug mapping or an invalid reference ended
// the shader (ray already updated).
template <bool kStats>
__device__ inline bool shade_begin(const FrameArgs& fa, Payload& ray, const HitInfo& h, float3 world_dir,
                                   ShadeState& s, PathCounters& pc) {
    const SceneView& sv = fa.sv;
    const InstanceRecord& ir = fa.tlas.inst[h.inst];
    if (kStats) pc.shade++;
    uint32_t ci = ir.custom_index;
    ray.t = h.t;
    ray.need_scatter = false;
    if (ci >= sv.num_prim_infos) { pc.bad++; ray.emittance = mk3(0.f, 0.f, 0.f); return false; }
    const rt_prim_info pi = sv.prim_infos[ci];
    if (pi.material_id >= sv.num_materials) { pc.bad++; ray.emittance = mk3(0.f, 0.f, 0.f); return false; }
    const rt_material_raw& mat = sv.materials[pi.material_id];
    float3 last_hit = ray.hit_point;
    uint32_t io = pi.i_offset + 3u * h.prim;
    if (io + 2u >= sv.num_indices) { pc.bad++; ray.emittance = mk3(0.f, 0.f, 0.f); return false; }
    uint32_t i0 = pi.v_offset + sv.indices[io], i1 = pi.v_offset + sv.indices[io + 1], i2 = pi.v_offset + sv.indices[io + 2];
    if (i0 >= sv.num_vertices || i1 >= sv.num_vertices || i2 >= sv.num_vertices) {
        pc.bad++;
        ray.emittance = mk3(0.f, 0.f, 0.f);
        return false;
    }
    const float4* va = reinterpret_cast<const float4*>(sv.vertices + i0);
    const float4* vb = reinterpret_cast<const float4*>(sv.vertices + i1);
    const float4* vc = reinterpret_cast<const float4*>(sv.vertices + i2);
    // getMixVertexAndGeoNormal, lib/RayTracingCommons.glsl:86-101
    const float bu = h.u, bv = h.v, b0 = 1.f - bu - bv;
    float3 p0 = xyz(va[0]), p1 = xyz(vb[0]), p2 = xyz(vc[0]);
    float4 ua = va[6], ub = vb[6], uc = vc[6];
    float4 uv = make_float4(ua.x * b0 + ub.x * bu + uc.x * bv, ua.y * b0 + ub.y * bu + uc.y * bv,
                            ua.z * b0 + ub.z * bu + uc.z * bv, ua.w * b0 + ub.w * bu + uc.w * bv);
    float3 pos = p0 * b0 + p1 * bu + p2 * bv;
    float4 ca = va[3], cb = vb[3], cc = vc[3];
    float4 vcol = make_float4(ca.x * b0 + cb.x * bu + cc.x * bv, ca.y * b0 + cb.y * bu + cc.y * bv,
                              ca.z * b0 + cb.z * bu + cc.z * bv, ca.w * b0 + cb.w * bu + cc.w * bv);
    float3 normal = normalize3(normalize3(xyz(va[1])) * b0 + normalize3(xyz(vb[1])) * bu + normalize3(xyz(vc[1])) * bv);
    float3 geo_normal = normalize3(o2w_dir(ir, cross3(p1 - p0, p2 - p0)));
    float3 origin = o2w_point(ir, pos);
    float4 color4 = mul4(vcol, ld4(mat.base_color));
    if (mat.base_color_texture.index >= 0) {
        color4 = mul4(color4, tex_uv(sv, mat.base_color_texture, uv));
        if (kStats) pc.tex++;
    }
    float3 color = mk3(color4.x, color4.y, color4.z);
    ray.hit_point = pos;
    if (mat.normal_texture.index >= 0) {  // rchit:178-185, getNormal :124-129
        float4 tx = tex_uv(sv, mat.normal_texture, uv);
        if (kStats) pc.tex++;
        float3 nt = normalize3(mk3(tx.x * 2.f - 1.f, tx.y * 2.f - 1.f, tx.z * 2.f - 1.f));
        float4 ta = va[2], tb = vb[2], tc = vc[2];
        float4 t4 = make_float4(ta.x * b0 + tb.x * bu + tc.x * bv, ta.y * b0 + tb.y * bu + tc.y * bv,
                                ta.z * b0 + tb.z * bu + tc.z * bv, ta.w * b0 + tb.w * bu + tc.w * bv);
        float tl = sqrtf(t4.x * t4.x + t4.y * t4.y + t4.z * t4.z + t4.w * t4.w);
        t4 = make_float4(t4.x / tl, t4.y / tl, t4.z / tl, t4.w / tl);
        float3 tm = xyz(t4);
        float3 tangent = normalize3(tm - normal * dot3(tm, normal));
        float3 bit = normalize3(cross3(normal, tangent) * t4.w);
        normal = tangent * nt.x + bit * nt.y + normal * nt.z;
    }
    normal = normalize3(o2w_dir(ir, normal));
    float3 V = -normalize3(world_dir);
    bool front = dot3(V, geo_normal) >= 0.f;
    geo_normal = front ? geo_normal : -geo_normal;
    float3 outward = dot3(geo_normal, normal) < 0.f ? -normal : normal;
    float3 emit = xyz(ld4(mat.emissive_factor));
    if (mat.emissive_texture.index >= 0) {
        float4 tx = tex_uv(sv, mat.emissive_texture, uv);
        if (kStats) pc.tex++;
        emit = emit * xyz(tx);
    }
    float metallic = mat.metallic_factor, roughness = mat.roughness_factor;
    if (mat.metallic_roughness_texture.index >= 0) {
        float4 tx = tex_uv(sv, mat.metallic_roughness_texture, uv);
        if (kStats) pc.tex++;
        roughness *= tx.y;
        metallic *= tx.z;
    }
    float3 spec_wf = mk3(1.f, 1.f, 1.f);
    const bool sg = mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS;
    if (sg) {  // rchit:215-239
        float4 df = ld4(mat.sg_diffuse_factor);
        float4 sgf = ld4(mat.sg_specular_glossiness_factor);
        if (mat.sg_diffuse_texture.index >= 0) { df = mul4(df, tex_uv(sv, mat.sg_diffuse_texture, uv)); if (kStats) pc.tex++; }
        if (mat.sg_specular_glossiness_texture.index >= 0) {
            sgf = mul4(sgf, tex_uv(sv, mat.sg_specular_glossiness_texture, uv));
            if (kStats) pc.tex++;
        }
        spec_wf = xyz(sgf);
        roughness = 1.f - sgf.w;
        float mx = fmaxf(fmaxf(spec_wf.x, spec_wf.y), spec_wf.z);
        color = xyz(vcol) * xyz(df);
        metallic = convert_metallic(color, spec_wf, mx);
    }
    float trans = 0.f;
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
        float3 e = mk3(0.f, 0.f, 0.f);
        switch (mapping) {
            case RT_MAPPING_ALBEDO: e = color; break;
            case RT_MAPPING_TRIANGLE: e = mk3(b0, bu, bv); break;
            case RT_MAPPING_METALLIC: e = mk3(metallic, metallic, metallic); break;
            case RT_MAPPING_ROUGHNESS: e = mk3(roughness, roughness, roughness); break;
            case RT_MAPPING_NORMAL: e = (outward + mk3(1.f, 1.f, 1.f)) / 2.f; break;
            case RT_MAPPING_TRANSMISSION: e = mk3(trans, trans, trans); break;
            case RT_MAPPING_INSTANCE:
            case RT_MAPPING_GEO_ID: {
                uint32_t k = mapping == RT_MAPPING_INSTANCE ? ir.instance_id : ci;
                k += k << 10; k ^= k >> 6; k += k << 3; k ^= k >> 11; k += k << 15;  // jenkinsHash
                e = mk3(float(k & 255u) / 255.f, float((k >> 8) & 255u) / 255.f, float((k >> 16) & 255u) / 255.f);
                break;
            }
            case RT_MAPPING_TANGENT: {
                float4 ta = va[2], tb = vb[2], tc = vc[2];
                float4 t4 = make_float4(ta.x * b0 + tb.x * bu + tc.x * bv, ta.y * b0 + tb.y * bu + tc.y * bv,
                                        ta.z * b0 + tb.z * bu + tc.z * bv, ta.w * b0 + tb.w * bu + tc.w * bv);
                float tl = sqrtf(t4.x * t4.x + t4.y * t4.y + t4.z * t4.z + t4.w * t4.w);
                e = (normalize3(o2w_dir(ir, xyz(t4) / tl)) + mk3(1.f, 1.f, 1.f)) / 2.f;
                break;
            }
            default: break;
        }
        ray.emittance = e;
        return false;
    }
    float spec_factor = mat.specular_factor;
    float3 spec_color = xyz(ld4(mat.specular_color_factor));
    if (mat.specular_texture.index >= 0) { spec_factor *= tex_uv(sv, mat.specular_texture, uv).w; if (kStats) pc.tex++; }
    if (mat.specular_color_texture.index >= 0) {
        spec_color = spec_color * xyz(tex_uv(sv, mat.specular_color_texture, uv));
        if (kStats) pc.tex++;
    }
    ray.hit_point = origin;
    ray.emittance = emit * fa.exposure;
    MatBrdf& m = s.m;
    m.base_color = color;
    m.metallic = metallic;
    m.roughness = roughness;
    m.ior = mat.volume_exists ? mat.ior : 1.f;
    m.transmission = trans;
    m.specular_factor = spec_factor;
    m.specular_color_factor = spec_color;
    m.front_face = front;
    mat_build(m);
    if (sg) {
        float mx = fmaxf(fmaxf(spec_wf.x, spec_wf.y), spec_wf.z);
        m.c_diff = color * (1.f - mx);
        m.F0 = spec_wf;
    }
    m.attenuation_color = mk3(mat.attenuation_color[0], mat.attenuation_color[1], mat.attenuation_color[2]);
    m.attenuation_distance = mat.attenuation_distance;
    m.volume = mat.volume_exists != 0;
    float displacement = length3(origin - last_hit);
    s.origin = origin;
    s.geo_normal = geo_normal;
    s.outward = outward;
    s.V = V;
    s.front = front;
    // sampleLightRIS, rchit:77-122 (uniform candidate index, Yuksel-attenuated weights)
    Rng4 rng = ray.rng;
    s.has_light = false;
    uint32_t num = sv.num_plights;
    if (num > 0) {
        float total = 0.f, sel_pdf = 0.f;
        uint32_t sel = 0;
        uint32_t cand = num < 3u ? num : 3u;
        for (uint32_t i = 0; i < cand; i++) {
            const rt_light& li = sv.plights[i];
            if (kStats) pc.lights++;
            if (luminance(mk3(li.color[0], li.color[1], li.color[2]) * li.intensity) < 0.1f) continue;
            uint32_t idx = min(num - 1u, uint32_t(rng_next(rng) * float(num)));
            const rt_light& cl = sv.plights[idx];
            if (kStats) pc.lights++;
            float3 lv = mk3(cl.transform[0], cl.transform[1], cl.transform[2]) - origin;
            float ld = length3(lv);
            float3 L = lv / ld;
            if (dot3(geo_normal, L) < 0.00001f) continue;
            float pdf = luminance(light_intensity_at(cl, ld));
            float ris = pdf * float(num);
            total += ris;
            if (rng_next(rng) < ris / total) {
                sel = idx;
                sel_pdf = pdf;
            }
        }
        if (total != 0.f) {
            const rt_light& l = sv.plights[sel];
            float w = (total / 3.f) / sel_pdf;
            float3 lv = mk3(l.transform[0], l.transform[1], l.transform[2]) - origin;
            s.light_dist = length3(lv);
            s.light_dir = normalize3(lv);
            s.light_scale = mk3(l.color[0], l.color[1], l.color[2]) * (w * l.intensity);
            s.has_light = true;
        }
    }
    s.rng = rng;
    s.displacement = displacement;
    return true;
}

// Second half of RayTracing.rchit main (:346-477) after the shadow ray: NEE
// contribution, lobe choice and the next direction.
__device__ inline void shade_end(Payload& ray, ShadeState& s, bool unoccluded) {
    const MatBrdf& m = s.m;
    if (s.has_light && unoccluded)
        ray.emittance = ray.emittance + eval_combined_brdf(s.outward, s.light_dir, s.V, m) * s.light_scale;
    Rng4 rng = s.rng;
    uint32_t type;
    float3 thr = mk3(1.f, 1.f, 1.f);
    if (m.metallic == 1.f && m.roughness == 0.f) {
        type = kSpecular;
    } else {
        float3 p = brdf_probability(m, s.V, s.outward);
        float r = rng_next(rng);
        if (r < p.x) {
            type = kSpecular;
            thr = thr / p.x;
            if (ray.volume_dis >= 0.f) ray.volume_dis += s.displacement;
        } else if (r >= p.x && r <= p.x + p.y) {
            type = kDiffuse;
            thr = thr / p.y;
            if (ray.volume_dis >= 0.f) ray.volume_dis += s.displacement;
        } else {
            type = kTransmission;
            if (m.volume) {
                if (ray.volume_dis >= 0.f) ray.volume_dis += s.displacement;
                else if (s.front) ray.volume_dis = 0.f;
            }
            thr = thr / p.z;
        }
    }
    float3 origin = s.origin;
    if (type == kTransmission) origin = offset_ray(origin, -s.geo_normal);
    float ux = rng_next(rng);
    float uy = rng_next(rng);
    float3 dir = mk3(0.f, 0.f, 0.f), w = mk3(0.f, 0.f, 0.f);
    ray.need_scatter = sample_indirect(ux, uy, s.outward, s.geo_normal, s.V, m, type, dir, w, ray.volume_dis);
    ray.hit_point = origin;
    ray.scatter_dir = dir;
    ray.hit_value = thr * w;
    ray.rng = rng;
}

// RayTracing.rgen main for one pixel (megakernel form); returns the summed radiance.
template <bool kAnyHit, bool kStats>
__device__ inline float3 render_pixel(const FrameArgs& fa, uint32_t x, uint32_t y, uint32_t* stack, uint32_t stride,
                                      PathCounters& pc) {
    Rng4 rng_a{x, y, fa.frame_count, 0u};
    Payload ray;
    ray.rng = Rng4{x, y, fa.seed_b, 0u};
    ray.hit_point = mk3(0.f, 0.f, 0.f);
    ray.hit_value = mk3(0.f, 0.f, 0.f);
    ray.emittance = mk3(0.f, 0.f, 0.f);
    ray.need_scatter = false;
    ray.volume_dis = -1.f;
    uint32_t lcg = tea16(tea16(x, y), fa.seed_b);
    float3 radiance = mk3(0.f, 0.f, 0.f);
    const uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
    const uint32_t shadow_flags = RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT | RT_RAY_FLAG_SKIP_CLOSEST_HIT | flags;
    for (uint32_t s = 0; s < fa.samples; s++) {
        float3 o, d;
        float tfar;
        camera_ray(fa, x, y, rng_a, lcg, o, d, tfar);
        float3 thr = mk3(1.f, 1.f, 1.f);
        ray.t = 0.f;
        ray.volume_dis = -1.f;
        for (uint32_t b = 0; b < fa.bounces; b++) {
            if (kStats) { if (b == 0) pc.primary++; else pc.extension++; }
            AlphaAnyHit ah{&fa.sv, &fa.tlas, &ray.rng};
            HitInfo h;
            bool hit = trace_ray<kAnyHit, kStats>(fa.tlas, o, d, kTMin, tfar, flags, ah, stack, stride, h, pc.tr);
            if (hit) {
                ShadeState ss;
                if (shade_begin<kStats>(fa, ray, h, d, ss, pc)) {
                    bool unocc = false;
                    if (ss.has_light) {
                        if (kStats) pc.shadow++;
                        Rng4 shadow_rng = ray.rng;  // copy of Ray.rngState, discarded at rchit:476
                        AlphaAnyHit sah{&fa.sv, &fa.tlas, &shadow_rng};
                        HitInfo sh;
                        unocc = !trace_ray<kAnyHit, kStats>(fa.tlas, ss.origin, ss.light_dir, kShadowTMin, ss.light_dist,
                                                           shadow_flags, sah, stack, stride, sh, pc.tr);
                    }
                    shade_end(ray, ss, unocc);
                }
            } else {
                miss_shader(fa.sv, ray, d);
            }
            float3 hit_color = ray.hit_value;
            float t = ray.t;
            bool scattered = ray.need_scatter;
            radiance = radiance + thr * ray.emittance;
            ray.emittance = mk3(0.f, 0.f, 0.f);
            if (b + 1 == fa.bounces) break;
            if (b > kMinBounces) {  // Russian roulette, rgen:109-118
                float rr = clampf(luminance(thr), 0.01f, 0.95f);
                float prop = rng_next(rng_a);
                if (rr < prop) break;
                thr = thr / rr;
            }
            thr = thr * hit_color;
            if (!scattered || t < 0.f) break;
            o = ray.hit_point;
            d = ray.scatter_dir;
        }
    }
    return radiance;
}

}  // namespace rt
