// adjoint.hip — path-replay adjoint of the trace: the gradient of
// sum(grad_accum * accum) with respect to the texels of RGBA32F parameter textures
// (base colour RGB; metallic-roughness G = roughness, B = metallic) of
// metallic-roughness materials, and to the intensities of the point and
// directional lights.
//
// Path replay backpropagation (Vicini, Speierer & Jakob, SIGGRAPH 2021): nothing of
// a path is stored between its two passes but its radiance and one bit per bounce.
// One lane per pixel at a time (persistent lanes taking pixels in turn); per sample:
//   pass 1  the forward path (traversal, shading, shadow rays) gives the sample's
//           radiance L and the shadow results (one bit per bounce);
//   pass 2  the same path traced and shaded again from the same RNG state.  At
//           vertex k with throughput T_k, radiance added E_k and throughput factor
//           M_k (T_{k+1} = T_k M_k), the radiance still to come is
//             L_k = L - sum_{j<=k} T_j E_j = T_{k+1} R_{k+1},
//           so the vertex's gradient  T_k dE_k/dp + T_k R_{k+1} dM_k/dp  is
//             T_k dE_k/dp + L_k (dM_k/dp) / M_k      (per colour channel).
// The Jacobians of the NEE term and of M_k with respect to (base colour, roughness,
// metallic) come from dual numbers (brdf_dual.h); those with respect to a light's
// intensity are the NEE and miss terms over that intensity.  A colour channel whose
// M_k is exactly zero carries no radiance past vertex k, and its dM_k/dp term is
// dropped (the known limit of path replay).
//
// Lobe choices, directions, sampling densities (the RIS light weight included) and
// Russian roulette are held at their forward values (detached sampling), so the
// replay sees the forward paths exactly.
//
// Scatter.  A vertex's contributions (texel x channel, or light) are summed per
// wave in an LDS table keyed by the target address before any float atomic: the
// camera-ray vertices of a wave's 8x8 pixels and 16 samples share a few hundred
// texels, and every lane's NEE adds to the same few lights.  The table is drained
// into the gradient buffers with one atomic per key when it is half full at the end
// of a sample, and at the end of the kernel; a key that finds no free slot within
// a few probes goes to the atomics directly.
//
// Follows the raygen loop of crates/examples/gltf_viewer/shaders/RayTracing.rgen
// (:85-128), the closest-hit shading of RayTracing.rchit (:136-416, NEE :338-350)
// and the miss shader RayTracing.rmiss (:14-45).
#include <algorithm>

#include "adjoint.h"
#include "brdf_dual.h"
#include "path.h"

namespace rt {

namespace {

constexpr int kAdjBlock = 128;
constexpr int kAdjWaves = kAdjBlock / 64;
constexpr int kSlots = 256;  // LDS table entries per wave
constexpr int kProbe = 4;    // slots a key tries before its atomics go out directly

struct AdjView {
    const float4* grad_accum;  // dL/d accum per pixel
    AdjointOut out;
    uint32_t table_bounces;    // vertices at bounces below this sum texel taps in LDS
};

struct GradCounts {
    uint32_t taps = 0, atomics = 0;
};

// A wave's table of pending gradient sums: key = the target's address (its first
// float), value = the sums for the target's first three floats.
struct WaveTable {
    unsigned long long* keys;
    float4* vals;
    uint32_t* used;
    uint32_t lane;

    // every lane of the wave
    __device__ void clear() {
        for (int s = int(lane); s < kSlots; s += 64) {
            keys[s] = 0ull;
            vals[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        if (lane == 0) *used = 0u;
        __threadfence_block();
    }
    // any lane: false when no slot was free within kProbe probes
    __device__ bool add(float* a, float v0, float v1, float v2) {
        const unsigned long long key = reinterpret_cast<unsigned long long>(a);
        const uint32_t h = uint32_t(((key >> 2) * 0x9E3779B97F4A7C15ull) >> 56);
        for (int i = 0; i < kProbe; i++) {
            const uint32_t s = (h + uint32_t(i)) & uint32_t(kSlots - 1);
            const unsigned long long old = atomicCAS(&keys[s], 0ull, key);
            if (old == 0ull) atomicAdd(used, 1u);
            if (old == 0ull || old == key) {
                if (v0 != 0.f) atomicAdd(&vals[s].x, v0);
                if (v1 != 0.f) atomicAdd(&vals[s].y, v1);
                if (v2 != 0.f) atomicAdd(&vals[s].z, v2);
                return true;
            }
        }
        return false;
    }
    __device__ uint32_t occupancy() {
        __threadfence_block();
        return *used;
    }
    // every lane of the wave: one float atomic per non-zero sum, table cleared
    __device__ uint32_t drain() {
        __threadfence_block();
        uint32_t n = 0;
        for (int s = int(lane); s < kSlots; s += 64) {
            const unsigned long long k = keys[s];
            if (k == 0ull) continue;
            float* a = reinterpret_cast<float*>(k);
            const float4 v = vals[s];
            if (v.x != 0.f) { atomicAdd(a, v.x); n++; }
            if (v.y != 0.f) { atomicAdd(a + 1, v.y); n++; }
            if (v.z != 0.f) { atomicAdd(a + 2, v.z); n++; }
            keys[s] = 0ull;
            vals[s] = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __threadfence_block();
        if (lane == 0) *used = 0u;
        __threadfence_block();
        return n;
    }
};

// Bilinear footprint of sample_texture (same coordinates and wrap): texel indices
// and weights.  Returns the number of taps.
__device__ inline int tex_taps(const rt_texture& t, double u, double v, uint32_t idx[4], double w[4]) {
    const int W = int(t.width), H = int(t.height);
    if (t.mag_filter == RT_FILTER_NEAREST) {
        int x = wrap_coord(int(floor(u * W)), W, t.wrap_s);
        int y = wrap_coord(int(floor(v * H)), H, t.wrap_t);
        idx[0] = uint32_t(y) * t.width + uint32_t(x);
        w[0] = 1.0;
        return 1;
    }
    double fu = u * W - 0.5, fv = v * H - 0.5;
    double iu = floor(fu), iv = floor(fv);
    double a = fu - iu, b = fv - iv;
    int x0 = wrap_coord(int(iu), W, t.wrap_s), x1 = wrap_coord(int(iu) + 1, W, t.wrap_s);
    int y0 = wrap_coord(int(iv), H, t.wrap_t), y1 = wrap_coord(int(iv) + 1, H, t.wrap_t);
    idx[0] = uint32_t(y0) * t.width + uint32_t(x0);
    idx[1] = uint32_t(y0) * t.width + uint32_t(x1);
    idx[2] = uint32_t(y1) * t.width + uint32_t(x0);
    idx[3] = uint32_t(y1) * t.width + uint32_t(x1);
    w[0] = (1.0 - a) * (1.0 - b);
    w[1] = a * (1.0 - b);
    w[2] = (1.0 - a) * b;
    w[3] = a * b;
    return 4;
}

// Gradient buffer of texture `index` if it is a parameter texture.
__device__ inline float* param_grad(const SceneView& sv, const AdjView& av, int32_t index) {
    if (index < 0 || uint32_t(index) >= av.out.num_textures || uint32_t(index) >= sv.num_textures) return nullptr;
    if (sv.textures[index].format != RT_FMT_RGBA32F) return nullptr;
    return av.out.grads[index];
}

__device__ inline D4 tex_coords(rt_texture_info ti, D4 uv) {
    return ti.tex_coord == 1 ? D4{uv.z, uv.w, 0.0, 0.0} : D4{uv.x, uv.y, 0.0, 0.0};
}

// Adds (v0, v1, v2) to a[0..2]: through the wave's table when `table`, else (or when
// the table has no room) with float atomics.
__device__ inline void scatter(WaveTable& tab, bool table, GradCounts& gc, float* a, double v0, double v1, double v2) {
    gc.taps++;
    const float f0 = float(v0), f1 = float(v1), f2 = float(v2);
    if (table && tab.add(a, f0, f1, f2)) return;
    if (f0 != 0.f) { atomicAdd(a, f0); gc.atomics++; }
    if (f1 != 0.f) { atomicAdd(a + 1, f1); gc.atomics++; }
    if (f2 != 0.f) { atomicAdd(a + 2, f2); gc.atomics++; }
}

// Pass-2 observer of one shaded vertex.
struct ReplayTap {
    const FrameArgs* fa;
    const AdjView* av;
    WaveTable* tab;
    GradCounts* gc;
    D3 g;        // dLoss/d radiance of the pixel
    D3 thr;      // T_k
    D3 l_rem;    // radiance of the sample not yet accounted for before this vertex
    bool unocc;  // the forward's shadow ray reached the light
    bool cont;   // the forward path went on past this vertex
    bool table;  // texel taps go through the wave's table

    __device__ void vertex(const ShadeState& ss, const Payload& ray, D3 f, D3 nee) {
        // L_k: the forward added thr * (emittance + NEE) when unoccluded (same expression)
        const D3 rem = l_rem - (unocc ? thr * (ray.emittance + nee) : thr * ray.emittance);
        const D3 gT = g * thr;
        const SceneView& sv = fa->sv;
        // point-light intensity: NEE = f * colour * (w * I) with the RIS weight w held
        if (ss.has_light && unocc && av->out.grad_plight) {
            const rt_light& l = sv.plights[ss.light_index];
            const double v = (gT.x * f.x * double(l.color[0]) + gT.y * f.y * double(l.color[1]) +
                              gT.z * f.z * double(l.color[2])) * ss.light_w;
            scatter(*tab, true, *gc, av->out.grad_plight + ss.light_index, v, 0.0, 0.0);
        }
        const rt_material_raw& mat = *ss.mat;
        if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) return;
        float* gb = param_grad(sv, *av, mat.base_color_texture.index);
        float* gm = param_grad(sv, *av, mat.metallic_roughness_texture.index);
        if (!gb && !gm) return;
        const bool nee_on = ss.has_light && unocc;
        const bool m_on = cont && ray.need_scatter && (rem.x != 0.0 || rem.y != 0.0 || rem.z != 0.0);
        if (!nee_on && !m_on) return;
        ad::MatIn in;
        in.base[0] = ss.m.base_color.x;
        in.base[1] = ss.m.base_color.y;
        in.base[2] = ss.m.base_color.z;
        in.metallic = ss.m.metallic;
        in.roughness = ss.m.roughness;
        in.ior = ss.m.ior;
        in.transmission = ss.m.transmission;
        in.specular_factor = ss.m.specular_factor;
        in.spec_color[0] = ss.m.specular_color_factor.x;
        in.spec_color[1] = ss.m.specular_color_factor.y;
        in.spec_color[2] = ss.m.specular_color_factor.z;
        const ad::MatD md = ad::mat_dual(in);
        const ad::V3 N = ad::v3(ss.outward.x, ss.outward.y, ss.outward.z);
        const ad::V3 V = ad::v3(ss.V.x, ss.V.y, ss.V.z);
        double q[ad::kParams];
        for (int k = 0; k < ad::kParams; k++) q[k] = 0.0;
        if (nee_on) {
            ad::Dv3 fd = ad::eval_brdf_d(N, ad::v3(ss.light_dir.x, ss.light_dir.y, ss.light_dir.z), V, md);
            const double sx = gT.x * ss.light_scale.x, sy = gT.y * ss.light_scale.y, sz = gT.z * ss.light_scale.z;
            for (int k = 0; k < ad::kParams; k++) q[k] += sx * fd.x.d[k] + sy * fd.y.d[k] + sz * fd.z.d[k];
        }
        if (m_on) {
            // T_k R_{k+1} dM_k/dp = L_k (dw/dp) / w: M_k is the sampled weight w over
            // factors that do not depend on p (Russian roulette, Beer attenuation)
            ad::Dv3 w = ad::sample_weight_d(int(ss.lobe), N, ad::v3(ray.scatter_dir.x, ray.scatter_dir.y, ray.scatter_dir.z),
                                            V, md, ss.p_lobe, ss.delta);
            const double sx = w.x.v != 0.0 ? g.x * rem.x / w.x.v : 0.0;
            const double sy = w.y.v != 0.0 ? g.y * rem.y / w.y.v : 0.0;
            const double sz = w.z.v != 0.0 ? g.z * rem.z / w.z.v : 0.0;
            for (int k = 0; k < ad::kParams; k++) q[k] += sx * w.x.d[k] + sy * w.y.d[k] + sz * w.z.d[k];
        }
        uint32_t idx[4];
        double tw[4];
        if (gb) {
            // base colour = vertex colour * factor * texel (rchit:166-174)
            const rt_texture& t = sv.textures[mat.base_color_texture.index];
            D4 c = tex_coords(mat.base_color_texture, ss.uv);
            int n = tex_taps(t, c.x, c.y, idx, tw);
            const double s0 = q[ad::kBaseR] * ss.vcol.x * double(mat.base_color[0]);
            const double s1 = q[ad::kBaseG] * ss.vcol.y * double(mat.base_color[1]);
            const double s2 = q[ad::kBaseB] * ss.vcol.z * double(mat.base_color[2]);
            for (int i = 0; i < n; i++)
                scatter(*tab, table, *gc, gb + 4 * size_t(idx[i]), s0 * tw[i], s1 * tw[i], s2 * tw[i]);
        }
        if (gm) {
            // roughness = factor * texel.g, metallic = factor * texel.b (rchit:204-213)
            const rt_texture& t = sv.textures[mat.metallic_roughness_texture.index];
            D4 c = tex_coords(mat.metallic_roughness_texture, ss.uv);
            int n = tex_taps(t, c.x, c.y, idx, tw);
            const double sr = q[ad::kRough] * double(mat.roughness_factor);
            const double sm = q[ad::kMetal] * double(mat.metallic_factor);
            for (int i = 0; i < n; i++)
                scatter(*tab, table, *gc, gm + 4 * size_t(idx[i]), 0.0, sr * tw[i], sm * tw[i]);
        }
    }
};

// Directional-light intensities at a miss after the first bounce (rmiss:18-26): the
// emittance holds max(0, -cos_i) colour_i I_i, doubled under a sky (rmiss:27-33).
__device__ inline void miss_light_grads(const SceneView& sv, const AdjView& av, D3 dir, D3 gT, WaveTable& tab,
                                        GradCounts& gc) {
    const D3 rd = dnormalize(dir);
    const double k = sv.has_sky ? 2.0 : 1.0;
    for (uint32_t i = 0; i < sv.num_dlights; i++) {
        const rt_light& li = sv.dlights[i];
        const double c = ddot(dnormalize(ld3(li.transform)), rd);
        if (c >= 0.0) continue;
        const double v = -c * k * (gT.x * double(li.color[0]) + gT.y * double(li.color[1]) + gT.z * double(li.color[2]));
        scatter(tab, true, gc, av.out.grad_dlight + i, v, 0.0, 0.0);
    }
}

// Persistent: as many workgroups as can be resident; each lane takes a pixel, runs
// its samples one per loop turn and then takes the next pixel (one atomic per wave
// on `next_pixel` for the lanes that need one; pixels numbered 8x8 tile by tile, so
// the lanes a wave serves together sit in one or two tiles).  A wave no longer
// lasts as long as its slowest pixel.  Each pixel's samples stay in order in one
// lane, so the gradients are the same up to the order of the float sums.
template <bool kAnyHit, bool kStats>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint(FrameArgs fa, AdjView av, uint32_t* next_pixel) {
    __shared__ uint32_t s_stack[kStackDepth * kAdjBlock];
    __shared__ unsigned long long s_keys[kAdjWaves][kSlots];
    __shared__ float4 s_vals[kAdjWaves][kSlots];
    __shared__ uint32_t s_used[kAdjWaves];
    const uint32_t tiles_x = (fa.width + 7) / 8;
    const uint32_t npix = tiles_x * ((fa.height + 7) / 8) * 64u;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    WaveTable tab{s_keys[wave], s_vals[wave], &s_used[wave], lane};
    tab.clear();
    PathCounters pc;
    GradCounts gc;
    uint32_t* stack = s_stack + threadIdx.x;
    const uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
    const bool dl_grads = av.out.grad_dlight != nullptr && fa.sv.num_dlights != 0u;
    // the lane's pixel, its next sample and its path state between samples
    bool have = false, exhausted = false;
    uint32_t x = 0, y = 0, s = 0;
    D3 g = d3(0.0);
    PathState p;
    for (;;) {
        if (!exhausted) {
            // lanes without a pixel take the next ones: one atomic per wave
            const uint64_t m = __ballot(!have);
            if (m) {
                const uint32_t first = __ffsll((unsigned long long)m) - 1;
                uint32_t base = 0;
                if (lane == first) base = atomicAdd(next_pixel, uint32_t(__popcll(m)));
                base = __shfl(base, first, 64);
                if (base + uint32_t(__popcll(m)) >= npix) exhausted = true;
                const uint32_t i = base + uint32_t(__popcll(m & lanemask_lt()));
                if (!have && i < npix) {
                    const uint32_t tile = i >> 6, within = i & 63u;
                    x = (tile % tiles_x) * 8 + (within & 7u);
                    y = (tile / tiles_x) * 8 + (within >> 3);
                    g = d3(0.0);
                    if (x < fa.width && y < fa.height) {
                        const float4 g4 = av.grad_accum[size_t(y) * fa.width + x];
                        g = d3(g4.x, g4.y, g4.z);
                    }
                    // a pixel outside the image or with a zero gradient does no work
                    have = (g.x != 0.0 || g.y != 0.0 || g.z != 0.0) && fa.bounces != 0u && fa.samples != 0u;
                    if (have) {
                        init_path(fa, x, y, p);
                        s = 0;
                    }
                }
            }
        }
        if (!__ballot(have)) {
            if (exhausted) break;
            continue;
        }
        if (have) {
            // sample s of the lane's pixel
            p.sample = s;
            begin_sample(fa, x, y, p);
            const PathState p0 = p;
            // pass 1: the sample's radiance and shadow results
            D3 L = d3(0.0);
            uint64_t vis = 0;
            uint32_t nb = 0;
            for (;;) {
                if (kStats) { if (p.bounce == 0u) pc.primary++; else pc.extension++; }
                Rng4 rb{x, y, fa.seed_b, p.wb};
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &rb};
                HitInfo h;
                bool found = trace_ray<kAnyHit, kStats>(fa.tlas, to_f3(p.o), to_f3(p.d), RayFromF64{p.o, p.d},
                                                        float(kTMin), float(p.tfar), flags, ah, stack, kAdjBlock, h, pc.tr);
                p.wb = rb.w;
                BounceOut r;
                shade_bounce<false>(fa, x, y, p, found, h, r, pc);
                const bool unocc = r.shadow && shadow_unoccluded<kAnyHit, kStats>(fa, x, y, r, stack, kAdjBlock, pc);
                L = L + (unocc ? r.pend_with : r.pend_without);
                if (unocc) vis |= 1ull << nb;
                nb++;
                if (!r.cont) break;
            }
            // pass 2: the same path again, differentiated vertex by vertex
            PathState q = p0;
            ReplayTap tap{&fa, &av, &tab, &gc, g, d3(0.0), L, false, false, false};
            for (uint32_t k = 0; k < nb; k++) {
                if (kStats) { if (k == 0u) pc.primary++; else pc.extension++; }
                Rng4 rb{x, y, fa.seed_b, q.wb};
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &rb};
                HitInfo h;
                bool found = trace_ray<kAnyHit, kStats>(fa.tlas, to_f3(q.o), to_f3(q.d), RayFromF64{q.o, q.d},
                                                        float(kTMin), float(q.tfar), flags, ah, stack, kAdjBlock, h, pc.tr);
                q.wb = rb.w;
                tap.thr = q.thr;
                tap.unocc = ((vis >> k) & 1ull) != 0ull;
                tap.cont = k + 1u < nb;
                tap.table = k < av.table_bounces;
                const D3 dir = q.d;
                BounceOut r;
                shade_bounce<false>(fa, x, y, q, found, h, r, pc, tap);
                if (!found && k > 0u && dl_grads) miss_light_grads(fa.sv, av, dir, g * tap.thr, tab, gc);
                tap.l_rem = tap.l_rem - (tap.unocc ? r.pend_with : r.pend_without);
            }
            if (++s == fa.samples) have = false;
        }
        // the whole wave: drain the table before it fills
        if (tab.occupancy() > uint32_t(kSlots / 2)) gc.atomics += tab.drain();
    }
    gc.atomics += tab.drain();
    if (kStats) {
        flush_counters(fa.counters, pc);
        const unsigned long long t = wave_sum<unsigned long long>(gc.taps);
        const unsigned long long a = wave_sum<unsigned long long>(gc.atomics);
        if (lane == 0) {
            if (t) atomicAdd(&fa.counters[kCntGradTaps], t);
            if (a) atomicAdd(&fa.counters[kCntGradAtomics], a);
        }
    }
}

}  // namespace

int launch_adjoint(const FrameArgs& fa, const float* grad_accum, const AdjointOut& out, uint32_t table_bounces,
                   hipStream_t st, bool any_hit, bool stats, uint32_t* next_pixel) {
    const uint32_t tiles = ((fa.width + 7) / 8) * ((fa.height + 7) / 8);
    if (tiles == 0) return 0;
    AdjView av{reinterpret_cast<const float4*>(grad_accum), out, table_bounces};
    const void* f = any_hit ? (stats ? (const void*)k_adjoint<true, true> : (const void*)k_adjoint<true, false>)
                            : (stats ? (const void*)k_adjoint<false, true> : (const void*)k_adjoint<false, false>);
    int dev = 0, cus = 0, per_cu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kAdjBlock, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    // resident workgroups, no more than it takes to start every 8x8 tile at once
    const uint32_t g = std::min(uint32_t(cus) * uint32_t(per_cu), (tiles + kAdjWaves - 1) / kAdjWaves);
    hipError_t e = hipMemsetAsync(next_pixel, 0, sizeof(uint32_t), st);
    if (e != hipSuccess) return int(e);
    if (stats) {
        if (any_hit) hipLaunchKernelGGL((k_adjoint<true, true>), dim3(g), dim3(kAdjBlock), 0, st, fa, av, next_pixel);
        else hipLaunchKernelGGL((k_adjoint<false, true>), dim3(g), dim3(kAdjBlock), 0, st, fa, av, next_pixel);
    } else {
        if (any_hit) hipLaunchKernelGGL((k_adjoint<true, false>), dim3(g), dim3(kAdjBlock), 0, st, fa, av, next_pixel);
        else hipLaunchKernelGGL((k_adjoint<false, false>), dim3(g), dim3(kAdjBlock), 0, st, fa, av, next_pixel);
    }
    return int(hipGetLastError());
}

}  // namespace rt
