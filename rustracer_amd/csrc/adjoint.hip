// adjoint.hip — path-replay adjoint of the trace: the gradient of
// sum(grad_accum * accum) with respect to the texels of RGBA32F parameter textures
// (base colour RGB; metallic-roughness G = roughness, B = metallic) of
// metallic-roughness materials.
//
// One lane per pixel, the megakernel's tiling.  Per sample:
//   pass 1  the forward path (traversal, shading, shadow rays), recording per
//           bounce the hit, the shadow result, the radiance the bounce adds per
//           unit throughput (E_k) and the factor to the next throughput (M_k) --
//           or these records read back from the forward trace that wrote them
//           (RT_OPTION_RECORD_PATHS);
//   back    R_k = E_k + M_k R_{k+1}: the radiance from bounce k on per unit of
//           throughput at k;
//   pass 2  the path replayed from the recorded hits (no traversal): at each
//           vertex whose material reads a parameter texture, the Jacobians of the
//           NEE term and of M_k with respect to (base colour, roughness, metallic)
//           come from dual numbers (brdf_dual.h), and
//             dL/dp = T_k [unoccluded dNEE/dp + dM_k/dp R_{k+1}]
//           is scattered to the texels of the bilinear footprint with atomics.
// Lobe choices, directions, sampling densities and Russian roulette are held at
// their forward values (detached sampling), so the replay sees the forward paths
// exactly.  Follows the raygen loop of
// crates/examples/gltf_viewer/shaders/RayTracing.rgen (:85-128) and the closest-hit
// shading of RayTracing.rchit (:136-416).
#include "adjoint.h"
#include "brdf_dual.h"
#include "path.h"

namespace rt {

namespace {

constexpr int kAdjBlock = 128;

struct AdjView {
    const float4* grad_accum;  // dL/d accum per pixel
    float* const* grads;       // per texture: gradient buffer (RGBA32F layout) or NULL
    uint32_t num_textures;
    const PathRecord* rec;     // the forward's path records (pass 1 skipped), or NULL
};

struct BounceRec {
    HitInfo h;
    uint32_t wb;     // stream-B counter after the traversal (any-hit draws)
    uint32_t flags;  // 1: hit, 2: shadow ray unoccluded
    D3 e, m;         // E_k, M_k
    double rr;       // Russian-roulette probability applied at this bounce (1: none)
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
    if (index < 0 || uint32_t(index) >= av.num_textures || uint32_t(index) >= sv.num_textures) return nullptr;
    if (sv.textures[index].format != RT_FMT_RGBA32F) return nullptr;
    return av.grads[index];
}

__device__ inline D4 tex_coords(rt_texture_info ti, D4 uv) {
    return ti.tex_coord == 1 ? D4{uv.z, uv.w, 0.0, 0.0} : D4{uv.x, uv.y, 0.0, 0.0};
}

// Pass-2 observer: the vertex's contribution to the parameter gradient.
struct GradTap {
    const FrameArgs* fa;
    const AdjView* av;
    D3 gT;       // dL/d radiance at this bounce: pixel gradient * throughput T_k
    D3 r_next;   // R_{k+1}
    bool unocc;
    double rr;
    uint32_t atomics;

    __device__ void vertex(const ShadeState& ss, const Payload& ray) {
        const rt_material_raw& mat = *ss.mat;
        if (mat.workflow == RT_WORKFLOW_SPECULAR_GLOSSINESS) return;
        const SceneView& sv = fa->sv;
        float* gb = param_grad(sv, *av, mat.base_color_texture.index);
        float* gm = param_grad(sv, *av, mat.metallic_roughness_texture.index);
        if (!gb && !gm) return;
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
        if (ss.has_light && unocc) {
            ad::Dv3 f = ad::eval_brdf_d(N, ad::v3(ss.light_dir.x, ss.light_dir.y, ss.light_dir.z), V, md);
            const double sx = gT.x * ss.light_scale.x, sy = gT.y * ss.light_scale.y, sz = gT.z * ss.light_scale.z;
            for (int k = 0; k < ad::kParams; k++) q[k] += sx * f.x.d[k] + sy * f.y.d[k] + sz * f.z.d[k];
        }
        if (ray.need_scatter && (r_next.x != 0.0 || r_next.y != 0.0 || r_next.z != 0.0)) {
            ad::Dv3 w = ad::sample_weight_d(int(ss.lobe), N, ad::v3(ray.scatter_dir.x, ray.scatter_dir.y, ray.scatter_dir.z),
                                            V, md, ss.p_lobe, ss.delta);
            const double sx = gT.x * r_next.x / rr, sy = gT.y * r_next.y / rr, sz = gT.z * r_next.z / rr;
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
            for (int i = 0; i < n; i++) {
                float* g = gb + 4 * size_t(idx[i]);
                atomicAdd(g + 0, float(s0 * tw[i]));
                atomicAdd(g + 1, float(s1 * tw[i]));
                atomicAdd(g + 2, float(s2 * tw[i]));
            }
            atomics += 3u * uint32_t(n);
        }
        if (gm) {
            // roughness = factor * texel.g, metallic = factor * texel.b (rchit:204-213)
            const rt_texture& t = sv.textures[mat.metallic_roughness_texture.index];
            D4 c = tex_coords(mat.metallic_roughness_texture, ss.uv);
            int n = tex_taps(t, c.x, c.y, idx, tw);
            const double sr = q[ad::kRough] * double(mat.roughness_factor);
            const double sm = q[ad::kMetal] * double(mat.metallic_factor);
            for (int i = 0; i < n; i++) {
                float* g = gm + 4 * size_t(idx[i]);
                atomicAdd(g + 1, float(sr * tw[i]));
                atomicAdd(g + 2, float(sm * tw[i]));
            }
            atomics += 2u * uint32_t(n);
        }
    }
};

template <bool kAnyHit>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint(FrameArgs fa, AdjView av) {
    __shared__ uint32_t s_stack[kStackDepth * kAdjBlock];
    const uint32_t tiles_x = (fa.width + 15) / 16;
    const uint32_t bid = blockIdx.x;
    const uint32_t tx = bid % tiles_x, ty = bid / tiles_x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = tx * 16 + wave * 8 + (lane & 7);
    const uint32_t y = ty * 8 + (lane >> 3);
    PathCounters pc;
    if (x >= fa.width || y >= fa.height) return;
    const float4 g4 = av.grad_accum[size_t(y) * fa.width + x];
    const D3 g = d3(g4.x, g4.y, g4.z);
    uint32_t* stack = s_stack + threadIdx.x;
    const uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
    BounceRec rec[kMaxAdjointBounces];
    D3 R[kMaxAdjointBounces + 1];
    uint32_t atomics = 0;
    PathState p;
    init_path(fa, x, y, p);
    for (; p.sample < fa.samples; p.sample++) {
        begin_sample(fa, x, y, p);
        if (fa.bounces == 0u) continue;
        const PathState p0 = p;
        uint32_t nb = 0;
        if (av.rec) {
            // pass 1 from the records the forward trace wrote
            const PathRecord* pr = av.rec + ((size_t(y) * fa.width + x) * fa.samples + p.sample) * fa.bounces;
            for (;;) {
                const PathRecord& q = pr[nb];
                BounceRec& br = rec[nb];
                br.h = HitInfo{q.t, q.u, q.v, q.inst, q.prim, q.rec};
                br.wb = q.wb;
                br.flags = q.flags & (kRecHit | kRecUnoccluded);
                br.e = d3(q.e[0], q.e[1], q.e[2]);
                br.m = d3(q.m[0], q.m[1], q.m[2]);
                br.rr = q.rr;
                nb++;
                if ((q.flags & kRecLast) || nb == kMaxAdjointBounces || nb == fa.bounces) break;
            }
        }
        // pass 1 without records: the forward path traced again
        while (!av.rec) {
            Rng4 rb{x, y, fa.seed_b, p.wb};
            AlphaAnyHit ah{&fa.sv, &fa.tlas, &rb};
            HitInfo h;
            bool found = trace_ray<kAnyHit, false>(fa.tlas, to_f3(p.o), to_f3(p.d), RayFromF64{p.o, p.d}, float(kTMin),
                                                   float(p.tfar), flags, ah, stack, kAdjBlock, h, pc.tr);
            p.wb = rb.w;
            BounceRec& br = rec[nb];
            br.h = h;
            br.wb = p.wb;
            BounceOut r;
            shade_bounce<false>(fa, x, y, p, found, h, r, pc);
            bool unocc = r.shadow && shadow_unoccluded<kAnyHit, false>(fa, x, y, r, stack, kAdjBlock, pc);
            br.flags = (found ? 1u : 0u) | (unocc ? 2u : 0u);
            br.e = unocc ? r.emit + r.nee : r.emit;
            br.m = r.mult;
            br.rr = r.rr;
            nb++;
            if (!r.cont || nb == kMaxAdjointBounces) break;
        }
        // radiance from each bounce on, per unit throughput there
        R[nb] = d3(0.0);
        for (int k = int(nb) - 1; k >= 0; k--) R[k] = rec[k].e + rec[k].m * R[k + 1];
        // pass 2: replay the shading from the recorded hits
        p = p0;
        GradTap tap{&fa, &av, d3(0.0), d3(0.0), false, 1.0, 0u};
        for (uint32_t k = 0; k < nb; k++) {
            const BounceRec& br = rec[k];
            p.wb = br.wb;
            tap.gT = g * p.thr;
            tap.r_next = R[k + 1];
            tap.unocc = (br.flags & 2u) != 0u;
            tap.rr = br.rr;
            BounceOut r;
            shade_bounce<false>(fa, x, y, p, (br.flags & 1u) != 0u, br.h, r, pc, tap);
            if (!r.cont) break;
        }
        atomics += tap.atomics;
    }
    (void)atomics;
}

}  // namespace

int launch_adjoint(const FrameArgs& fa, const float* grad_accum, float* const* d_grads, uint32_t num_textures,
                   const PathRecord* rec, hipStream_t st, bool any_hit) {
    uint32_t tiles = ((fa.width + 15) / 16) * ((fa.height + 7) / 8);
    if (tiles == 0) return 0;
    AdjView av{reinterpret_cast<const float4*>(grad_accum), d_grads, num_textures, rec};
    if (any_hit) hipLaunchKernelGGL((k_adjoint<true>), dim3(tiles), dim3(kAdjBlock), 0, st, fa, av);
    else hipLaunchKernelGGL((k_adjoint<false>), dim3(tiles), dim3(kAdjBlock), 0, st, fa, av);
    return int(hipGetLastError());
}

}  // namespace rt
