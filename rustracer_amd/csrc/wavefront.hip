// wavefront.hip — the wavefront form of the trace (RT_MODE_WAVEFRONT).
//
// The raygen loop of the reference (crates/examples/gltf_viewer/shaders/RayTracing.rgen)
// is split into queue kernels that each do one stage for every live path:
//   init     path state of every pixel, camera ray of sample 0           -> queue
//   extend   closest-hit traversal (fp32, small register footprint)
//   shade    closest-hit / miss shaders in fp64, NEE set-up, next ray     -> shadow queue
//   shadow   shadow-ray traversal with terminate-on-first-hit
//   resolve  radiance update; a path that ended starts its pixel's next
//            sample or retires the pixel into the accumulation image       -> next queue
// One slot per pixel keeps each pixel's RNG streams sequential, so the image is
// the megakernel's bit for bit (both call the same device functions in path.h).
// Queues are compacted with wave ballots (one atomic per wave).  Kernels are
// persistent grid-stride loops that read their queue length from device memory,
// so the host only checks for completion every few iterations.
#include <algorithm>
#include <unordered_map>
#include <vector>

#include "path.h"
#include "wavefront.h"

namespace rt {

namespace {

constexpr int kWfBlock = 256;
constexpr uint32_t kCheckEvery = 8;

struct WfArrays {
    double *ox, *oy, *oz, *dx, *dy, *dz;  // ray
    double *tx, *ty, *tz, *rx, *ry, *rz;  // throughput, radiance
    double *hx, *hy, *hz, *vol, *tfar;    // Ray.hitPoint, Ray.volume_dis, tMax
    double *pwx, *pwy, *pwz, *pox, *poy, *poz;  // pending radiance with / without NEE
    uint32_t *wa, *wb, *lcg, *sample, *bounce;
    uint32_t *h_inst, *h_prim, *h_rec;
    float *h_t, *h_u, *h_v;
    float* last_t;     // payload t after the last trace (DISTANCE mapping only)
    uint32_t *flags;   // kFlagShadow | kFlagCont
    uint32_t *unocc;   // shadow result
    float4 *sq_o, *sq_d;  // shadow queue: origin + tmax, direction + slot bits
    uint32_t* sq_wb;
    uint32_t *q0, *q1;
    uint32_t* counts;  // [0], [1]: path queues (ping-pong), [2]: shadow queue,
                       // [3]: extend fetch cursor, [4]: shadow fetch cursor
    uint32_t refill_ext, refill_sh;  // busy lanes below which a wave takes new rays (1..64)
};

enum : uint32_t { kFlagShadow = 1u, kFlagCont = 2u };

__device__ inline void slot_pixel(const FrameArgs& fa, uint32_t s, uint32_t& x, uint32_t& y) {
    // 8x8 pixel tiles: consecutive slots are neighbouring pixels
    uint32_t tiles_x = (fa.width + 7) / 8;
    uint32_t tile = s >> 6, within = s & 63u;
    x = (tile % tiles_x) * 8 + (within & 7u);
    y = (tile / tiles_x) * 8 + (within >> 3);
}

__device__ inline void store_ray(const WfArrays& a, uint32_t s, const PathState& p) {
    a.ox[s] = p.o.x; a.oy[s] = p.o.y; a.oz[s] = p.o.z;
    a.dx[s] = p.d.x; a.dy[s] = p.d.y; a.dz[s] = p.d.z;
}
__device__ inline void load_full(const WfArrays& a, uint32_t s, PathState& p) {
    p.o = d3(a.ox[s], a.oy[s], a.oz[s]);
    p.d = d3(a.dx[s], a.dy[s], a.dz[s]);
    p.thr = d3(a.tx[s], a.ty[s], a.tz[s]);
    p.rad = d3(a.rx[s], a.ry[s], a.rz[s]);
    p.hit_point = d3(a.hx[s], a.hy[s], a.hz[s]);
    p.volume_dis = a.vol[s];
    p.tfar = a.tfar[s];
    p.wa = a.wa[s];
    p.wb = a.wb[s];
    p.lcg = a.lcg[s];
    p.sample = a.sample[s];
    p.bounce = a.bounce[s];
}
__device__ inline void store_full(const WfArrays& a, uint32_t s, const PathState& p) {
    store_ray(a, s, p);
    a.tx[s] = p.thr.x; a.ty[s] = p.thr.y; a.tz[s] = p.thr.z;
    a.rx[s] = p.rad.x; a.ry[s] = p.rad.y; a.rz[s] = p.rad.z;
    a.hx[s] = p.hit_point.x; a.hy[s] = p.hit_point.y; a.hz[s] = p.hit_point.z;
    a.vol[s] = p.volume_dis;
    a.tfar[s] = p.tfar;
    a.wa[s] = p.wa;
    a.wb[s] = p.wb;
    a.lcg[s] = p.lcg;
    a.sample[s] = p.sample;
    a.bounce[s] = p.bounce;
}

// The fp64 ray of a path slot, re-read from the path state when traversal enters
// an instance (keeps it out of the traversal's registers).
struct RayFromSlot {
    const WfArrays& a;
    uint32_t s;
    __device__ void get(D3& O, D3& Dd) const {
        O = d3(a.ox[s], a.oy[s], a.oz[s]);
        Dd = d3(a.dx[s], a.dy[s], a.dz[s]);
    }
};

// Dynamic ray fetch for the traversal kernels: a lane whose ray has ended takes the
// next queue entry (one atomic per wave); a wave keeps stepping its rays while at
// least `refill` of its lanes are busy, then refills.  Traversal lengths vary by an
// order of magnitude between rays, so a wave that only took new rays when all 64
// had ended (refill 1) would idle many lanes; refilling often costs a ballot and an
// atomic per refill.  Returns the queue index for the lane, or kEmpty (sets
// `exhausted`, uniformly over the wave, once the queue has run out).
__device__ inline uint32_t wave_fetch(bool need, uint32_t* cursor, uint32_t n, bool& exhausted) {
    uint64_t m = __ballot(need);
    if (!m) return kEmpty;
    uint32_t first = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (__lane_id() == first) base = atomicAdd(cursor, uint32_t(__popcll(m)));
    base = __shfl(base, first, 64);
    if (base + uint32_t(__popcll(m)) >= n) exhausted = true;
    uint32_t i = base + uint32_t(__popcll(m & lanemask_lt()));
    return (need && i < n) ? i : kEmpty;
}

// Appends `item` to queue q (count c) for the lanes with want set: one atomic per wave.
__device__ inline uint32_t wave_append(bool want, uint32_t* c) {
    uint64_t m = __ballot(want);
    uint32_t base = 0;
    if (m) {
        uint32_t first = __ffsll((unsigned long long)m) - 1;
        if (__lane_id() == first) base = atomicAdd(c, uint32_t(__popcll(m)));
        base = __shfl(base, first, 64);
    }
    return base + uint32_t(__popcll(m & lanemask_lt()));
}

// Where a slot's frame arguments come from: one frame (kernel argument), or a
// batch of frames in device memory with the slots of frame i in
// [base[i], base[i + 1]) (rt_trace_rays_batch).
struct OneFrame {
    FrameArgs fa;
    __device__ const FrameArgs& at(uint32_t s, uint32_t& x, uint32_t& y) const {
        slot_pixel(fa, s, x, y);
        return fa;
    }
    __device__ unsigned long long* counters() const { return fa.counters; }
};
struct ManyFrames {
    const FrameArgs* fas;
    const uint32_t* base;  // n + 1 slot offsets
    uint32_t n;
    unsigned long long* cnt;
    __device__ const FrameArgs& at(uint32_t s, uint32_t& x, uint32_t& y) const {
        uint32_t lo = 0, hi = n;
        while (hi - lo > 1) {
            uint32_t m = (lo + hi) >> 1;
            if (base[m] <= s) lo = m; else hi = m;
        }
        const FrameArgs& fa = fas[lo];
        slot_pixel(fa, s - base[lo], x, y);
        return fa;
    }
    __device__ unsigned long long* counters() const { return cnt; }
};

__device__ inline void write_accum(const FrameArgs& fa, uint32_t x, uint32_t y, D3 rad, float last_t) {
    float4* acc = reinterpret_cast<float4*>(fa.accum) + size_t(y) * fa.width + x;
    float4 prev = fa.reset ? make_float4(0.f, 0.f, 0.f, 0.f) : *acc;
    const float4 sum = make_float4(prev.x + float(rad.x), prev.y + float(rad.y), prev.z + float(rad.z), 0.f);
    *acc = sum;
    if (fa.o.out) write_output(fa.o, fa.width, x, y, sum, last_t, 0u);
}

template <class Frames>
__global__ __launch_bounds__(kWfBlock) void k_wf_init(Frames fr, WfArrays a, uint32_t nslots) {
    uint32_t* q = a.q0;
    for (uint32_t base = blockIdx.x * kWfBlock; base < nslots; base += gridDim.x * kWfBlock) {
        uint32_t s = base + threadIdx.x;
        uint32_t x = 0, y = 0;
        bool valid = s < nslots;
        if (valid) {
            const FrameArgs& fa = fr.at(s, x, y);
            valid = x < fa.width && y < fa.height;
        }
        if (valid) {
            const FrameArgs& fa = fr.at(s, x, y);
            PathState p;
            init_path(fa, x, y, p);
            begin_sample(fa, x, y, p);
            store_full(a, s, p);
        }
        uint32_t idx = wave_append(valid, &a.counts[0]);
        if (valid) q[idx] = s;
    }
}

// Traversal variants of the extend and shadow kernels: LDS stack depth and the
// waves per SIMD the register allocation targets (the compiler may spill above
// its natural occupancy; the any-hit kernels keep their natural allocation).
// Selected per context (RT_OPTION_TRAVERSAL_VARIANT).
struct TravVariant {
    int depth, waves;
};
constexpr TravVariant kTravVariants[] = {{32, 5}, {16, 6}, {16, 8}, {64, 2}};
constexpr int kNumTravVariants = int(sizeof(kTravVariants) / sizeof(kTravVariants[0]));

template <class Frames, bool kAnyHit, bool kStats, int kDepth, int kWaves>
__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(kAnyHit ? 1 : kWaves)))
void k_wf_extend(Frames fr, WfArrays a, const uint32_t* q, const uint32_t* n_in, uint32_t* zero0, uint32_t* zero1) {
    __shared__ uint32_t s_stack[kDepth * kWfBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *zero0 = 0u;        // next path queue
        *zero1 = 0u;        // shadow queue
        a.counts[4] = 0u;   // shadow fetch cursor
    }
    const uint32_t n = *n_in;
    uint32_t* const stack = s_stack + threadIdx.x;
    PathCounters pc;
    Traversal<kAnyHit, kStats, kDepth> tr;
    uint32_t s = 0, x = 0, y = 0;
    Rng4 rb{0u, 0u, 0u, 0u};
    bool active = false, exhausted = false;
    for (;;) {
        if (!exhausted) {
            uint32_t i = wave_fetch(!active, a.counts + 3, n, exhausted);
            if (i != kEmpty) {
                s = q[i];
                const FrameArgs& fa = fr.at(s, x, y);
                const uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
                float3 o = make_float3(float(a.ox[s]), float(a.oy[s]), float(a.oz[s]));
                float3 d = make_float3(float(a.dx[s]), float(a.dy[s]), float(a.dz[s]));
                if (kStats) { if (a.bounce[s] == 0u) pc.primary++; else pc.extension++; }
                rb = Rng4{x, y, fa.seed_b, kAnyHit ? a.wb[s] : 0u};
                tr.begin(fa.tlas, o, d, float(kTMin), float(a.tfar[s]), flags, stack, kWfBlock);
                active = true;
            }
        }
        if (!__ballot(active)) break;
        const uint32_t keep = exhausted ? 1u : a.refill_ext;
        do {
            if (active) {
                uint32_t xx, yy;
                const FrameArgs& fa = fr.at(s, xx, yy);
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &rb};
                if (tr.step(fa.tlas, RayFromSlot{a, s}, ah, stack, kWfBlock, pc.tr)) {
                    if (kAnyHit) a.wb[s] = rb.w;
                    a.h_inst[s] = tr.found ? tr.hit.inst : kEmpty;
                    a.h_prim[s] = tr.hit.prim;
                    a.h_rec[s] = tr.hit.rec;
                    a.h_t[s] = tr.hit.t;
                    a.h_u[s] = tr.hit.u;
                    a.h_v[s] = tr.hit.v;
                    active = false;
                }
            }
        } while (uint32_t(__popcll(__ballot(active))) >= keep);
    }
    if (kStats) flush_counters(fr.counters(), pc);
}

template <class Frames, bool kStats>
__global__ __launch_bounds__(kWfBlock) void k_wf_shade(Frames fr, WfArrays a, const uint32_t* q, const uint32_t* n_in) {
    const uint32_t n = *n_in;
    PathCounters pc;
    for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool active = i < n;
        BounceOut r;
        r.shadow = false;
        uint32_t s = 0;
        if (active) {
            s = q[i];
            uint32_t x, y;
            const FrameArgs& fa = fr.at(s, x, y);
            PathState p;
            load_full(a, s, p);
            HitInfo h;
            h.inst = a.h_inst[s];
            bool found = h.inst != kEmpty;
            h.prim = a.h_prim[s];
            h.rec = a.h_rec[s];
            h.t = a.h_t[s];
            h.u = a.h_u[s];
            h.v = a.h_v[s];
            shade_bounce<kStats>(fa, x, y, p, found, h, r, pc);
            store_full(a, s, p);
            if (fa.o.out && fa.o.mapping == RT_MAPPING_DISTANCE) a.last_t[s] = p.last_t;
            a.pwx[s] = r.pend_with.x; a.pwy[s] = r.pend_with.y; a.pwz[s] = r.pend_with.z;
            a.pox[s] = r.pend_without.x; a.poy[s] = r.pend_without.y; a.poz[s] = r.pend_without.z;
            a.flags[s] = (r.shadow ? kFlagShadow : 0u) | (r.cont ? kFlagCont : 0u);
        }
        bool want = active && r.shadow;
        uint32_t idx = wave_append(want, &a.counts[2]);
        if (want) {
            float3 so = to_f3(r.so), sd = to_f3(r.sd);
            a.sq_o[idx] = make_float4(so.x, so.y, so.z, float(r.sdist));
            a.sq_d[idx] = make_float4(sd.x, sd.y, sd.z, __uint_as_float(s));
            a.sq_wb[idx] = r.shadow_wb;
        }
    }
    if (kStats) flush_counters(fr.counters(), pc);
}

template <class Frames, bool kAnyHit, bool kStats, int kDepth, int kWaves>
__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(kAnyHit ? 1 : kWaves)))
void k_wf_shadow(Frames fr, WfArrays a) {
    __shared__ uint32_t s_stack[kDepth * kWfBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counts[3] = 0u;  // next extend's fetch cursor
    const uint32_t n = a.counts[2];
    uint32_t* const stack = s_stack + threadIdx.x;
    PathCounters pc;
    Traversal<kAnyHit, kStats, kDepth> tr;
    uint32_t s = 0, x = 0, y = 0;
    Rng4 shadow_rng{0u, 0u, 0u, 0u};
    bool active = false, exhausted = false;
    for (;;) {
        if (!exhausted) {
            uint32_t i = wave_fetch(!active, a.counts + 4, n, exhausted);
            if (i != kEmpty) {
                float4 o = a.sq_o[i], d = a.sq_d[i];
                s = __float_as_uint(d.w);
                const FrameArgs& fa = fr.at(s, x, y);
                const uint32_t flags = kShadowFlags | (fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : 0u);
                if (kStats) pc.shadow++;
                shadow_rng = Rng4{x, y, fa.seed_b, a.sq_wb[i]};
                tr.begin(fa.tlas, make_float3(o.x, o.y, o.z), make_float3(d.x, d.y, d.z), float(kShadowTMin), o.w,
                         flags, stack, kWfBlock);
                active = true;
            }
        }
        if (!__ballot(active)) break;
        const uint32_t keep = exhausted ? 1u : a.refill_sh;
        do {
            if (active) {
                uint32_t xx, yy;
                const FrameArgs& fa = fr.at(s, xx, yy);
                AlphaAnyHit sah{&fa.sv, &fa.tlas, &shadow_rng};
                if (tr.step(fa.tlas, RayFromF32{tr.o, tr.d}, sah, stack, kWfBlock, pc.trs)) {
                    a.unocc[s] = tr.found ? 0u : 1u;
                    active = false;
                }
            }
        } while (uint32_t(__popcll(__ballot(active))) >= keep);
    }
    if (kStats) flush_counters(fr.counters(), pc);
}

template <class Frames>
__global__ __launch_bounds__(kWfBlock) void k_wf_resolve(Frames fr, WfArrays a, const uint32_t* q, const uint32_t* n_in,
                                                          uint32_t* q_out, uint32_t* n_out) {
    const uint32_t n = *n_in;
    for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool alive = false;
        uint32_t s = 0;
        if (i < n) {
            s = q[i];
            uint32_t x, y;
            const FrameArgs& fa = fr.at(s, x, y);
            uint32_t f = a.flags[s];
            bool unocc = (f & kFlagShadow) && a.unocc[s];
            D3 rad = d3(a.rx[s], a.ry[s], a.rz[s]);
            D3 pend = unocc ? d3(a.pwx[s], a.pwy[s], a.pwz[s]) : d3(a.pox[s], a.poy[s], a.poz[s]);
            rad = rad + pend;
            a.rx[s] = rad.x; a.ry[s] = rad.y; a.rz[s] = rad.z;
            if (f & kFlagCont) {
                alive = true;
            } else {
                uint32_t sample = a.sample[s] + 1u;
                a.sample[s] = sample;
                if (sample < fa.samples) {
                    PathState p;
                    p.wa = a.wa[s];
                    p.lcg = a.lcg[s];
                    begin_sample(fa, x, y, p);
                    store_ray(a, s, p);
                    a.wa[s] = p.wa;
                    a.lcg[s] = p.lcg;
                    a.tx[s] = p.thr.x; a.ty[s] = p.thr.y; a.tz[s] = p.thr.z;
                    a.vol[s] = p.volume_dis;
                    a.tfar[s] = p.tfar;
                    a.bounce[s] = 0u;
                    alive = true;
                } else {
                    const bool dist = fa.o.out && fa.o.mapping == RT_MAPPING_DISTANCE;
                    write_accum(fa, x, y, rad, dist ? a.last_t[s] : 0.f);
                }
            }
        }
        uint32_t idx = wave_append(alive, n_out);
        if (alive) q_out[idx] = s;
    }
}

template <class T> size_t carve(char*& p, T*& out, size_t n) {
    size_t bytes = (n * sizeof(T) + 255) & ~size_t(255);
    out = reinterpret_cast<T*>(p);
    p += bytes;
    return bytes;
}

size_t layout(WfArrays& a, char* base, uint32_t n) {
    char* p = base;
    size_t total = 0;
    double** dptrs[] = {&a.ox, &a.oy, &a.oz, &a.dx, &a.dy, &a.dz, &a.tx, &a.ty, &a.tz, &a.rx, &a.ry, &a.rz,
                        &a.hx, &a.hy, &a.hz, &a.vol, &a.tfar, &a.pwx, &a.pwy, &a.pwz, &a.pox, &a.poy, &a.poz};
    for (double** d : dptrs) total += carve(p, *d, n);
    uint32_t** uptrs[] = {&a.wa, &a.wb, &a.lcg, &a.sample, &a.bounce, &a.h_inst, &a.h_prim, &a.h_rec, &a.flags,
                          &a.unocc, &a.sq_wb, &a.q0, &a.q1};
    for (uint32_t** u : uptrs) total += carve(p, *u, n);
    total += carve(p, a.h_t, n);
    total += carve(p, a.h_u, n);
    total += carve(p, a.h_v, n);
    total += carve(p, a.last_t, n);
    total += carve(p, a.sq_o, n);
    total += carve(p, a.sq_d, n);
    total += carve(p, a.counts, 8);
    return total;
}

__global__ void k_zero_counts(uint32_t* c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0u;
}

template <class Frames> struct Kernels {
    using ExtendFn = void (*)(Frames, WfArrays, const uint32_t*, const uint32_t*, uint32_t*, uint32_t*);
    using ShadowFn = void (*)(Frames, WfArrays);
    template <int V> static ExtendFn extend_v(bool ah, bool st) {
        constexpr int D = kTravVariants[V].depth, W = kTravVariants[V].waves;
        return ah ? (st ? k_wf_extend<Frames, true, true, D, W> : k_wf_extend<Frames, true, false, D, W>)
                  : (st ? k_wf_extend<Frames, false, true, D, W> : k_wf_extend<Frames, false, false, D, W>);
    }
    template <int V> static ShadowFn shadow_v(bool ah, bool st) {
        constexpr int D = kTravVariants[V].depth, W = kTravVariants[V].waves;
        return ah ? (st ? k_wf_shadow<Frames, true, true, D, W> : k_wf_shadow<Frames, true, false, D, W>)
                  : (st ? k_wf_shadow<Frames, false, true, D, W> : k_wf_shadow<Frames, false, false, D, W>);
    }
    static ExtendFn extend(int v, bool ah, bool st) {
        switch (v) {
            case 1: return extend_v<1>(ah, st);
            case 2: return extend_v<2>(ah, st);
            case 3: return extend_v<3>(ah, st);
            default: return extend_v<0>(ah, st);
        }
    }
    static ShadowFn shadow(int v, bool ah, bool st) {
        switch (v) {
            case 1: return shadow_v<1>(ah, st);
            case 2: return shadow_v<2>(ah, st);
            case 3: return shadow_v<3>(ah, st);
            default: return shadow_v<0>(ah, st);
        }
    }
};
static_assert(kNumTravVariants == 4, "Kernels::extend / shadow list every variant");

}  // namespace

int wavefront_num_variants() { return kNumTravVariants; }

struct WavefrontState {
    uint32_t cap = 0;
    char* mem = nullptr;
    WfArrays a{};
    char* batch_mem = nullptr;  // frame arguments + slot offsets of a batch
    size_t batch_cap = 0;
    bool use_graph = false;     // replay captured chunks (RT_OPTION_GRAPH)
    uint32_t refill_ext = 40, refill_sh = 1;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec = nullptr;
    int cus = 0;
    std::unordered_map<const void*, uint32_t> grids;  // persistent grid per kernel (occupancy x CUs)
    // grid of a persistent launch of kernel `f`: as many blocks as can be resident
    uint32_t grid_for(const void* f) {
        auto it = grids.find(f);
        if (it != grids.end()) return it->second;
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kWfBlock, 0) != hipSuccess || per_cu < 1) per_cu = 4;
        uint32_t g = uint32_t(std::max(cus, 1)) * uint32_t(per_cu);
        grids.emplace(f, g);
        return g;
    }
};

WavefrontState* wavefront_create() { return new WavefrontState(); }

void wavefront_set_graph(WavefrontState* s, bool on) { s->use_graph = on; }

void wavefront_set_refill(WavefrontState* s, uint32_t extend, uint32_t shadow) {
    s->refill_ext = std::min(std::max(extend, 1u), 64u);
    s->refill_sh = std::min(std::max(shadow, 1u), 64u);
    s->a.refill_ext = s->refill_ext;
    s->a.refill_sh = s->refill_sh;
}

void wavefront_destroy(WavefrontState* s) {
    if (!s) return;
    if (s->mem) (void)hipFree(s->mem);
    if (s->batch_mem) (void)hipFree(s->batch_mem);
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
    delete s;
}

namespace {

// Grows the path state to `nslots` slots (one per pixel).
int ensure_slots(WavefrontState* ws, uint32_t nslots, std::string* err) {
    if (nslots <= ws->cap) return 0;
    if (ws->mem) (void)hipFree(ws->mem);
    ws->mem = nullptr;
    ws->cap = 0;
    WfArrays probe{};
    size_t bytes = layout(probe, nullptr, nslots) + 4096;
    hipError_t e = hipMalloc(&ws->mem, bytes);
    if (e != hipSuccess) {
        *err = "cannot allocate the path state";
        return int(e);
    }
    ws->cap = nslots;
    layout(ws->a, ws->mem, nslots);
    ws->a.refill_ext = ws->refill_ext;
    ws->a.refill_sh = ws->refill_sh;
    return 0;
}

// The queue loop over `nslots` slots whose frames `fr` gives; every path ends within
// max_samples * max_bounces bounces.
template <class Frames>
int run_wavefront(WavefrontState* ws, const Frames& fr, uint32_t nslots, uint64_t max_steps, hipStream_t st,
                  bool any_hit, bool stats, int variant, StageTimer* tm, std::string* err) {
    if (ws->cus == 0) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        ws->cus = std::max(cus, 1);
    }
    const WfArrays& a = ws->a;
    using K = Kernels<Frames>;
    const typename K::ExtendFn k_extend = K::extend(variant, any_hit, stats);
    const typename K::ShadowFn k_shadow = K::shadow(variant, any_hit, stats);
    const void* k_shade = stats ? (const void*)k_wf_shade<Frames, true> : (const void*)k_wf_shade<Frames, false>;
    const uint32_t g_extend = ws->grid_for((const void*)k_extend);
    const uint32_t g_shadow = ws->grid_for((const void*)k_shadow);
    const uint32_t g_shade = ws->grid_for(k_shade);
    const uint32_t g_resolve = ws->grid_for((const void*)k_wf_resolve<Frames>);
    hipError_t e = hipSuccess;
    // stage marks: 0 extend, 1 shade, 2 shadow, 3 resolve (init counts as resolve)
    auto mark = [&](int stage) {
        if (tm && e == hipSuccess) e = tm->mark(st, stage);
    };
    mark(-1);
    hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, st, a.counts);
    hipLaunchKernelGGL(k_wf_init<Frames>, dim3(std::min(g_resolve, (nslots + kWfBlock - 1) / kWfBlock)), dim3(kWfBlock),
                       0, st, fr, a, nslots);
    mark(3);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        *err = "launch k_wf_init";
        return int(e);
    }
    const uint64_t max_iter = max_steps + kCheckEvery;
    // one iteration: extend, shade, shadow, resolve, with the path queues of parity k
    auto iteration = [&](uint64_t k, hipStream_t s, bool marks) {
        uint32_t* q_in = (k & 1) ? a.q1 : a.q0;
        uint32_t* q_out = (k & 1) ? a.q0 : a.q1;
        uint32_t* n_in = a.counts + (k & 1);
        uint32_t* n_out = a.counts + ((k + 1) & 1);
        hipLaunchKernelGGL(k_extend, dim3(g_extend), dim3(kWfBlock), 0, s, fr, a, q_in, n_in, n_out, a.counts + 2);
        if (marks) mark(0);
        if (stats) hipLaunchKernelGGL((k_wf_shade<Frames, true>), dim3(g_shade), dim3(kWfBlock), 0, s, fr, a, q_in, n_in);
        else hipLaunchKernelGGL((k_wf_shade<Frames, false>), dim3(g_shade), dim3(kWfBlock), 0, s, fr, a, q_in, n_in);
        if (marks) mark(1);
        hipLaunchKernelGGL(k_shadow, dim3(g_shadow), dim3(kWfBlock), 0, s, fr, a);
        if (marks) mark(2);
        hipLaunchKernelGGL(k_wf_resolve<Frames>, dim3(g_resolve), dim3(kWfBlock), 0, s, fr, a, q_in, n_in, q_out, n_out);
        if (marks) mark(3);
    };
    // the live-path count after an even number of iterations sits in counts[0]
    auto live_paths = [&](uint32_t& live) -> hipError_t {
        hipError_t r = hipMemcpyAsync(&live, a.counts, sizeof(uint32_t), hipMemcpyDeviceToHost, st);
        if (r == hipSuccess) r = hipStreamSynchronize(st);
        return r;
    };
    static_assert(kCheckEvery % 2 == 0, "a chunk ends on queue 0");
    if (ws->use_graph && !tm) {
        // the kCheckEvery iterations of a chunk captured once as a HIP graph and
        // replayed until no path is alive: one launch per chunk instead of 32
        if (!ws->cap_stream) {
            e = hipStreamCreateWithFlags(&ws->cap_stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                *err = "capture stream";
                return int(e);
            }
        }
        e = hipStreamBeginCapture(ws->cap_stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            *err = "hipStreamBeginCapture";
            return int(e);
        }
        for (uint64_t k = 0; k < kCheckEvery; k++) iteration(k, ws->cap_stream, false);
        hipGraph_t graph = nullptr;
        hipError_t launch_err = hipGetLastError();
        e = hipStreamEndCapture(ws->cap_stream, &graph);
        if (e == hipSuccess) e = launch_err;
        if (e == hipSuccess && ws->exec) {
            hipGraphNode_t bad = nullptr;
            hipGraphExecUpdateResult res;
            if (hipGraphExecUpdate(ws->exec, graph, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                (void)hipGetLastError();
                (void)hipGraphExecDestroy(ws->exec);
                ws->exec = nullptr;
            }
        }
        if (e == hipSuccess && !ws->exec) e = hipGraphInstantiate(&ws->exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            *err = "capture of the wavefront chunk";
            return int(e);
        }
        for (uint64_t k = 0;; k += kCheckEvery) {
            e = hipGraphLaunch(ws->exec, st);
            uint32_t live = 0;
            if (e == hipSuccess) e = live_paths(live);
            if (e != hipSuccess) {
                *err = "replay of the wavefront chunk";
                return int(e);
            }
            if (live == 0) break;
            if (k > max_iter) {
                *err = "paths did not terminate";
                return int(hipErrorUnknown);
            }
        }
        return 0;
    }
    for (uint64_t k = 0;; k++) {
        iteration(k, st, true);
        if (e == hipSuccess) e = hipGetLastError();
        if (e != hipSuccess) {
            *err = "launch of a wavefront kernel";
            return int(e);
        }
        if ((k + 1) % kCheckEvery == 0) {
            uint32_t live = 0;
            e = live_paths(live);
            if (e != hipSuccess) {
                *err = "reading the queue length";
                return int(e);
            }
            if (live == 0) break;
        }
        if (k > max_iter) {
            *err = "paths did not terminate";
            return int(hipErrorUnknown);
        }
    }
    return 0;
}

uint32_t frame_slots(const FrameArgs& fa) { return ((fa.width + 7) / 8) * ((fa.height + 7) / 8) * 64u; }

}  // namespace

int launch_wavefront(WavefrontState* ws, const FrameArgs& fa, hipStream_t st, bool any_hit, bool stats, int variant,
                     StageTimer* tm, std::string* err) {
    if (variant < 0 || variant >= kNumTravVariants) {
        *err = "unknown traversal variant";
        return kWavefrontUnsupported;
    }
    if (fa.samples == 0u || fa.bounces == 0u) return launch_megakernel(fa, st, any_hit, stats, tm);
    const uint32_t nslots = frame_slots(fa);
    int rc = ensure_slots(ws, nslots, err);
    if (rc) return rc;
    return run_wavefront(ws, OneFrame{fa}, nslots, uint64_t(fa.samples) * fa.bounces, st, any_hit, stats, variant, tm,
                         err);
}

int launch_wavefront_batch(WavefrontState* ws, const FrameArgs* frames, uint32_t count, hipStream_t st, bool any_hit,
                           bool stats, int variant, StageTimer* tm, std::string* err) {
    if (variant < 0 || variant >= kNumTravVariants) {
        *err = "unknown traversal variant";
        return kWavefrontUnsupported;
    }
    std::vector<uint32_t> base(count + 1, 0u);
    uint64_t total = 0, steps = 0;
    for (uint32_t i = 0; i < count; i++) {
        base[i] = uint32_t(total);
        if (frames[i].samples && frames[i].bounces) total += frame_slots(frames[i]);
        steps = std::max(steps, uint64_t(frames[i].samples) * frames[i].bounces);
        if (total >= (1ull << 31)) {
            *err = "the batch exceeds 2^31 path slots";
            return kWavefrontUnsupported;
        }
    }
    base[count] = uint32_t(total);
    if (total == 0) return 0;
    int rc = ensure_slots(ws, uint32_t(total), err);
    if (rc) return rc;
    // frame arguments and slot offsets in device memory
    const size_t fbytes = sizeof(FrameArgs) * count, bbytes = sizeof(uint32_t) * (count + 1);
    if (fbytes + bbytes > ws->batch_cap) {
        if (ws->batch_mem) (void)hipFree(ws->batch_mem);
        ws->batch_mem = nullptr;
        ws->batch_cap = 0;
        hipError_t e = hipMalloc(&ws->batch_mem, fbytes + bbytes + 256);
        if (e != hipSuccess) {
            *err = "cannot allocate the batch table";
            return int(e);
        }
        ws->batch_cap = fbytes + bbytes;
    }
    FrameArgs* d_frames = reinterpret_cast<FrameArgs*>(ws->batch_mem);
    uint32_t* d_base = reinterpret_cast<uint32_t*>(ws->batch_mem + ((fbytes + 255) & ~size_t(255)));
    // frames with no samples or bounces own no slots: the image reset still applies
    hipError_t e = hipMemcpyAsync(d_frames, frames, fbytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_base, base.data(), bbytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        *err = "copy of the batch table";
        return int(e);
    }
    ManyFrames fr{d_frames, d_base, count, frames[0].counters};
    rc = run_wavefront(ws, fr, uint32_t(total), steps, st, any_hit, stats, variant, tm, err);
    // the host arrays the copies read must outlive them
    if (rc == 0) {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = int(e);
    }
    return rc;
}

}  // namespace rt
