// wavefront.hip — the wavefront form of the trace (RT_MODE_WAVEFRONT).
//
// The raygen loop of the reference (crates/examples/gltf_viewer/shaders/RayTracing.rgen)
// is split into queue kernels that each do one stage for every live path:
//   init   path state of every pixel, camera ray of sample 0            -> path queue
//   trace  traversal of the queue's extension rays (closest hit) and of the
//          previous bounce's shadow rays (terminate on first hit), served by one
//          persistent grid from one fetch cursor
//   shade  the previous bounce's NEE radiance (its shadow ray's result), then the
//          closest-hit / miss shaders in fp64, NEE set-up (shadow ray), next ray;
//          a path that ended starts its pixel's next sample or retires the pixel
//          into the accumulation image                                   -> path queue
// Adding a bounce's NEE term one iteration late (once its shadow ray is traced)
// changes nothing in the order of the radiance sums or of the RNG draws, so the image
// is the megakernel's bit for bit (both call the same device functions in path.h).
// One slot per pixel keeps each pixel's RNG streams sequential.  Queues are
// compacted with wave ballots (one atomic per wave).  Kernels are persistent
// grid-stride loops that read their queue lengths from device memory, so the host
// only checks for completion every few iterations.
#include <algorithm>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "path.h"
#include "wavefront.h"

namespace rt {

namespace {

constexpr int kWfBlock = 256;
constexpr uint32_t kCheckEvery = 8;
constexpr int kMaxStreams = 4;  // slot sets traced concurrently (RT_OPTION_WAVEFRONT_STREAMS)

struct WfArrays {
    double *ox, *oy, *oz, *dx, *dy, *dz;  // ray
    double *tx, *ty, *tz, *rx, *ry, *rz;  // throughput, radiance
    double *hx, *hy, *hz, *vol, *tfar;    // Ray.hitPoint, Ray.volume_dis, tMax
    double *pwx, *pwy, *pwz, *pox, *poy, *poz;  // pending radiance with / without NEE
    uint32_t *wa, *wb, *lcg, *sample, *bounce;
    uint32_t *h_inst, *h_prim, *h_rec;
    float *h_t, *h_u, *h_v;
    float* last_t;     // payload t after the last trace (DISTANCE mapping only)
    uint32_t *flags;   // kFlagShadow
    uint32_t *unocc;   // shadow result
    float4 *sq_o, *sq_d;  // pending shadow ray per slot: origin + tmax, direction
    uint32_t* sq_wb;      // its stream-B counter
    uint32_t *q0, *q1;    // path queues (ping-pong)
    uint32_t *qs0, *qs1;  // shadow-ray queues (ping-pong): slot | kRetireBit
    uint32_t* counts;  // [0], [1]: path queue lengths, [2], [3]: shadow queue lengths
                       // (ping-pong), [4]: trace fetch cursor (8 per slot set)
    uint32_t refill_ext, refill_sh;  // busy lanes below which a wave takes new rays (1..64)
};

// kFlagShadow: the slot's last shaded bounce has a shadow ray whose result (unocc)
// decides its NEE radiance (pw* or po*), added by the next shade.  A shadow queue
// entry with kRetireBit: that bounce ended the pixel's last sample, and the slot
// has no path queue entry; the next shade adds the radiance and writes the pixel.
enum : uint32_t { kFlagShadow = 1u };
constexpr uint32_t kRetireBit = 0x80000000u;
constexpr int kCursor = 4;  // counts[kCursor]: the trace kernel's fetch cursor

__host__ __device__ inline uint32_t* shadow_count(const WfArrays& a, uint64_t parity) { return a.counts + 2 + (parity & 1); }

// Slot -> pixel: slots come in 8x8 pixel tiles (consecutive slots are neighbouring
// pixels); a set of slots traced on one stream holds every `step`-th tile from `first`.
__device__ inline void slot_pixel(const FrameArgs& fa, uint32_t s, uint32_t step, uint32_t first, uint32_t& x,
                                  uint32_t& y) {
    uint32_t tiles_x = (fa.width + 7) / 8;
    uint32_t tile = (s >> 6) * step + first, within = s & 63u;
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
    uint32_t tile_step = 1, tile_first = 0;  // the frame's tiles this slot set holds
    __device__ uint32_t frame_of(uint32_t) const { return 0u; }
    __device__ const FrameArgs& frame(uint32_t) const { return fa; }
    __device__ const FrameArgs& at(uint32_t s, uint32_t& x, uint32_t& y) const {
        slot_pixel(fa, s, tile_step, tile_first, x, y);
        return fa;
    }
    __device__ unsigned long long* counters() const { return fa.counters; }
};
struct ManyFrames {
    const FrameArgs* fas;
    const uint32_t* base;  // n + 1 slot offsets
    uint32_t n;
    unsigned long long* cnt;
    const uint32_t* slot_frame;  // frame of every slot (k_slot_frames)
    __device__ uint32_t frame_of(uint32_t s) const { return slot_frame[s]; }
    __device__ const FrameArgs& frame(uint32_t f) const { return fas[f]; }
    __device__ const FrameArgs& at(uint32_t s, uint32_t& x, uint32_t& y) const {
        const uint32_t lo = frame_of(s);
        const FrameArgs& fa = fas[lo];
        slot_pixel(fa, s - base[lo], 1u, 0u, x, y);
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
            a.flags[s] = 0u;
        }
        uint32_t idx = wave_append(valid, &a.counts[0]);
        if (valid) q[idx] = s;
    }
}

// Traversal variants of the trace kernel: LDS stack depth and the waves per SIMD the
// register allocation targets (the compiler may spill above its natural occupancy).
// Selected per context (RT_OPTION_TRAVERSAL_VARIANT).
struct TravVariant {
    int depth, waves;
};
constexpr TravVariant kTravVariants[] = {{32, 3}, {32, 4}, {32, 5}, {64, 2}};
constexpr int kNumTravVariants = int(sizeof(kTravVariants) / sizeof(kTravVariants[0]));
// The any-hit kernels (alpha-tested geometry) hold more state: a variant's target
// maps to 3 waves (no spills) or 2.
constexpr int anyhit_waves(int w) { return w < 3 ? w : 3; }

// The ray a trace-kernel lane works on, in fp64 for instance entry, re-read from the
// slot (keeps it out of the traversal's registers): an extension ray from the fp64
// path state, a shadow ray from its fp32 record.
struct RayFromSlotOrShadow {
    const WfArrays& a;
    uint32_t s;
    bool shadow;
    __device__ void get(D3& O, D3& Dd) const {
        if (shadow) {
            const float4 o = a.sq_o[s], d = a.sq_d[s];
            O = d3(make_float3(o.x, o.y, o.z));
            Dd = d3(make_float3(d.x, d.y, d.z));
        } else {
            O = d3(a.ox[s], a.oy[s], a.oz[s]);
            Dd = d3(a.dx[s], a.dy[s], a.dz[s]);
        }
    }
};

// Iteration k's traversal: the extension rays of path queue q (n_e entries) then the
// shadow rays of queue qs (n_s entries, from the previous shade), as one index
// range served by one fetch cursor.  Zeroes the queue lengths the next shade
// appends to.
template <class Frames, bool kAnyHit, bool kStats, int kDepth, int kWaves>
__global__ __launch_bounds__(kWfBlock) __attribute__((amdgpu_waves_per_eu(kAnyHit ? anyhit_waves(kWaves) : kWaves)))
void k_wf_trace(Frames fr, WfArrays a, const uint32_t* q, const uint32_t* n_e_in, const uint32_t* qs,
                const uint32_t* n_s_in, uint32_t* zero_e, uint32_t* zero_s) {
    __shared__ uint32_t s_stack[kDepth * kWfBlock];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *zero_e = 0u;  // next path queue
        *zero_s = 0u;  // next shadow queue
    }
    const uint32_t n_e = *n_e_in, n = n_e + *n_s_in;
    uint32_t* const stack = s_stack + threadIdx.x;
    PathCounters pc;
    TraceStats ts;
    Traversal<kAnyHit, kStats, kDepth> tr;
    uint32_t s = 0, x = 0, y = 0, f = 0;  // f: the slot's frame (a batch's frames)
    Rng4 rng{0u, 0u, 0u, 0u};
    bool active = false, exhausted = false, shadow = false;
    for (;;) {
        if (!exhausted) {
            uint32_t i = wave_fetch(!active, a.counts + kCursor, n, exhausted);
            if (i != kEmpty) {
                shadow = i >= n_e;
                s = shadow ? (qs[i - n_e] & ~kRetireBit) : q[i];
                const FrameArgs& fa = fr.at(s, x, y);
                f = fr.frame_of(s);
                const uint32_t opq = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
                if (shadow) {
                    // castShadowRay, rchit:35-59: tMin 0.1, tMax the light distance
                    const float4 o = a.sq_o[s], d = a.sq_d[s];
                    if (kStats) pc.shadow++;
                    rng = Rng4{x, y, fa.seed_b, a.sq_wb[s]};
                    tr.begin(fa.tlas, make_float3(o.x, o.y, o.z), make_float3(d.x, d.y, d.z), float(kShadowTMin), o.w,
                             kShadowFlags | opq, stack, kWfBlock);
                } else {
                    float3 o = make_float3(float(a.ox[s]), float(a.oy[s]), float(a.oz[s]));
                    float3 d = make_float3(float(a.dx[s]), float(a.dy[s]), float(a.dz[s]));
                    if (kStats) { if (a.bounce[s] == 0u) pc.primary++; else pc.extension++; }
                    rng = Rng4{x, y, fa.seed_b, kAnyHit ? a.wb[s] : 0u};
                    tr.begin(fa.tlas, o, d, float(kTMin), float(a.tfar[s]), opq, stack, kWfBlock);
                }
                active = true;
            }
        }
        if (!__ballot(active)) break;
        // the wave steps its rays while at least refill_ext of its lanes are busy, then
        // takes new rays: traversal lengths vary by an order of magnitude, so waiting
        // for all 64 idles many lanes, and each refill costs a ballot and an atomic
        const uint32_t keep = exhausted ? 1u : a.refill_ext;
        do {
            if (active) {
                const FrameArgs& fa = fr.frame(f);  // found once per ray, not per step
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &rng};
                if (tr.step(fa.tlas, RayFromSlotOrShadow{a, s, shadow}, ah, stack, kWfBlock, ts)) {
                    if (shadow) {
                        a.unocc[s] = tr.found ? 0u : 1u;
                    } else {
                        if (kAnyHit) a.wb[s] = rng.w;
                        a.h_inst[s] = tr.found ? tr.hit.inst : kEmpty;
                        a.h_prim[s] = tr.hit.prim;
                        a.h_rec[s] = tr.hit.rec;
                        a.h_t[s] = tr.hit.t;
                        a.h_u[s] = tr.hit.u;
                        a.h_v[s] = tr.hit.v;
                    }
                    if (kStats) {
                        TraceStats& to = shadow ? pc.trs : pc.tr;
                        to.nodes += ts.nodes;
                        to.tris += ts.tris;
                        to.instances += ts.instances;
                        to.anyhits += ts.anyhits;
                        to.overflows += ts.overflows;
                        ts = TraceStats{};
                    }
                    active = false;
                }
            }
        } while (uint32_t(__popcll(__ballot(active))) >= keep);
    }
    if (kStats) flush_counters(fr.counters(), pc);
}

// Adds the NEE radiance of the slot's last shaded bounce, now that its shadow ray
// has been traced (the radiance update of rgen:98-103 for that bounce).
__device__ inline void add_pending(const WfArrays& a, uint32_t s, uint32_t f, D3& rad) {
    if (f & kFlagShadow) {
        rad = rad + (a.unocc[s] ? d3(a.pwx[s], a.pwy[s], a.pwz[s]) : d3(a.pox[s], a.poy[s], a.poz[s]));
    }
}

// Iteration k's shading of path queue q: pending radiance, then the shaders for the
// slot's traced ray; appends the next path queue and the next shadow-ray queue.
// The retiring entries of the shadow queue the trace just served (qs_in) get their
// last radiance and are written.
template <class Frames, bool kStats>
__global__ __launch_bounds__(kWfBlock) void k_wf_shade(Frames fr, WfArrays a, const uint32_t* q, const uint32_t* n_in,
                                                        const uint32_t* qs_in, const uint32_t* ns_in, uint32_t* q_out,
                                                        uint32_t* n_out, uint32_t* qs_out, uint32_t* ns_out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) a.counts[kCursor] = 0u;  // next trace's fetch cursor
    const uint32_t n_e = *n_in, n = n_e + *ns_in;
    PathCounters pc;
    for (uint32_t base = blockIdx.x * kWfBlock; base < n; base += gridDim.x * kWfBlock) {
        uint32_t i = base + threadIdx.x;
        bool next = false, want_shadow = false, retire = false;
        uint32_t s = 0;
        if (i >= n_e && i < n) {
            const uint32_t entry = qs_in[i - n_e];
            if (entry & kRetireBit) {
                // the pixel's last bounce: its radiance, then the pixel
                s = entry & ~kRetireBit;
                uint32_t x, y;
                const FrameArgs& fa = fr.at(s, x, y);
                D3 rad = d3(a.rx[s], a.ry[s], a.rz[s]);
                add_pending(a, s, kFlagShadow, rad);
                const bool dist = fa.o.out && fa.o.mapping == RT_MAPPING_DISTANCE;
                write_accum(fa, x, y, rad, dist ? a.last_t[s] : 0.f);
            }
        } else if (i < n_e) {
            s = q[i];
            uint32_t x, y;
            const FrameArgs& fa = fr.at(s, x, y);
            const uint32_t f = a.flags[s];
            PathState p;
            load_full(a, s, p);
            add_pending(a, s, f, p.rad);
            HitInfo h;
            h.inst = a.h_inst[s];
            bool found = h.inst != kEmpty;
            h.prim = a.h_prim[s];
            h.rec = a.h_rec[s];
            h.t = a.h_t[s];
            h.u = a.h_u[s];
            h.v = a.h_v[s];
            BounceOut r;
            shade_bounce<kStats>(fa, x, y, p, found, h, r, pc);
            uint32_t nf = 0u;
            if (r.shadow) {
                // its radiance waits for the shadow ray (next trace, next shade)
                a.pwx[s] = r.pend_with.x; a.pwy[s] = r.pend_with.y; a.pwz[s] = r.pend_with.z;
                a.pox[s] = r.pend_without.x; a.poy[s] = r.pend_without.y; a.poz[s] = r.pend_without.z;
                const float3 so = to_f3(r.so), sd = to_f3(r.sd);
                a.sq_o[s] = make_float4(so.x, so.y, so.z, float(r.sdist));
                a.sq_d[s] = make_float4(sd.x, sd.y, sd.z, 0.f);
                a.sq_wb[s] = r.shadow_wb;
                nf = kFlagShadow;
                want_shadow = true;
            } else {
                p.rad = p.rad + r.pend_without;
            }
            if (fa.o.out && fa.o.mapping == RT_MAPPING_DISTANCE) a.last_t[s] = p.last_t;
            if (r.cont) {
                next = true;
            } else if (++p.sample < fa.samples) {
                begin_sample(fa, x, y, p);  // the pixel's next sample
                next = true;
            } else if (nf) {
                retire = true;  // written once its shadow ray is traced
            } else {
                write_accum(fa, x, y, p.rad, p.last_t);
            }
            store_full(a, s, p);
            a.flags[s] = nf;
        }
        uint32_t idx = wave_append(next, n_out);
        if (next) q_out[idx] = s;
        uint32_t sidx = wave_append(want_shadow, ns_out);
        if (want_shadow) qs_out[sidx] = s | (retire ? kRetireBit : 0u);
    }
    if (kStats) flush_counters(fr.counters(), pc);
}

// The tail of a trace: once few paths are alive, each iteration's launches last as
// long as their slowest ray while most of the GPU idles.  This kernel takes, after a
// shade, the paths of path queue q and the retiring slots of shadow queue qs and runs
// each one, in one lane, through its pending shadow ray and the rest of its pixel's
// samples (the megakernel's loop from the slot's state), then writes the pixel.  Its
// lanes are PathLanes: traversals stepped, the wave's shading done together, and a
// lane that finishes its pixel takes the next queue entry (one atomic per wave, the
// trace cursor, which the last shade cleared).
constexpr int kFinishBlock = 128;

template <class Frames, bool kAnyHit, bool kStats>
__global__ __launch_bounds__(kFinishBlock) void k_wf_finish(Frames fr, WfArrays a, const uint32_t* q,
                                                            const uint32_t* n_e_in, const uint32_t* qs,
                                                            const uint32_t* n_s_in) {
    __shared__ uint32_t s_stack[kStackDepth * kFinishBlock];
    uint32_t* const stack = s_stack + threadIdx.x;
    const uint32_t n_e = *n_e_in, n = n_e + *n_s_in;
    PathCounters pc;
    PathLane<kAnyHit, kStats> ln;
    uint32_t f = 0;  // the lane's frame (a batch's frames)
    bool exhausted = false;
    for (;;) {
        // lanes without a pixel take the next entry of q, then of qs (its retiring ones)
        if (!exhausted) {
            const uint32_t i = wave_fetch(ln.state == kLaneIdle, a.counts + kCursor, n, exhausted);
            if (i != kEmpty) {
                const uint32_t entry = i < n_e ? q[i] : qs[i - n_e];
                if (i < n_e || (entry & kRetireBit)) {  // a shadow entry without the bit has a path entry too
                    const uint32_t s = entry & ~kRetireBit;
                    const FrameArgs& fa = fr.at(s, ln.x, ln.y);
                    f = fr.frame_of(s);
                    load_full(a, s, ln.p);
                    ln.p.last_t = (fa.o.out && fa.o.mapping == RT_MAPPING_DISTANCE) ? a.last_t[s] : 0.f;
                    const uint32_t then = i >= n_e ? kAfterDone : kAfterExt;
                    if (a.flags[s] & kFlagShadow) {
                        // the last shaded bounce's shadow ray, as the trace kernel would take it
                        const float4 o = a.sq_o[s], d = a.sq_d[s];
                        ln.r.pend_with = d3(a.pwx[s], a.pwy[s], a.pwz[s]);
                        ln.r.pend_without = d3(a.pox[s], a.poy[s], a.poz[s]);
                        ln.start_shadow(fa, make_float3(o.x, o.y, o.z), make_float3(d.x, d.y, d.z), o.w, a.sq_wb[s],
                                        then, stack, kFinishBlock, pc);
                    } else if (then == kAfterExt) {
                        ln.start_ext(fa, stack, kFinishBlock, pc);
                    } else {
                        ln.state = kLaneDone;
                    }
                }
            }
        }
        const uint64_t holding = __ballot(ln.state != kLaneIdle);
        if (!holding && exhausted) break;
        const uint32_t keep = (uint32_t(__popcll(holding)) + 1u) / 2u;
        while (keep && uint32_t(__popcll(__ballot(ln.tracing()))) >= keep) {
            if (ln.tracing()) {
                const FrameArgs& fa = fr.frame(f);
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &ln.rng};
                ln.step(fa, ah, stack, kFinishBlock, pc);
            }
        }
        if (ln.state != kLaneIdle && !ln.tracing()) {
            const FrameArgs& fa = fr.frame(f);
            ln.advance(fa, stack, kFinishBlock, pc);
            if (ln.state == kLaneDone) {
                write_accum(fa, ln.x, ln.y, ln.p.rad, ln.p.last_t);
                ln.state = kLaneIdle;
            }
        }
    }
    if (kStats) flush_counters(fr.counters(), pc);
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
                          &a.unocc, &a.sq_wb, &a.q0, &a.q1, &a.qs0, &a.qs1};
    for (uint32_t** u : uptrs) total += carve(p, *u, n);
    total += carve(p, a.h_t, n);
    total += carve(p, a.h_u, n);
    total += carve(p, a.h_v, n);
    total += carve(p, a.last_t, n);
    total += carve(p, a.sq_o, n);
    total += carve(p, a.sq_d, n);
    total += carve(p, a.counts, 8 * kMaxStreams);
    return total;
}

// The arrays of the slots [base, base + n) of a slot set with its own counters: the
// per-slot arrays and both queues are offset, so the set's kernels index them from 0.
WfArrays slice(const WfArrays& a, uint32_t base, uint32_t* counts) {
    WfArrays r = a;
    double** d[] = {&r.ox, &r.oy, &r.oz, &r.dx, &r.dy, &r.dz, &r.tx, &r.ty, &r.tz, &r.rx, &r.ry, &r.rz,
                    &r.hx, &r.hy, &r.hz, &r.vol, &r.tfar, &r.pwx, &r.pwy, &r.pwz, &r.pox, &r.poy, &r.poz};
    for (double** x : d) *x += base;
    uint32_t** u[] = {&r.wa, &r.wb, &r.lcg, &r.sample, &r.bounce, &r.h_inst, &r.h_prim, &r.h_rec, &r.flags,
                      &r.unocc, &r.sq_wb, &r.q0, &r.q1, &r.qs0, &r.qs1};
    for (uint32_t** x : u) *x += base;
    r.h_t += base;
    r.h_u += base;
    r.h_v += base;
    r.last_t += base;
    r.sq_o += base;
    r.sq_d += base;
    r.counts = counts;
    return r;
}

// The frame of every slot of a batch's slot set: the frame i with base[i] <= s < base[i + 1].
__global__ void k_slot_frames(const uint32_t* base, uint32_t n, uint32_t nslots, uint32_t* out) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t m = (lo + hi) >> 1;
        if (base[m] <= s) lo = m; else hi = m;
    }
    out[s] = lo;
}

__global__ void k_zero_counts(uint32_t* c) {
    if (threadIdx.x < 8) c[threadIdx.x] = 0u;
}

template <class Frames> struct Kernels {
    using TraceFn = void (*)(Frames, WfArrays, const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*,
                             uint32_t*, uint32_t*);
    using FinishFn = void (*)(Frames, WfArrays, const uint32_t*, const uint32_t*, const uint32_t*, const uint32_t*);
    static FinishFn finish(bool ah, bool st) {
        return ah ? (st ? k_wf_finish<Frames, true, true> : k_wf_finish<Frames, true, false>)
                  : (st ? k_wf_finish<Frames, false, true> : k_wf_finish<Frames, false, false>);
    }
    template <int V> static TraceFn trace_v(bool ah, bool st) {
        constexpr int D = kTravVariants[V].depth, W = kTravVariants[V].waves;
        return ah ? (st ? k_wf_trace<Frames, true, true, D, W> : k_wf_trace<Frames, true, false, D, W>)
                  : (st ? k_wf_trace<Frames, false, true, D, W> : k_wf_trace<Frames, false, false, D, W>);
    }
    static TraceFn trace(int v, bool ah, bool st) {
        switch (v) {
            case 1: return trace_v<1>(ah, st);
            case 2: return trace_v<2>(ah, st);
            case 3: return trace_v<3>(ah, st);
            default: return trace_v<0>(ah, st);
        }
    }
};
static_assert(kNumTravVariants == 4, "Kernels::trace lists every variant");

}  // namespace

int wavefront_num_variants() { return kNumTravVariants; }

struct WavefrontState {
    uint32_t cap = 0;
    char* mem = nullptr;
    WfArrays a{};
    char* batch_mem = nullptr;  // frame arguments + slot offsets of a batch
    size_t batch_cap = 0;
    bool use_graph = false;     // replay captured chunks (RT_OPTION_GRAPH)
    uint32_t refill_ext = 16, refill_sh = 1;
    int streams = 2;            // slot sets traced concurrently (RT_OPTION_WAVEFRONT_STREAMS)
    uint32_t finish_below = 0;  // live paths below which k_wf_finish ends a set (RT_OPTION_WAVEFRONT_FINISH)
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t exec[kMaxStreams] = {};
    hipStream_t set_stream[kMaxStreams] = {};  // streams of slot sets 1.. (set 0 uses the caller's)
    hipEvent_t fork = nullptr, join[kMaxStreams] = {}, counted[kMaxStreams][2] = {};
    uint32_t* host_live = nullptr;  // pinned: [set][parity][4] queue lengths read back
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

void wavefront_set_streams(WavefrontState* s, int n) { s->streams = std::min(std::max(n, 1), kMaxStreams); }

void wavefront_set_finish(WavefrontState* s, uint32_t live_paths) { s->finish_below = live_paths; }

int wavefront_max_streams() { return kMaxStreams; }

void wavefront_destroy(WavefrontState* s) {
    if (!s) return;
    if (s->mem) (void)hipFree(s->mem);
    if (s->batch_mem) (void)hipFree(s->batch_mem);
    if (s->host_live) (void)hipHostFree(s->host_live);
    for (int g = 0; g < kMaxStreams; g++) {
        if (s->exec[g]) (void)hipGraphExecDestroy(s->exec[g]);
        if (s->set_stream[g]) (void)hipStreamDestroy(s->set_stream[g]);
        if (s->join[g]) (void)hipEventDestroy(s->join[g]);
        for (int k = 0; k < 2; k++)
            if (s->counted[g][k]) (void)hipEventDestroy(s->counted[g][k]);
    }
    if (s->fork) (void)hipEventDestroy(s->fork);
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

// Streams, events and pinned read-back words for `n` slot sets.
int ensure_streams(WavefrontState* ws, int n, std::string* err) {
    hipError_t e = hipSuccess;
    if (!ws->host_live) e = hipHostMalloc((void**)&ws->host_live, sizeof(uint32_t) * 8 * kMaxStreams, hipHostMallocDefault);
    if (e == hipSuccess && !ws->fork) e = hipEventCreateWithFlags(&ws->fork, hipEventDisableTiming);
    for (int g = 0; g < n && e == hipSuccess; g++) {
        if (g > 0 && !ws->set_stream[g]) e = hipStreamCreateWithFlags(&ws->set_stream[g], hipStreamNonBlocking);
        if (e == hipSuccess && !ws->join[g]) e = hipEventCreateWithFlags(&ws->join[g], hipEventDisableTiming);
        for (int k = 0; k < 2 && e == hipSuccess; k++)
            if (!ws->counted[g][k]) e = hipEventCreateWithFlags(&ws->counted[g][k], hipEventDisableTiming);
    }
    if (e != hipSuccess) *err = "cannot create the slot-set streams";
    return int(e);
}

// One set of slots traced by its own chain of queue kernels.
template <class Frames> struct SlotSet {
    Frames fr;
    WfArrays a;       // sliced to the set's slots, with its own counters
    uint32_t nslots;
};

// Iteration k of a slot set on stream s: trace, then shade, with the path queues of
// parity k; the shadow queue the trace reads is the one the previous shade wrote.
// mark(0) follows the trace launch and mark(1) the shade launch (stage timing).
template <class Frames, class Mark>
void enqueue_iteration(const SlotSet<Frames>& set, uint64_t k, hipStream_t s, typename Kernels<Frames>::TraceFn k_trace,
                       uint32_t g_trace, uint32_t g_shade, bool stats, Mark&& mark) {
    const WfArrays& a = set.a;
    uint32_t* q_in = (k & 1) ? a.q1 : a.q0;
    uint32_t* q_out = (k & 1) ? a.q0 : a.q1;
    uint32_t* n_in = a.counts + (k & 1);
    uint32_t* n_out = a.counts + ((k + 1) & 1);
    uint32_t* qs_in = (k & 1) ? a.qs0 : a.qs1;  // parity of k - 1
    uint32_t* qs_out = (k & 1) ? a.qs1 : a.qs0;
    uint32_t* s_in = shadow_count(a, k + 1);
    uint32_t* s_out = shadow_count(a, k);
    hipLaunchKernelGGL(k_trace, dim3(g_trace), dim3(kWfBlock), 0, s, set.fr, a, q_in, n_in, qs_in, s_in, n_out, s_out);
    mark(0);
    if (stats)
        hipLaunchKernelGGL((k_wf_shade<Frames, true>), dim3(g_shade), dim3(kWfBlock), 0, s, set.fr, a, q_in, n_in,
                           qs_in, s_in, q_out, n_out, qs_out, s_out);
    else
        hipLaunchKernelGGL((k_wf_shade<Frames, false>), dim3(g_shade), dim3(kWfBlock), 0, s, set.fr, a, q_in, n_in,
                           qs_in, s_in, q_out, n_out, qs_out, s_out);
    mark(1);
}

// The queue loop.  Each slot set runs init, then chunks of kCheckEvery iterations
// (extend, shade, shadow, resolve) on its own stream, so one set's kernels fill the
// CUs another set's kernel tails leave idle; every set's stream waits for `st` at
// the start and `st` waits for every set at the end.  The host reads each set's
// live-path count one chunk behind the chunk it enqueues, so the device never waits
// for the host.  Every path ends within max_steps bounces.
template <class Frames>
int run_wavefront(WavefrontState* ws, const std::vector<SlotSet<Frames>>& sets, uint64_t max_steps, hipStream_t st,
                  bool any_hit, bool stats, int variant, StageTimer* tm, std::string* err) {
    if (ws->cus == 0) {
        int dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        ws->cus = std::max(cus, 1);
    }
    const int nsets = int(sets.size());
    int rc = ensure_streams(ws, nsets, err);
    if (rc) return rc;
    using K = Kernels<Frames>;
    const typename K::TraceFn k_trace = K::trace(variant, any_hit, stats);
    const void* k_shade = stats ? (const void*)k_wf_shade<Frames, true> : (const void*)k_wf_shade<Frames, false>;
    const uint32_t g_trace = ws->grid_for((const void*)k_trace);
    const uint32_t g_shade = ws->grid_for(k_shade);
    const uint32_t g_init = ws->grid_for((const void*)k_wf_init<Frames>);
    hipError_t e = hipSuccess;
    auto stream_of = [&](int g) { return g == 0 ? st : ws->set_stream[g]; };
    // stage marks (one slot set only): 0 trace, 1 shade, 3 init
    auto mark = [&](int stage) {
        if (tm && e == hipSuccess) e = tm->mark(st, stage);
    };
    auto iteration = [&](const SlotSet<Frames>& set, uint64_t k, hipStream_t s, bool marks) {
        enqueue_iteration(set, k, s, k_trace, g_trace, g_shade, stats, [&](int stage) {
            if (marks) mark(stage);
        });
    };
    static_assert(kCheckEvery % 2 == 0, "a chunk ends on queue 0");
    const bool graph = ws->use_graph && !tm;
    if (graph) {
        // the kCheckEvery iterations of a chunk of each set captured once as a HIP
        // graph (one launch per chunk instead of 16)
        if (!ws->cap_stream) {
            e = hipStreamCreateWithFlags(&ws->cap_stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                *err = "capture stream";
                return int(e);
            }
        }
        for (int g = 0; g < nsets; g++) {
            e = hipStreamBeginCapture(ws->cap_stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                *err = "hipStreamBeginCapture";
                return int(e);
            }
            for (uint64_t k = 0; k < kCheckEvery; k++) iteration(sets[g], k, ws->cap_stream, false);
            hipGraph_t cg = nullptr;
            hipError_t launch_err = hipGetLastError();
            e = hipStreamEndCapture(ws->cap_stream, &cg);
            if (e == hipSuccess) e = launch_err;
            if (e == hipSuccess && ws->exec[g]) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult res;
                if (hipGraphExecUpdate(ws->exec[g], cg, &bad, &res) != hipSuccess || res != hipGraphExecUpdateSuccess) {
                    (void)hipGetLastError();
                    (void)hipGraphExecDestroy(ws->exec[g]);
                    ws->exec[g] = nullptr;
                }
            }
            if (e == hipSuccess && !ws->exec[g]) e = hipGraphInstantiate(&ws->exec[g], cg, nullptr, nullptr, 0);
            if (cg) (void)hipGraphDestroy(cg);
            if (e != hipSuccess) {
                *err = "capture of the wavefront chunk";
                return int(e);
            }
        }
    }
    // fork: the sets' streams start after everything already enqueued on st
    if (nsets > 1) {
        e = hipEventRecord(ws->fork, st);
        for (int g = 1; g < nsets && e == hipSuccess; g++) e = hipStreamWaitEvent(ws->set_stream[g], ws->fork, 0);
        if (e != hipSuccess) {
            *err = "fork of the slot-set streams";
            return int(e);
        }
    }
    mark(-1);
    for (int g = 0; g < nsets; g++) {
        const SlotSet<Frames>& set = sets[g];
        hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, stream_of(g), set.a.counts);
        hipLaunchKernelGGL(k_wf_init<Frames>, dim3(std::min(g_init, (set.nslots + kWfBlock - 1) / kWfBlock)),
                           dim3(kWfBlock), 0, stream_of(g), set.fr, set.a, set.nslots);
    }
    mark(3);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) {
        *err = "launch k_wf_init";
        rc = int(e);
    }
    const uint64_t max_chunks = max_steps / kCheckEvery + 2;
    bool alive[kMaxStreams];
    for (int g = 0; g < kMaxStreams; g++) alive[g] = g < nsets;
    for (uint64_t c = 0; rc == 0; c++) {
        // enqueue chunk c of every live set, then its live-path count read-back
        for (int g = 0; g < nsets && rc == 0; g++) {
            if (!alive[g]) continue;
            hipStream_t s = stream_of(g);
            if (graph) {
                e = hipGraphLaunch(ws->exec[g], s);
            } else {
                for (uint64_t k = 0; k < kCheckEvery; k++) iteration(sets[g], k, s, true);
                if (e == hipSuccess) e = hipGetLastError();
            }
            if (e == hipSuccess)
                e = hipMemcpyAsync(ws->host_live + 8 * g + 4 * (c & 1), sets[g].a.counts, 4 * sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, s);
            if (e == hipSuccess) e = hipEventRecord(ws->counted[g][c & 1], s);
            if (e != hipSuccess) {
                *err = "launch of a wavefront chunk";
                rc = int(e);
            }
        }
        if (rc || c == 0) continue;
        // sets whose count after chunk c - 1 was zero are done (their chunk c found empty queues)
        bool any = false;
        for (int g = 0; g < nsets && rc == 0; g++) {
            if (!alive[g]) continue;
            e = hipEventSynchronize(ws->counted[g][(c - 1) & 1]);
            if (e != hipSuccess) {
                *err = "reading the queue length";
                rc = int(e);
                break;
            }
            // after an even number of iterations: the path queue (counts[0]) and the
            // shadow queue of the last shade (odd parity), whose retiring entries
            // are pixels still to be written
            const uint32_t* hl = ws->host_live + 8 * g + 4 * ((c - 1) & 1);
            const uint32_t live = hl[0] + hl[3];
            if (live == 0u) {
                alive[g] = false;
            } else if (live < ws->finish_below && !tm) {
                // the paths left after chunk c (at most `live`) end in one launch
                const typename K::FinishFn k_finish = K::finish(any_hit, stats);
                const uint32_t blocks = std::min((live + kFinishBlock - 1) / kFinishBlock, 65535u);
                hipLaunchKernelGGL(k_finish, dim3(blocks), dim3(kFinishBlock), 0, stream_of(g), sets[g].fr, sets[g].a,
                                   sets[g].a.q0, sets[g].a.counts, sets[g].a.qs1, shadow_count(sets[g].a, 1));
                e = hipGetLastError();
                if (e != hipSuccess) {
                    *err = "launch k_wf_finish";
                    rc = int(e);
                    break;
                }
                alive[g] = false;
            }
            any = any || alive[g];
        }
        if (rc == 0 && !any) break;
        if (rc == 0 && c > max_chunks) {
            *err = "paths did not terminate";
            rc = int(hipErrorUnknown);
        }
    }
    // join: st continues after every set's work
    for (int g = 1; g < nsets; g++) {
        hipError_t j = hipEventRecord(ws->join[g], ws->set_stream[g]);
        if (j == hipSuccess) j = hipStreamWaitEvent(st, ws->join[g], 0);
        if (j != hipSuccess && rc == 0) {
            *err = "join of the slot-set streams";
            rc = int(j);
        }
    }
    return rc;
}

uint32_t frame_tiles(const FrameArgs& fa) { return ((fa.width + 7) / 8) * ((fa.height + 7) / 8); }

}  // namespace

int launch_wavefront(WavefrontState* ws, const FrameArgs& fa, hipStream_t st, bool any_hit, bool stats, int variant,
                     StageTimer* tm, std::string* err) {
    if (variant < 0 || variant >= kNumTravVariants) {
        *err = "unknown traversal variant";
        return kWavefrontUnsupported;
    }
    if (fa.samples == 0u || fa.bounces == 0u) return launch_megakernel(fa, st, any_hit, stats, tm);
    const uint32_t tiles = frame_tiles(fa);
    int rc = ensure_slots(ws, tiles * 64u, err);
    if (rc) return rc;
    // the frame's tiles dealt round-robin to the slot sets (stage timing: one set)
    const uint32_t nsets = tm ? 1u : std::min(uint32_t(ws->streams), tiles);
    std::vector<SlotSet<OneFrame>> sets;
    uint32_t base = 0;
    for (uint32_t g = 0; g < nsets; g++) {
        const uint32_t n = ((tiles - g + nsets - 1) / nsets) * 64u;
        sets.push_back(SlotSet<OneFrame>{OneFrame{fa, nsets, g}, slice(ws->a, base, ws->a.counts + 8 * g), n});
        base += n;
    }
    return run_wavefront(ws, sets, uint64_t(fa.samples) * fa.bounces, st, any_hit, stats, variant, tm, err);
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
        if (frames[i].samples && frames[i].bounces) total += frame_tiles(frames[i]) * 64u;
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
    // slot sets: runs of whole frames with about total / nsets slots each
    const uint32_t want = tm ? 1u : uint32_t(ws->streams);
    std::vector<uint32_t> cut{0u};  // first frame of each set, then count
    for (uint32_t i = 1; i < count && cut.size() < want; i++)
        if (uint64_t(base[i]) * want >= uint64_t(total) * cut.size() && base[i] > base[cut.back()] && base[i] < total)
            cut.push_back(i);  // every set owns at least one slot
    cut.push_back(count);
    const uint32_t nsets = uint32_t(cut.size()) - 1u;
    // frame arguments and per-set slot offsets (relative to the set) in device memory
    std::vector<uint32_t> set_base;
    std::vector<size_t> set_off;
    for (uint32_t g = 0; g < nsets; g++) {
        set_off.push_back(set_base.size());
        for (uint32_t i = cut[g]; i <= cut[g + 1]; i++) set_base.push_back(base[i] - base[cut[g]]);
    }
    const size_t fbytes = (sizeof(FrameArgs) * count + 255) & ~size_t(255);
    const size_t bbytes = (sizeof(uint32_t) * set_base.size() + 255) & ~size_t(255);
    const size_t sbytes = sizeof(uint32_t) * size_t(total);
    if (fbytes + bbytes + sbytes > ws->batch_cap) {
        if (ws->batch_mem) (void)hipFree(ws->batch_mem);
        ws->batch_mem = nullptr;
        ws->batch_cap = 0;
        hipError_t e = hipMalloc(&ws->batch_mem, fbytes + bbytes + sbytes);
        if (e != hipSuccess) {
            *err = "cannot allocate the batch table";
            return int(e);
        }
        ws->batch_cap = fbytes + bbytes + sbytes;
    }
    FrameArgs* d_frames = reinterpret_cast<FrameArgs*>(ws->batch_mem);
    uint32_t* d_base = reinterpret_cast<uint32_t*>(ws->batch_mem + fbytes);
    uint32_t* d_slot_frame = reinterpret_cast<uint32_t*>(ws->batch_mem + fbytes + bbytes);
    // frames with no samples or bounces own no slots: the image reset still applies
    hipError_t e = hipMemcpyAsync(d_frames, frames, fbytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_base, set_base.data(), bbytes, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        *err = "copy of the batch table";
        return int(e);
    }
    std::vector<SlotSet<ManyFrames>> sets;
    for (uint32_t g = 0; g < nsets; g++) {
        const uint32_t f0 = cut[g], nf = cut[g + 1] - cut[g];
        const uint32_t first = base[f0], n = base[cut[g + 1]] - first;
        if (n) {
            hipLaunchKernelGGL(k_slot_frames, dim3((n + 255) / 256), dim3(256), 0, st, d_base + set_off[g], nf, n,
                               d_slot_frame + first);
            e = hipGetLastError();
            if (e != hipSuccess) {
                *err = "launch k_slot_frames";
                return int(e);
            }
        }
        sets.push_back(SlotSet<ManyFrames>{
            ManyFrames{d_frames + f0, d_base + set_off[g], nf, frames[0].counters, d_slot_frame + first},
            slice(ws->a, first, ws->a.counts + 8 * g), n});
    }
    rc = run_wavefront(ws, sets, steps, st, any_hit, stats, variant, tm, err);
    // the host arrays the copies read must outlive them
    if (rc == 0) {
        e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = int(e);
    }
    return rc;
}


}  // namespace rt
