// render.hip — the plain megakernel (one lane per pixel; raygen loop with the
// closest-hit, any-hit and miss shaders inlined) plus the ray-query and
// primary-hit kernels used for traversal parity.
//
// Mapping: 128-lane workgroups, each 64-lane wave owns an 8x8 pixel tile; the
// workgroup order is remapped so each of the 8 XCDs (which receive workgroups
// round-robin) renders one horizontal band of the image and its L2 sees the part
// of the scene that band looks at.
//
// Persistent form (default): a grid of as many workgroups as can be resident, each
// lane taking its next pixel from one counter when its pixel is done (the pixel's
// samples stay in order in one lane, so the image is the same), with its traversals
// stepped so that a wave neither lasts as long as its slowest pixel nor waits each
// bounce for its slowest traversal.  Pixels are numbered tile by tile
// (8x8, tiles in image order), so the pixels in flight at any time form one band of
// the image that every XCD works on: the caches the XCDs share hold the part of the
// scene that band sees.  (One band per XCD, each XCD's workgroups taking their own
// band first, was 3 % slower on the 1080p frame.)
#include <algorithm>
#include <mutex>
#include <unordered_map>

#include "path.h"

namespace rt {

namespace {

constexpr int kMegaBlock = 128;

__device__ inline uint32_t xcd_remap(uint32_t bid, uint32_t nblocks) {
    const uint32_t nx = 8;
    uint32_t full = (nblocks / nx) * nx;
    if (bid >= full) return bid;
    uint32_t per = full / nx;
    return (bid % nx) * per + bid / nx;
}

// The accumulation and display image writes of a finished pixel (RayTracing.rgen:
// 132-167); `cycles` is the pixel's clock count for the HEAT mapping.
__device__ inline void mega_write(const FrameArgs& fa, uint32_t x, uint32_t y, D3 radiance, float last_t,
                                  uint64_t cycles) {
    float4* acc = reinterpret_cast<float4*>(fa.accum) + size_t(y) * fa.width + x;
    float4 prev = fa.reset ? make_float4(0.f, 0.f, 0.f, 0.f) : *acc;
    const float4 sum =
        make_float4(prev.x + float(radiance.x), prev.y + float(radiance.y), prev.z + float(radiance.z), 0.f);
    *acc = sum;
    if (fa.o.out) write_output(fa.o, fa.width, x, y, sum, last_t, cycles);
}

// One pixel of the megakernel: the raygen loop, then the writes (RayTracing.rgen:25-167).
template <bool kAnyHit, bool kStats>
__device__ inline void mega_pixel(const FrameArgs& fa, uint32_t x, uint32_t y, uint32_t* stack, PathCounters& pc) {
    const bool heat = fa.o.out && fa.o.mapping == RT_MAPPING_HEAT;
    const uint64_t t0 = heat ? uint64_t(clock64()) : 0u;  // clockARB(), rgen:27
    float last_t;
    D3 radiance = render_pixel<kAnyHit, kStats>(fa, x, y, stack, kMegaBlock, pc, last_t);
    mega_write(fa, x, y, radiance, last_t, heat ? uint64_t(clock64()) - t0 : 0u);
}

// Persistent form: lanes take pixels from one counter (tile by tile, tiles in image
// order) and run them as PathLanes, traversals stepped and the wave's shading done
// together, so no lane waits each bounce for the slowest traversal of its wave; a
// lane that finishes its pixel takes the next (one atomic per wave).
template <bool kAnyHit, bool kStats>
__global__ __launch_bounds__(kMegaBlock) void k_megakernel_persistent(FrameArgs fa, uint32_t* next_pixel) {
    __shared__ uint32_t s_stack[kStackDepth * kMegaBlock];
    uint32_t* const stack = s_stack + threadIdx.x;
    const uint32_t tiles_x = (fa.width + 7) / 8;
    const uint32_t items = tiles_x * ((fa.height + 7) / 8) * 64u;
    const bool heat = fa.o.out && fa.o.mapping == RT_MAPPING_HEAT;
    PathCounters pc;
    PathLane<kAnyHit, kStats> ln;
    uint64_t t0 = 0;  // clockARB() at the pixel's start, rgen:27
    bool exhausted = false;
    for (;;) {
        if (!exhausted) {
            const uint32_t it = wave_fetch(ln.state == kLaneIdle, next_pixel, items, exhausted);
            if (it != kEmpty) {
                const uint32_t tile = it >> 6, within = it & 63u;
                const uint32_t x = (tile % tiles_x) * 8 + (within & 7u);
                const uint32_t y = (tile / tiles_x) * 8 + (within >> 3);
                if (x < fa.width && y < fa.height) {
                    if (heat) t0 = uint64_t(clock64());
                    ln.begin_pixel(fa, x, y, stack, kMegaBlock, pc);
                }
            }
        }
        const uint64_t holding = __ballot(ln.state != kLaneIdle);
        if (!holding && exhausted) break;
        const uint32_t keep = (uint32_t(__popcll(holding)) + 1u) / 2u;
        while (keep && uint32_t(__popcll(__ballot(ln.tracing()))) >= keep) {
            if (ln.tracing()) {
                AlphaAnyHit ah{&fa.sv, &fa.tlas, &ln.rng};
                ln.step(fa, ah, stack, kMegaBlock, pc);
            }
        }
        if (ln.state != kLaneIdle && !ln.tracing()) {
            ln.advance(fa, stack, kMegaBlock, pc);
            if (ln.state == kLaneDone) {
                mega_write(fa, ln.x, ln.y, ln.p.rad, ln.p.last_t, heat ? uint64_t(clock64()) - t0 : 0u);
                ln.state = kLaneIdle;
            }
        }
    }
    if (kStats) flush_counters(fa.counters, pc);
}

template <bool kAnyHit, bool kStats>
__global__ __launch_bounds__(kMegaBlock) void k_megakernel(FrameArgs fa) {
    __shared__ uint32_t s_stack[kStackDepth * kMegaBlock];
    const uint32_t tiles_x = (fa.width + 15) / 16;
    const uint32_t tiles_y = (fa.height + 7) / 8;
    const uint32_t bid = xcd_remap(blockIdx.x, tiles_x * tiles_y);
    const uint32_t tx = bid % tiles_x, ty = bid / tiles_x;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = tx * 16 + wave * 8 + (lane & 7);
    const uint32_t y = ty * 8 + (lane >> 3);
    PathCounters pc;
    if (x < fa.width && y < fa.height) mega_pixel<kAnyHit, kStats>(fa, x, y, s_stack + threadIdx.x, pc);
    if (kStats) flush_counters(fa.counters, pc);
}

template <bool kAnyHit>
__global__ __launch_bounds__(kMegaBlock) void k_ray_queries(TlasView tl, SceneView sv, const rt_ray* rays, uint32_t n,
                                                             uint32_t flags, rt_hit* hits, uint32_t seed_b) {
    __shared__ uint32_t s_stack[kStackDepth * kMegaBlock];
    uint32_t i = blockIdx.x * kMegaBlock + threadIdx.x;
    if (i >= n) return;
    rt_ray r = rays[i];
    Rng4 rng{i, 0u, seed_b, 0u};
    AlphaAnyHit ah{&sv, &tl, &rng};
    HitInfo h;
    TraceStats st;
    const float3 ro = mk3(r.origin[0], r.origin[1], r.origin[2]), rd = mk3(r.direction[0], r.direction[1], r.direction[2]);
    bool found = trace_ray<kAnyHit, false>(tl, ro, rd, RayFromF32{ro, rd}, r.tmin, r.tmax, flags, ah,
                                          s_stack + threadIdx.x, kMegaBlock, h, st);
    rt_hit out;
    out._pad[0] = out._pad[1] = 0;
    if (found) {
        const InstanceRecord& ir = tl.inst[h.inst];
        out.t = h.t;
        out.u = h.u;
        out.v = h.v;
        out.instance_id = ir.instance_id;
        out.primitive_id = h.prim;
        out.custom_index = ir.custom_index;
    } else {
        out.t = -1.f;
        out.u = out.v = 0.f;
        out.instance_id = out.primitive_id = out.custom_index = kEmpty;
    }
    hits[i] = out;
}

template <bool kAnyHit>
__global__ __launch_bounds__(kMegaBlock) void k_primary(FrameArgs fa, rt_hit* hits) {
    __shared__ uint32_t s_stack[kStackDepth * kMegaBlock];
    uint32_t i = blockIdx.x * kMegaBlock + threadIdx.x;
    if (i >= fa.width * fa.height) return;
    uint32_t x = i % fa.width, y = i / fa.width;
    Rng4 rng_a{x, y, fa.frame_count, 0u};
    Rng4 rng_b{x, y, fa.seed_b, 0u};
    uint32_t lcg = tea16(tea16(x, y), fa.seed_b);
    D3 o, d;
    double tfar;
    camera_ray(fa, x, y, rng_a, lcg, o, d, tfar);
    AlphaAnyHit ah{&fa.sv, &fa.tlas, &rng_b};
    HitInfo h;
    TraceStats st;
    uint32_t flags = fa.fully_opaque ? RT_RAY_FLAG_OPAQUE : RT_RAY_FLAG_NONE;
    bool found = trace_ray<kAnyHit, false>(fa.tlas, to_f3(o), to_f3(d), RayFromF64{o, d}, float(kTMin), float(tfar),
                                          flags, ah, s_stack + threadIdx.x, kMegaBlock, h, st);
    rt_hit out;
    out._pad[0] = out._pad[1] = 0;
    if (found) {
        const InstanceRecord& ir = fa.tlas.inst[h.inst];
        out.t = h.t;
        out.u = h.u;
        out.v = h.v;
        out.instance_id = ir.instance_id;
        out.primitive_id = h.prim;
        out.custom_index = ir.custom_index;
    } else {
        out.t = -1.f;
        out.u = out.v = 0.f;
        out.instance_id = out.primitive_id = out.custom_index = kEmpty;
    }
    hits[i] = out;
}

}  // namespace

// Resident workgroups of kernel f across the device (occupancy x CUs), per kernel.
static uint32_t resident_grid(const void* f) {
    static std::mutex mu;
    static std::unordered_map<const void*, uint32_t> grids;
    std::lock_guard<std::mutex> lk(mu);
    int dev = 0;
    (void)hipGetDevice(&dev);
    const void* key = static_cast<const char*>(f) + dev;  // per kernel and device
    auto it = grids.find(key);
    if (it != grids.end()) return it->second;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, kMegaBlock, 0) != hipSuccess || per_cu < 1) per_cu = 2;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus < 1) cus = 256;
    const uint32_t g = uint32_t(cus) * uint32_t(per_cu);
    grids.emplace(key, g);
    return g;
}

int launch_megakernel(const FrameArgs& fa, hipStream_t st, bool any_hit, bool stats, StageTimer* tm,
                      uint32_t* next_pixel) {
    uint32_t tiles = ((fa.width + 15) / 16) * ((fa.height + 7) / 8);
    if (tiles == 0) return 0;
    if (tm) {
        hipError_t e = tm->mark(st, -1);
        if (e != hipSuccess) return int(e);
    }
    if (next_pixel) {
        hipError_t e = hipMemsetAsync(next_pixel, 0, sizeof(uint32_t), st);
        if (e != hipSuccess) return int(e);
        const void* f = any_hit ? (stats ? (const void*)k_megakernel_persistent<true, true>
                                         : (const void*)k_megakernel_persistent<true, false>)
                                : (stats ? (const void*)k_megakernel_persistent<false, true>
                                         : (const void*)k_megakernel_persistent<false, false>);
        // no more workgroups than there are pixel tiles to start on
        const uint32_t g = std::min(resident_grid(f), tiles);
        if (any_hit) {
            if (stats) hipLaunchKernelGGL((k_megakernel_persistent<true, true>), dim3(g), dim3(kMegaBlock), 0, st, fa, next_pixel);
            else hipLaunchKernelGGL((k_megakernel_persistent<true, false>), dim3(g), dim3(kMegaBlock), 0, st, fa, next_pixel);
        } else {
            if (stats) hipLaunchKernelGGL((k_megakernel_persistent<false, true>), dim3(g), dim3(kMegaBlock), 0, st, fa, next_pixel);
            else hipLaunchKernelGGL((k_megakernel_persistent<false, false>), dim3(g), dim3(kMegaBlock), 0, st, fa, next_pixel);
        }
    } else if (any_hit) {
        if (stats) hipLaunchKernelGGL((k_megakernel<true, true>), dim3(tiles), dim3(kMegaBlock), 0, st, fa);
        else hipLaunchKernelGGL((k_megakernel<true, false>), dim3(tiles), dim3(kMegaBlock), 0, st, fa);
    } else {
        if (stats) hipLaunchKernelGGL((k_megakernel<false, true>), dim3(tiles), dim3(kMegaBlock), 0, st, fa);
        else hipLaunchKernelGGL((k_megakernel<false, false>), dim3(tiles), dim3(kMegaBlock), 0, st, fa);
    }
    if (tm) {
        // the megakernel's time is all traversal + shading: reported as the extend stage
        hipError_t e = tm->mark(st, 0);
        if (e != hipSuccess) return int(e);
    }
    return int(hipGetLastError());
}

int launch_ray_queries(const TlasView& tl, const SceneView& sv, const rt_ray* rays, uint32_t n, uint32_t flags,
                       rt_hit* hits, uint32_t seed_b, hipStream_t st, bool any_hit) {
    if (n == 0) return 0;
    uint32_t nb = (n + kMegaBlock - 1) / kMegaBlock;
    if (any_hit) hipLaunchKernelGGL((k_ray_queries<true>), dim3(nb), dim3(kMegaBlock), 0, st, tl, sv, rays, n, flags, hits, seed_b);
    else hipLaunchKernelGGL((k_ray_queries<false>), dim3(nb), dim3(kMegaBlock), 0, st, tl, sv, rays, n, flags, hits, seed_b);
    return int(hipGetLastError());
}

int launch_primary(const FrameArgs& fa, rt_hit* hits, hipStream_t st, bool any_hit) {
    uint32_t n = fa.width * fa.height;
    if (n == 0) return 0;
    uint32_t nb = (n + kMegaBlock - 1) / kMegaBlock;
    if (any_hit) hipLaunchKernelGGL((k_primary<true>), dim3(nb), dim3(kMegaBlock), 0, st, fa, hits);
    else hipLaunchKernelGGL((k_primary<false>), dim3(nb), dim3(kMegaBlock), 0, st, fa, hits);
    return int(hipGetLastError());
}

// ------------------------------------------------------- no-allocation form
// Callers that may not allocate (serving loops, harnesses that forbid it) trace
// through launch_noalloc: the persistent megakernel, whose only scratch is its pixel
// counter, with that counter and the frame's texture descriptors in one file-scope
// device array instead of memory the library allocates.  Every call on one device
// shares its array: calls that overlap on one device race.
namespace {
constexpr size_t kNoAllocTexBytes = size_t(kNoAllocMaxTextures) * sizeof(rt_texture);
constexpr size_t kNoAllocCounter = (kNoAllocTexBytes + 255) & ~size_t(255);  // offset of the pixel counter
constexpr size_t kNoAllocBytes = kNoAllocCounter + 256;
constexpr int kNoAllocMaxDevices = 64;
}  // namespace

static __device__ __align__(256) char scratch[kNoAllocBytes];

// Device address of `scratch` on each device (the host-side name is not one).
static char* g_scratch[kNoAllocMaxDevices] = {};

int noalloc_init(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return int(hipErrorInvalidDevice);
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (g_scratch[device]) return 0;
    char* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));  // the current device's copy
    if (e != hipSuccess) return int(e);
    if (!p) return int(hipErrorInvalidSymbol);
    g_scratch[device] = p;
    return 0;
}

int launch_noalloc(int device, const FrameArgs& fa_in, const rt_texture* host_textures, hipStream_t st, bool any_hit,
                   std::string* err) {
    if (device < 0 || device >= kNoAllocMaxDevices || !g_scratch[device]) {
        *err = "the device's scratch array was not set up";
        return kNoAllocUnsupported;
    }
    if (fa_in.sv.num_textures > kNoAllocMaxTextures) {
        *err = "more textures than the scratch array holds";
        return kNoAllocUnsupported;
    }
    char* base = g_scratch[device];
    FrameArgs fa = fa_in;
    if (fa.sv.num_textures) {
        // the descriptors of this frame (the caller's host array), ahead of its kernel
        const hipError_t e = hipMemcpyAsync(base, host_textures, sizeof(rt_texture) * fa.sv.num_textures,
                                            hipMemcpyHostToDevice, st);
        if (e != hipSuccess) {
            *err = "copy of the texture descriptors";
            return int(e);
        }
        fa.sv.textures = reinterpret_cast<rt_texture*>(base);
    }
    const int rc = launch_megakernel(fa, st, any_hit, false, nullptr, reinterpret_cast<uint32_t*>(base + kNoAllocCounter));
    if (rc) *err = "megakernel launch";
    return rc;
}

}  // namespace rt
