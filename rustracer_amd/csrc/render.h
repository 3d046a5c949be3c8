// render.h — launch interface of the trace kernels (render.hip, wavefront.hip).
#pragma once
#include <string>
#include <vector>

#include "output.h"
#include "shade.h"
#include "traverse.h"

namespace rt {

// Per-launch arguments: the UBO fields the pipeline reads, the scene bindings and
// the acceleration structure.  Passed by value (kernel arguments).
struct FrameArgs {
    TlasView tlas;
    SceneView sv;
    float* accum;
    uint32_t width, height;
    uint32_t seed_b;       // stream-B seed (replaces clockARB(), RayTracing.rgen:33,40)
    uint32_t reset;        // number_of_samples == total_number_of_samples
    unsigned long long* counters;  // kNumCounters, used by the counting variants
    float mvi[16], pin[16];        // ModelViewInverse, ProjectionInverse (column-major)
    float aperture, focus_distance, ortho, exposure;
    uint32_t samples, bounces, antialiasing, mapping, frame_count, fully_opaque;
    uint32_t debug;        // ubo.debug == 1: the legacy materials of RayTracing.rchit:437-475
    OutputArgs o;                  // display image (o.out NULL: none)
};

enum Counter : int {
    kCntPrimary = 0, kCntExtension, kCntShadow, kCntNodes, kCntTris, kCntInstances, kCntShade, kCntAnyHit,
    kCntTex, kCntLights, kCntOverflow, kCntBadIndex, kCntShadowNodes, kCntShadowTris, kCntShadowInstances,
    kCntGradTaps,     // adjoint: gradient contributions (texel or light) before the per-wave sums
    kCntGradAtomics,  // adjoint: float atomics issued to the gradient buffers
    kNumCounters
};

// Per-stage event timing of a trace (enabled by RT_OPTION_STAGE_TIMING).
struct StageTimer {
    std::vector<hipEvent_t> events;
    std::vector<int> stage;  // stage of the interval ending at events[i] (-1: none)
    size_t used = 0;
    // records an event closing an interval of `st` (0 extend, 1 shade, 2 shadow, 3 resolve, -1 start)
    hipError_t mark(hipStream_t s, int st) {
        if (used == events.size()) {
            hipEvent_t e;
            hipError_t r = hipEventCreate(&e);
            if (r != hipSuccess) return r;
            events.push_back(e);
            stage.push_back(-1);
        }
        stage[used] = st;
        return hipEventRecord(events[used++], s);
    }
    ~StageTimer() {
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

// next_pixel: one device counter the persistent form takes pixels from, zeroed on
// `st` here (NULL: one workgroup per 16x8 pixel tile, one pixel per lane).
int launch_megakernel(const FrameArgs& fa, hipStream_t st, bool any_hit, bool stats, StageTimer* timer = nullptr,
                      uint32_t* next_pixel = nullptr);
int launch_ray_queries(const TlasView& tl, const SceneView& sv, const rt_ray* rays, uint32_t n, uint32_t flags,
                       rt_hit* hits, uint32_t seed_b, hipStream_t st, bool any_hit);
int launch_primary(const FrameArgs& fa, rt_hit* hits, hipStream_t st, bool any_hit);

// No-allocation form (rt_trace_rays_noalloc): the persistent megakernel with its pixel
// counter and the frame's texture descriptors in a file-scope device array (about
// 41 KB per device), no stats.  noalloc_init(device) must have run once with that
// device current (it records the array's address).  Returns kNoAllocUnsupported (err
// set) for a frame it cannot take.
constexpr uint32_t kNoAllocMaxTextures = 1024;
constexpr int kNoAllocUnsupported = -1000;
int noalloc_init(int device);
int launch_noalloc(int device, const FrameArgs& fa, const rt_texture* host_textures, hipStream_t st, bool any_hit,
                   std::string* err);

}  // namespace rt
