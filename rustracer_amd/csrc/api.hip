// api.hip — the C ABI (include/rt/rt.h): contexts, acceleration-structure handles,
// trace entry points.
//
// Mirrors the reference's operator surface: create_as / create_top_as
// (crates/libs/asset_loader/src/acceleration_structures.rs:79-249) for builds and
// record_raytracing_commands -> trace_rays (crates/examples/gltf_viewer/src/main.rs:244-270,
// crates/libs/vulkan/src/command.rs:374-392) for frames.  Every entry point returns
// a status code; nothing aborts.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "lbvh.h"
#include "skin.h"
#include "render.h"
#include "wavefront.h"
#include "adjoint.h"

using namespace rt;

namespace {

// Samples per pixel of the autotuner's probe frames (the frame's own, up to this).  A
// pixel's samples run one after another, so the tail of a frame, and with it the best
// slot-set count and finishing threshold, grows with the samples: probes at fewer
// samples than the frame pick other settings (at 8 spp, 4 slot sets for the 1080p /
// 64 spp frame, which runs 3 % faster with 2).
constexpr uint32_t kAutotuneSamples = 64;
constexpr double kAutotuneMargin = 0.015;  // relative gain a probe needs to replace the incumbent
// Adjoint: vertices at bounces below this sum their texel taps per wave in LDS
// before the atomics (camera-ray vertices of neighbouring pixels and samples share
// texels; deeper vertices scatter over the scene).
constexpr uint32_t kAdjointTableBounces = 1;
constexpr uint32_t kBlasLeafMax = 2;  // default triangles per leaf (RT_OPTION_BLAS_LEAF_SIZE)
constexpr uint32_t kTreeletPasses = 2;  // default RT_OPTION_BVH_TREELET_PASSES
constexpr uint32_t kTlasLeafMax = 1;

std::string g_create_error;

struct DevFree {
    void operator()(void* p) const { if (p) (void)hipFree(p); }
};
using DevPtr = std::unique_ptr<void, DevFree>;

// Device memory of one batched BLAS build, shared by the BLASes of the batch and
// kept alive by every TLAS that references one of them.
struct BlasPool {
    DevPtr nodes, tris, descs;
    size_t bytes = 0;
    uint32_t count = 0, num_nodes = 0, num_tris = 0;
    std::vector<uint32_t> prim_count;   // per BLAS, in build order
    std::vector<uint32_t> level_start;  // node levels of the pool's trees (refit)
};

struct BlasEntry {
    std::shared_ptr<BlasPool> pool;
    BlasDesc* d_desc;
    BlasDesc h_desc;
    uint32_t num_nodes_pool;
};

struct TlasEntry {
    DevPtr nodes, inst;
    uint32_t num_nodes = 0, num_inst = 0;
    uint32_t num_tlas = 0;  // instances in the TLAS proper (inst[0, num_tlas))
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    DevPtr flat_nodes, flat_tris;  // flattened BVH over the identity instances
    uint32_t flat_num_nodes = 0, flat_num_tris = 0;
    float flat_bounds[6] = {0, 0, 0, 0, 0, 0};
    std::vector<uint32_t> level_start, flat_level_start;  // node ranges per tree level (refit)
    std::vector<std::shared_ptr<BlasPool>> pools;
    bool any_non_opaque = false;
};

TlasView tlas_view(const TlasEntry& t, uint32_t* err) {
    TlasView v{};
    v.nodes = (const QNode*)t.nodes.get();
    v.inst = (const InstanceRecord*)t.inst.get();
    v.root = 0;
    v.num_inst = t.num_inst;
    v.error_flags = err;
    v.num_tlas = t.num_tlas;
    v.flat_nodes = t.flat_num_tris ? (const QNode*)t.flat_nodes.get() : nullptr;
    v.flat_tris = (const float4*)t.flat_tris.get();
    v.flat_root = 0;
    return v;
}

bool is_identity(const float (&m)[3][4]) {
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 4; c++)
            if (m[r][c] != (r == c ? 1.0f : 0.0f)) return false;
    return true;
}

}  // namespace

struct rt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    uint64_t next_handle = 1;
    std::unordered_map<uint64_t, BlasEntry> blas;
    std::unordered_map<uint64_t, uint64_t> blas_by_addr;
    std::unordered_map<uint64_t, TlasEntry> tlas;
    unsigned long long* d_counters = nullptr;
    uint32_t* d_err = nullptr;  // kErr* flags raised by kernels, checked at rt_sync
    float* d_srgb = nullptr;
    rt_texture* d_textures = nullptr;
    uint32_t tex_cap = 0;
    std::vector<rt_texture> tex_cache;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    WavefrontState* wavefront = nullptr;
    bool flatten = true;  // RT_OPTION_FLATTEN_IDENTITY_INSTANCES
    float** d_grad_ptrs = nullptr;  // adjoint: device copy of the gradient pointers
    rt_texture* d_batch_tex = nullptr;  // batch: texture descriptors of every item, end to end
    size_t batch_tex_cap = 0;
    uint32_t grad_cap = 0;
    int variant = 0;      // RT_OPTION_TRAVERSAL_VARIANT
    bool use_graph = false;  // RT_OPTION_GRAPH
    uint32_t refill_ext = 16, refill_sh = 1;  // RT_OPTION_EXTEND_REFILL / RT_OPTION_SHADOW_REFILL
    bool out_bgra = false;   // RT_OPTION_OUTPUT_BGRA
    int streams = 2;         // RT_OPTION_WAVEFRONT_STREAMS
    uint32_t leaf_max = kBlasLeafMax;  // RT_OPTION_BLAS_LEAF_SIZE
    uint32_t collapse = kCollapseSah;  // RT_OPTION_BVH_COLLAPSE
    uint32_t treelets = kTreeletPasses;  // RT_OPTION_BVH_TREELET_PASSES
    bool mega_persistent = true;       // RT_OPTION_MEGAKERNEL_PERSISTENT
    uint32_t* d_next_pixel = nullptr;  // pixel counter of the persistent megakernel and adjoint
    uint32_t finish = 65536; // RT_OPTION_WAVEFRONT_FINISH
    bool stage_timing = false;
    StageTimer timer;  // events of the last trace when stage_timing is set
};

namespace {

int fail(rt_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    else g_create_error = msg;
    return code;
}
int hip_fail(rt_ctx* ctx, hipError_t e, const char* where) {
    return fail(ctx, e == hipErrorOutOfMemory ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP,
                std::string(where) + ": " + hipGetErrorString(e));
}
#define RT_HIP(ctx, call, where)                          \
    do {                                                  \
        hipError_t e_ = (call);                           \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, where); \
    } while (0)

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev) {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DeviceGuard() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

template <class T> T* dev_alloc(size_t count, hipError_t* e) {
    void* p = nullptr;
    *e = hipMalloc(&p, count * sizeof(T) + 16);
    return (T*)p;
}

void invert34(const double m[12], double r[12]) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double inv = det != 0.0 ? 1.0 / det : 0.0;
    double r3[9] = {A * inv, -(b * i - c * h) * inv, (b * f - c * e) * inv, B * inv, (a * i - c * g) * inv,
                    -(a * f - c * d) * inv, C * inv, -(a * h - b * g) * inv, (a * e - b * d) * inv};
    for (int row = 0; row < 3; row++) {
        for (int col = 0; col < 3; col++) r[row * 4 + col] = r3[row * 3 + col];
        r[row * 4 + 3] = -(r3[row * 3 + 0] * m[3] + r3[row * 3 + 1] * m[7] + r3[row * 3 + 2] * m[11]);
    }
}

int validate_bindings(rt_ctx* ctx, const rt_bindings* b) {
    if (!b || !b->ubo) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "bindings and bindings->ubo are required");
    if (!b->vertices || !b->indices || !b->prim_infos || !b->materials)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "vertex, index, prim-info and material buffers are required");
    if (!b->num_vertices || !b->num_indices || !b->num_prim_infos || !b->num_materials)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "buffer element counts must be non-zero");
    if (b->num_dlights && !b->dlights) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "dlights is NULL");
    if (b->num_plights && !b->plights) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "plights is NULL");
    if (b->num_textures && !b->textures) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "textures is NULL");
    for (uint32_t i = 0; i < b->num_textures; i++) {
        const rt_texture& t = b->textures[i];
        if (!t.data || !t.width || !t.height || t.format > RT_FMT_RGBA32F || t.wrap_s > RT_WRAP_MIRRORED_REPEAT ||
            t.wrap_t > RT_WRAP_MIRRORED_REPEAT || t.mag_filter > RT_FILTER_LINEAR || t.min_filter > RT_FILTER_LINEAR)
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + " has an invalid descriptor");
    }
    if (b->ubo->has_sky) {
        const rt_texture* t = b->skybox;
        if (!t) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "ubo.has_sky needs the skybox binding");
        if (!t->data || !t->width || t->width != t->height || t->width > kMaxCubeFace || t->format > RT_FMT_RGBA32F ||
            t->mag_filter > RT_FILTER_LINEAR)
            return fail(ctx, RT_ERR_INVALID_ARGUMENT,
                        "the skybox needs square faces of 1.." + std::to_string(kMaxCubeFace) +
                            " texels and a valid format and filter");
    }
    return RT_OK;
}

// Device copy of the texture descriptors; re-uploaded only when they change.
int upload_textures(rt_ctx* ctx, const rt_bindings* b) {
    uint32_t n = b->num_textures;
    if (n == 0) return RT_OK;
    bool same = ctx->tex_cache.size() == n &&
                std::memcmp(ctx->tex_cache.data(), b->textures, sizeof(rt_texture) * n) == 0;
    if (same) return RT_OK;
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (n > ctx->tex_cap) {
        if (ctx->d_textures) (void)hipFree(ctx->d_textures);
        ctx->d_textures = nullptr;
        hipError_t e;
        ctx->d_textures = dev_alloc<rt_texture>(n, &e);
        if (e != hipSuccess) { ctx->tex_cap = 0; return hip_fail(ctx, e, "hipMalloc textures"); }
        ctx->tex_cap = n;
    }
    RT_HIP(ctx, hipMemcpy(ctx->d_textures, b->textures, sizeof(rt_texture) * n, hipMemcpyHostToDevice), "copy textures");
    ctx->tex_cache.assign(b->textures, b->textures + n);
    return RT_OK;
}

int make_frame(rt_ctx* ctx, const TlasEntry& tl, const rt_bindings* b, float* accum, uint32_t w, uint32_t h,
               FrameArgs* fa) {
    const rt_ubo& u = *b->ubo;
    std::memset(fa, 0, sizeof(*fa));
    fa->tlas = tlas_view(tl, ctx->d_err);
    SceneView& sv = fa->sv;
    sv.vertices = b->vertices;
    sv.indices = b->indices;
    sv.prim_infos = b->prim_infos;
    sv.materials = b->materials;
    sv.dlights = b->dlights;
    sv.plights = b->plights;
    sv.textures = ctx->d_textures;
    sv.srgb_lut = ctx->d_srgb;
    sv.num_dlights = b->num_dlights;
    sv.num_plights = b->num_plights;
    sv.num_textures = b->num_textures;
    sv.num_vertices = b->num_vertices;
    sv.num_indices = b->num_indices;
    sv.num_prim_infos = b->num_prim_infos;
    sv.num_materials = b->num_materials;
    sv.has_sky = u.has_sky ? 1u : 0u;
    if (sv.has_sky) sv.sky = *b->skybox;
    fa->accum = accum;
    fa->width = w;
    fa->height = h;
    // stream B / LCG seed: replaces clockARB() of RayTracing.rgen:33,40
    fa->seed_b = tea16(u.random_seed, u.frame_count);
    fa->reset = u.number_of_samples == u.total_number_of_samples;
    fa->counters = ctx->d_counters;
    std::memcpy(fa->mvi, u.model_view_inverse, sizeof(fa->mvi));
    std::memcpy(fa->pin, u.projection_inverse, sizeof(fa->pin));
    fa->aperture = u.aperture;
    fa->focus_distance = u.focus_distance;
    fa->ortho = u.orthographic_fov_dis;
    fa->exposure = u.exposure;
    fa->samples = u.number_of_samples;
    fa->bounces = u.number_of_bounces;
    fa->antialiasing = u.antialiasing;
    fa->mapping = u.mapping;
    fa->debug = u.debug == 1u ? 1u : 0u;
    fa->frame_count = u.frame_count;
    fa->fully_opaque = u.fully_opaque;
    fa->o.out = nullptr;
    fa->o.heat_scale = u.heatmap_scale;
    fa->o.tone_mode = u.tone_mapping_mode;
    fa->o.mapping = u.mapping;
    fa->o.total = u.total_number_of_samples;
    fa->o.bgra = 0u;
    return RT_OK;
}

// Triangle segments of a BLAS batch from the caller's inputs (validated).
int make_segments(rt_ctx* ctx, const rt_blas_input* inputs, uint32_t count, std::vector<TriSegment>& segs,
                  std::vector<uint32_t>& seg_start, uint64_t& total) {
    segs.resize(count);
    seg_start.resize(count + 1);
    total = 0;
    for (uint32_t i = 0; i < count; i++) {
        const rt_blas_input& in = inputs[i];
        if (in.primitive_count > 0) {
            if (!in.vertex_data || !in.index_data)
                return fail(ctx, RT_ERR_INVALID_ARGUMENT, "input " + std::to_string(i) + ": NULL vertex or index data");
            if (in.vertex_stride < 12 || (in.vertex_stride & 3u))
                return fail(ctx, RT_ERR_INVALID_ARGUMENT, "input " + std::to_string(i) + ": vertex_stride must be >= 12 and a multiple of 4");
            if (in.primitive_offset & 3u)
                return fail(ctx, RT_ERR_INVALID_ARGUMENT, "input " + std::to_string(i) + ": primitive_offset must be a multiple of 4");
        }
        seg_start[i] = uint32_t(total);
        segs[i] = TriSegment{(const uint8_t*)in.vertex_data,
                             in.index_data ? in.index_data + in.primitive_offset / 4 : nullptr,
                             in.vertex_stride, in.max_vertex, in.first_vertex, in.primitive_count, uint32_t(total), 0u};
        total += in.primitive_count;
        if (total >= (1ull << kLeafCountShift))
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "batch exceeds 2^27 triangles");
    }
    seg_start[count] = uint32_t(total);
    return RT_OK;
}

}  // namespace

// ----------------------------------------------------------------- context
extern "C" {

const char* rt_version(void) { return "rt-gfx950 0.1.0"; }

const char* rt_last_error(const rt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_error.c_str(); }

int rt_create(rt_ctx** out, int hip_device, void* hip_stream) {
    if (!out) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, RT_ERR_HIP, "no HIP device available");
    if (hip_device < 0 || hip_device >= n) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "hip_device out of range");
    auto* ctx = new rt_ctx();
    ctx->device = hip_device;
    DeviceGuard g(hip_device);
    // NULL: the device's default stream, which orders the calls with other work
    // the caller enqueues there (e.g. a framework's default stream)
    ctx->stream = (hipStream_t)hip_stream;
    ctx->d_counters = dev_alloc<unsigned long long>(kNumCounters, &e);
    if (e == hipSuccess) ctx->d_srgb = dev_alloc<float>(256, &e);
    if (e == hipSuccess) ctx->d_err = dev_alloc<uint32_t>(1, &e);
    if (e == hipSuccess) e = hipMemset(ctx->d_err, 0, sizeof(uint32_t));
    if (e == hipSuccess) ctx->d_next_pixel = dev_alloc<uint32_t>(1, &e);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    // the device address of the no-allocation scratch array (rt_trace_rays_noalloc)
    if (e == hipSuccess) e = hipError_t(noalloc_init(hip_device));
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) {
        float lut[256];
        for (int i = 0; i < 256; i++) {  // sRGB EOTF of the Vulkan spec, evaluated in double
            double c = i / 255.0;
            lut[i] = float(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
        }
        e = hipMemcpy(ctx->d_srgb, lut, sizeof(lut), hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        std::string msg = std::string("context set-up failed: ") + hipGetErrorString(e);
        rt_destroy(ctx);
        return fail(nullptr, RT_ERR_HIP, msg);
    }
    *out = ctx;
    return RT_OK;
}

void rt_destroy(rt_ctx* ctx) {
    if (!ctx) return;
    DeviceGuard g(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->tlas.clear();
    ctx->blas.clear();
    if (ctx->wavefront) wavefront_destroy(ctx->wavefront);
    if (ctx->d_counters) (void)hipFree(ctx->d_counters);
    if (ctx->d_srgb) (void)hipFree(ctx->d_srgb);
    if (ctx->d_grad_ptrs) (void)hipFree(ctx->d_grad_ptrs);
    if (ctx->d_batch_tex) (void)hipFree(ctx->d_batch_tex);
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->d_next_pixel) (void)hipFree(ctx->d_next_pixel);
    if (ctx->d_textures) (void)hipFree(ctx->d_textures);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    delete ctx;
}

// The status and message of a set of kErr* flags raised by kernels (RT_OK for none).
static int flags_error(rt_ctx* ctx, uint32_t flags) {
    if (flags & kErrStackOverflow)
        return fail(ctx, RT_ERR_STACK_OVERFLOW, "a traversal exceeded its stack restart budget; results are incomplete");
    if (flags & kErrBadSkin)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "a vertex's skin index or joint is out of range (left unskinned)");
    if (flags) return fail(ctx, RT_ERR_HIP, "a kernel raised an unknown error flag");
    return RT_OK;
}

int rt_sync(rt_ctx* ctx) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    DeviceGuard g(ctx->device);
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    uint32_t flags = 0;
    RT_HIP(ctx, hipMemcpy(&flags, ctx->d_err, sizeof(flags), hipMemcpyDeviceToHost), "read error flags");
    if (flags) {
        RT_HIP(ctx, hipMemset(ctx->d_err, 0, sizeof(uint32_t)), "reset error flags");
        return flags_error(ctx, flags);
    }
    return RT_OK;
}

// ------------------------------------------------------------------ BLAS
int rt_build_blas_batch(rt_ctx* ctx, const rt_blas_input* inputs, uint32_t count, rt_blas* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!inputs || !out || count == 0) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "inputs/out must be non-NULL, count > 0");
    DeviceGuard g(ctx->device);
    std::vector<TriSegment> segs;
    std::vector<uint32_t> seg_start;
    uint64_t total = 0;
    int src = make_segments(ctx, inputs, count, segs, seg_start, total);
    if (src) return src;
    const uint32_t n = uint32_t(total);
    hipStream_t st = ctx->stream;
    hipError_t e;
    DevPtr d_segs(dev_alloc<TriSegment>(count, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_seg_start(dev_alloc<uint32_t>(count + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_tri_in(dev_alloc<float4>(3 * size_t(n) + 3, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_lo(dev_alloc<float4>(n + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_hi(dev_alloc<float4>(n + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_bad(dev_alloc<uint32_t>(1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    RT_HIP(ctx, hipMemcpyAsync(d_segs.get(), segs.data(), sizeof(TriSegment) * count, hipMemcpyHostToDevice, st), "copy segments");
    RT_HIP(ctx, hipMemcpyAsync(d_seg_start.get(), seg_start.data(), sizeof(uint32_t) * (count + 1), hipMemcpyHostToDevice, st), "copy segments");
    RT_HIP(ctx, hipMemsetAsync(d_bad.get(), 0, sizeof(uint32_t), st), "memset");
    int rc = prepare_triangles(st, (const TriSegment*)d_segs.get(), count, n, (float4*)d_tri_in.get(),
                               (float4*)d_lo.get(), (float4*)d_hi.get(), (uint32_t*)d_bad.get());
    if (rc) return hip_fail(ctx, hipError_t(rc), "prepare_triangles");
    LbvhResult res;
    std::string err;
    rc = lbvh_build(st, (const float4*)d_lo.get(), (const float4*)d_hi.get(), n, (const uint32_t*)d_seg_start.get(),
                    seg_start, ctx->leaf_max, ctx->collapse, ctx->treelets, &res, &err);
    if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "BLAS build: " + err);
    DevPtr perm(res.perm), seg_bounds(res.seg_bounds);
    auto pool = std::make_shared<BlasPool>();
    pool->nodes.reset(res.nodes);
    pool->tris.reset(dev_alloc<float4>(3 * size_t(n) + 3, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    pool->descs.reset(dev_alloc<BlasDesc>(count, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    pool->bytes = sizeof(QNode) * size_t(res.num_nodes) + 48 * size_t(n) + sizeof(BlasDesc) * count;
    pool->count = count;
    pool->num_nodes = res.num_nodes;
    pool->num_tris = n;
    pool->level_start = res.level_start;
    for (uint32_t i = 0; i < count; i++) pool->prim_count.push_back(inputs[i].primitive_count);
    rc = reorder_triangles(st, (const float4*)d_tri_in.get(), (const uint32_t*)perm.get(), n, (float4*)pool->tris.get());
    if (rc) return hip_fail(ctx, hipError_t(rc), "reorder_triangles");
    std::vector<float> hb(6 * size_t(count));
    uint32_t bad = 0;
    RT_HIP(ctx, hipMemcpyAsync(hb.data(), seg_bounds.get(), sizeof(float) * hb.size(), hipMemcpyDeviceToHost, st), "copy bounds");
    RT_HIP(ctx, hipMemcpyAsync(&bad, d_bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, st), "copy flag");
    RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    if (bad) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "an index exceeds max_vertex");
    std::vector<BlasDesc> descs(count);
    for (uint32_t i = 0; i < count; i++) {
        BlasDesc& d = descs[i];
        d.nodes = (const QNode*)pool->nodes.get();
        d.tris = (const float4*)pool->tris.get();
        d.root = i;
        d.num_prims = inputs[i].primitive_count;
        d.first_prim = seg_start[i];
        d.flags = (inputs[i].geometry_flags & RT_GEOMETRY_OPAQUE) ? kBlasOpaque : 0u;
        std::memcpy(d.bounds, &hb[6 * size_t(i)], sizeof(d.bounds));
    }
    RT_HIP(ctx, hipMemcpy(pool->descs.get(), descs.data(), sizeof(BlasDesc) * count, hipMemcpyHostToDevice), "copy descs");
    for (uint32_t i = 0; i < count; i++) {
        uint64_t hnd = ctx->next_handle++;
        BlasEntry be{pool, (BlasDesc*)pool->descs.get() + i, descs[i], res.num_nodes};
        ctx->blas.emplace(hnd, be);
        ctx->blas_by_addr[uint64_t(uintptr_t(be.d_desc))] = hnd;
        out[i] = hnd;
    }
    return RT_OK;
}

int rt_build_blas(rt_ctx* ctx, const rt_blas_input* input, rt_blas* out) {
    return rt_build_blas_batch(ctx, input, 1, out);
}

int rt_skin_vertices(rt_ctx* ctx, const rt_vertex* d_in, rt_vertex* d_out, uint32_t count, const float* d_skins,
                     uint32_t num_skins) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (count && (!d_in || !d_out)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "vertex buffers must be non-NULL");
    if (num_skins && !d_skins) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "d_skins is NULL");
    if ((reinterpret_cast<uintptr_t>(d_in) | reinterpret_cast<uintptr_t>(d_out)) & 15u)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "vertex buffers must be 16-byte aligned");
    DeviceGuard g(ctx->device);
    int rc = launch_skin(ctx->stream, d_in, d_out, count, d_skins, num_skins, ctx->d_err);
    if (rc) return hip_fail(ctx, hipError_t(rc), "skinning launch");
    return RT_OK;
}

int rt_refit_blas_batch(rt_ctx* ctx, const rt_blas_input* inputs, uint32_t count, const rt_blas* blases) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!inputs || !blases || count == 0) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "inputs/blases must be non-NULL, count > 0");
    std::vector<BlasEntry*> ents(count);
    std::shared_ptr<BlasPool> pool;
    for (uint32_t i = 0; i < count; i++) {
        auto it = ctx->blas.find(blases[i]);
        if (it == ctx->blas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown BLAS handle");
        ents[i] = &it->second;
        if (i == 0) pool = it->second.pool;
        if (it->second.pool != pool || it->second.h_desc.root != i)
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "blases must be one rt_build_blas_batch call's handles, in its order");
    }
    if (pool->count != count)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "blases must be all of one rt_build_blas_batch call's handles");
    for (uint32_t i = 0; i < count; i++)
        if (inputs[i].primitive_count != pool->prim_count[i])
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "input " + std::to_string(i) + ": primitive_count differs from the build");
    DeviceGuard g(ctx->device);
    std::vector<TriSegment> segs;
    std::vector<uint32_t> seg_start;
    uint64_t total = 0;
    int rc = make_segments(ctx, inputs, count, segs, seg_start, total);
    if (rc) return rc;
    hipStream_t st = ctx->stream;
    hipError_t e;
    DevPtr d_segs(dev_alloc<TriSegment>(count, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_bad(dev_alloc<uint32_t>(1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    RT_HIP(ctx, hipMemcpyAsync(d_segs.get(), segs.data(), sizeof(TriSegment) * count, hipMemcpyHostToDevice, st), "copy segments");
    RT_HIP(ctx, hipMemsetAsync(d_bad.get(), 0, sizeof(uint32_t), st), "memset");
    std::string err;
    rc = lbvh_refit(st, (const TriSegment*)d_segs.get(), count, pool->num_tris, (float4*)pool->tris.get(),
                    (QNode*)pool->nodes.get(), pool->num_nodes, pool->level_start, (BlasDesc*)pool->descs.get(),
                    (uint32_t*)d_bad.get(), &err);
    if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "BLAS refit: " + err);
    uint32_t bad = 0;
    std::vector<BlasDesc> descs(count);
    RT_HIP(ctx, hipMemcpy(&bad, d_bad.get(), sizeof(uint32_t), hipMemcpyDeviceToHost), "copy flag");
    RT_HIP(ctx, hipMemcpy(descs.data(), pool->descs.get(), sizeof(BlasDesc) * count, hipMemcpyDeviceToHost), "copy descs");
    for (uint32_t i = 0; i < count; i++) std::memcpy(ents[i]->h_desc.bounds, descs[i].bounds, sizeof(descs[i].bounds));
    if (bad) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "an index exceeds max_vertex");
    return RT_OK;
}

int rt_blas_device_address(rt_ctx* ctx, rt_blas blas, uint64_t* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    auto it = ctx->blas.find(blas);
    if (it == ctx->blas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown BLAS handle");
    *out = uint64_t(uintptr_t(it->second.d_desc));
    return RT_OK;
}

void rt_destroy_blas(rt_ctx* ctx, rt_blas blas) {
    if (!ctx) return;
    auto it = ctx->blas.find(blas);
    if (it == ctx->blas.end()) return;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->blas_by_addr.erase(uint64_t(uintptr_t(it->second.d_desc)));
    ctx->blas.erase(it);
}

int rt_blas_info(rt_ctx* ctx, rt_blas blas, rt_bvh_info* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    auto it = ctx->blas.find(blas);
    if (it == ctx->blas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown BLAS handle");
    const BlasDesc& d = it->second.h_desc;
    out->nodes = uint64_t(uintptr_t(d.nodes));
    out->tris = uint64_t(uintptr_t(d.tris));
    out->root = d.root;
    out->num_nodes = it->second.num_nodes_pool;
    out->first_prim = d.first_prim;
    out->num_prims = d.num_prims;
    std::memcpy(out->bounds, d.bounds, sizeof(out->bounds));
    return RT_OK;
}

// ------------------------------------------------------------------ TLAS
int rt_build_tlas(rt_ctx* ctx, const rt_instance* d_instances, uint32_t count, rt_tlas* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out || (count && !d_instances)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "instances/out must be non-NULL");
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    std::vector<rt_instance> hi(count);
    if (count) {
        RT_HIP(ctx, hipMemcpyAsync(hi.data(), d_instances, sizeof(rt_instance) * count, hipMemcpyDeviceToHost, st), "copy instances");
        RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    TlasEntry te;
    std::vector<InstanceRecord> recs(count);
    std::vector<uint32_t> flat_prims(count, 0u);  // > 0: instance goes to the flattened BVH
    for (uint32_t i = 0; i < count; i++) {
        const rt_instance& in = hi[i];
        auto it = ctx->blas_by_addr.find(in.acceleration_structure_reference);
        if (it == ctx->blas_by_addr.end())
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": acceleration_structure_reference is not a live BLAS");
        const BlasEntry& be = ctx->blas.at(it->second);
        te.pools.push_back(be.pool);
        double m[12], w[12];
        for (int k = 0; k < 12; k++) {
            m[k] = in.transform[k / 4][k % 4];
            if (!std::isfinite(m[k])) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "instance " + std::to_string(i) + ": non-finite transform");
        }
        invert34(m, w);
        InstanceRecord& r = recs[i];
        std::memset(&r, 0, sizeof(r));
        for (int row = 0; row < 3; row++) {
            r.o2w[row] = make_float4(in.transform[row][0], in.transform[row][1], in.transform[row][2], in.transform[row][3]);
        }
        for (int k = 0; k < 12; k++) r.w2o_d[k] = w[k];
        r.blas = be.d_desc;
        r.nodes = be.h_desc.nodes;
        r.tris = be.h_desc.tris;
        r.root = be.h_desc.root;
        r.custom_index = in.custom_index_and_mask & 0xFFFFFFu;
        r.mask = in.custom_index_and_mask >> 24;
        uint32_t f = in.sbt_offset_and_flags >> 24;
        r.flags = ((be.h_desc.flags & kBlasOpaque) ? kInstOpaqueGeom : 0u) |
                  ((f & RT_INSTANCE_FORCE_OPAQUE) ? kInstForceOpaque : 0u) |
                  ((f & RT_INSTANCE_FORCE_NO_OPAQUE) ? kInstForceNoOpaque : 0u) |
                  (be.h_desc.num_prims == 0 ? kInstEmpty : 0u);
        r.instance_id = i;
        if (!(be.h_desc.flags & kBlasOpaque)) te.any_non_opaque = true;
        // identity instances that rays can hit are traversed in one world-space BVH
        if (ctx->flatten && is_identity(in.transform) && r.mask != 0u && be.h_desc.num_prims > 0)
            flat_prims[i] = be.h_desc.num_prims;
    }
    std::vector<uint32_t> tlas_ids, flat_ids;
    uint64_t flat_total = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (flat_prims[i]) {
            flat_ids.push_back(i);
            flat_total += flat_prims[i];
        } else {
            tlas_ids.push_back(i);
        }
    }
    if (flat_total >= (1ull << kLeafCountShift)) {
        // too large for one flattened build: keep every instance in the TLAS
        flat_ids.clear();
        flat_total = 0;
        tlas_ids.resize(count);
        for (uint32_t i = 0; i < count; i++) tlas_ids[i] = i;
    }
    const uint32_t nt = uint32_t(tlas_ids.size()), nf = uint32_t(flat_ids.size());
    std::vector<InstanceRecord> trecs(nt);
    for (uint32_t k = 0; k < nt; k++) trecs[k] = recs[tlas_ids[k]];
    hipError_t e;
    DevPtr d_rec(dev_alloc<InstanceRecord>(nt + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_lo(dev_alloc<float4>(nt + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    DevPtr d_hi(dev_alloc<float4>(nt + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    std::vector<uint32_t> seg_start = {0u, nt};
    DevPtr d_seg(dev_alloc<uint32_t>(2, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    if (nt) RT_HIP(ctx, hipMemcpyAsync(d_rec.get(), trecs.data(), sizeof(InstanceRecord) * nt, hipMemcpyHostToDevice, st), "copy records");
    RT_HIP(ctx, hipMemcpyAsync(d_seg.get(), seg_start.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, st), "copy segments");
    int rc = prepare_instances(st, (const InstanceRecord*)d_rec.get(), nt, (float4*)d_lo.get(), (float4*)d_hi.get());
    if (rc) return hip_fail(ctx, hipError_t(rc), "prepare_instances");
    LbvhResult res;
    std::string err;
    rc = lbvh_build(st, (const float4*)d_lo.get(), (const float4*)d_hi.get(), nt, (const uint32_t*)d_seg.get(),
                    seg_start, kTlasLeafMax, ctx->collapse, ctx->treelets, &res, &err);
    if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "TLAS build: " + err);
    DevPtr perm(res.perm), seg_bounds(res.seg_bounds);
    te.nodes.reset(res.nodes);
    te.num_nodes = res.num_nodes;
    te.level_start = res.level_start;
    te.inst.reset(dev_alloc<InstanceRecord>(count + 1, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    te.num_inst = count;
    te.num_tlas = nt;
    // records: TLAS leaf order first, then the flattened instances in caller order
    rc = reorder_instances(st, (const InstanceRecord*)d_rec.get(), (const uint32_t*)perm.get(), nt,
                           (InstanceRecord*)te.inst.get());
    if (rc) return hip_fail(ctx, hipError_t(rc), "reorder_instances");
    std::vector<InstanceRecord> frecs(nf);
    std::vector<FlatSource> fsrc(nf);
    uint32_t pos = 0;
    for (uint32_t k = 0; k < nf; k++) {
        const InstanceRecord& r = recs[flat_ids[k]];
        frecs[k] = r;
        const BlasEntry& be = ctx->blas.at(ctx->blas_by_addr.at(uint64_t(uintptr_t(r.blas))));
        bool opq = (r.flags & kInstOpaqueGeom) != 0;
        if (r.flags & kInstForceOpaque) opq = true;
        if (r.flags & kInstForceNoOpaque) opq = false;
        fsrc[k] = FlatSource{be.h_desc.tris, be.h_desc.first_prim, pos, (nt + k) | (opq ? kFlatOpaque : 0u), 0u};
        pos += be.h_desc.num_prims;
    }
    if (nf)
        RT_HIP(ctx, hipMemcpyAsync((InstanceRecord*)te.inst.get() + nt, frecs.data(), sizeof(InstanceRecord) * nf,
                                   hipMemcpyHostToDevice, st), "copy records");
    RT_HIP(ctx, hipMemcpyAsync(te.bounds, seg_bounds.get(), sizeof(te.bounds), hipMemcpyDeviceToHost, st), "copy bounds");
    if (nf) {
        const uint32_t n = uint32_t(flat_total);
        DevPtr d_src(dev_alloc<FlatSource>(nf, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        DevPtr d_tri_in(dev_alloc<float4>(3 * size_t(n) + 3, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        DevPtr d_flo(dev_alloc<float4>(n + 1, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        DevPtr d_fhi(dev_alloc<float4>(n + 1, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        DevPtr d_fseg(dev_alloc<uint32_t>(2, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        std::vector<uint32_t> fseg = {0u, n};
        RT_HIP(ctx, hipMemcpyAsync(d_src.get(), fsrc.data(), sizeof(FlatSource) * nf, hipMemcpyHostToDevice, st), "copy sources");
        RT_HIP(ctx, hipMemcpyAsync(d_fseg.get(), fseg.data(), sizeof(uint32_t) * 2, hipMemcpyHostToDevice, st), "copy segments");
        rc = gather_flat(st, (const FlatSource*)d_src.get(), nf, n, (float4*)d_tri_in.get(), (float4*)d_flo.get(),
                         (float4*)d_fhi.get());
        if (rc) return hip_fail(ctx, hipError_t(rc), "gather_flat");
        LbvhResult fres;
        rc = lbvh_build(st, (const float4*)d_flo.get(), (const float4*)d_fhi.get(), n, (const uint32_t*)d_fseg.get(),
                        fseg, ctx->leaf_max, ctx->collapse, ctx->treelets, &fres, &err);
        if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "flattened BVH build: " + err);
        DevPtr fperm(fres.perm), fseg_bounds(fres.seg_bounds);
        te.flat_nodes.reset(fres.nodes);
        te.flat_num_nodes = fres.num_nodes;
        te.flat_level_start = fres.level_start;
        te.flat_tris.reset(dev_alloc<float4>(3 * size_t(n) + 3, &e));
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
        te.flat_num_tris = n;
        rc = reorder_triangles(st, (const float4*)d_tri_in.get(), (const uint32_t*)fperm.get(), n,
                               (float4*)te.flat_tris.get());
        if (rc) return hip_fail(ctx, hipError_t(rc), "reorder_triangles");
        RT_HIP(ctx, hipMemcpyAsync(te.flat_bounds, fseg_bounds.get(), sizeof(te.flat_bounds), hipMemcpyDeviceToHost, st), "copy bounds");
        RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    }
    RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    uint64_t hnd = ctx->next_handle++;
    ctx->tlas.emplace(hnd, std::move(te));
    *out = hnd;
    return RT_OK;
}

int rt_refit_tlas(rt_ctx* ctx, rt_tlas tlas) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    TlasEntry& te = it->second;
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    std::string err;
    if (te.flat_num_tris) {
        int rc = lbvh_refit_flat(st, (const InstanceRecord*)te.inst.get(), te.flat_num_tris, (float4*)te.flat_tris.get(),
                                 (QNode*)te.flat_nodes.get(), te.flat_num_nodes, te.flat_level_start, te.flat_bounds,
                                 &err);
        if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "flattened BVH refit: " + err);
    }
    if (te.num_tlas) {
        int rc = lbvh_refit_instances(st, (const InstanceRecord*)te.inst.get(), te.num_tlas, (QNode*)te.nodes.get(),
                                      te.num_nodes, te.level_start, te.bounds, &err);
        if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "TLAS refit: " + err);
    }
    return RT_OK;
}

void rt_destroy_tlas(rt_ctx* ctx, rt_tlas tlas) {
    if (!ctx) return;
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return;
    DeviceGuard g(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->tlas.erase(it);
}

int rt_tlas_info(rt_ctx* ctx, rt_tlas tlas, rt_bvh_info* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    const TlasEntry& t = it->second;
    out->nodes = uint64_t(uintptr_t(t.nodes.get()));
    out->tris = uint64_t(uintptr_t(t.inst.get()));
    out->root = 0;
    out->num_nodes = t.num_nodes;
    out->first_prim = 0;
    out->num_prims = t.num_tlas;
    std::memcpy(out->bounds, t.bounds, sizeof(out->bounds));
    return RT_OK;
}

int rt_tlas_flat_info(rt_ctx* ctx, rt_tlas tlas, rt_bvh_info* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    const TlasEntry& t = it->second;
    out->nodes = uint64_t(uintptr_t(t.flat_nodes.get()));
    out->tris = uint64_t(uintptr_t(t.flat_tris.get()));
    out->root = 0;
    out->num_nodes = t.flat_num_nodes;
    out->first_prim = 0;
    out->num_prims = t.flat_num_tris;
    std::memcpy(out->bounds, t.flat_bounds, sizeof(out->bounds));
    return RT_OK;
}

// ----------------------------------------------------------------- trace
// Counters and device time of the trace just enqueued (ev0 recorded before it).
static int read_stats(rt_ctx* ctx, rt_trace_stats* stats) {
    hipStream_t st = ctx->stream;
    RT_HIP(ctx, hipEventRecord(ctx->ev1, st), "hipEventRecord");
    unsigned long long c[kNumCounters];
    RT_HIP(ctx, hipMemcpyAsync(c, ctx->d_counters, sizeof(c), hipMemcpyDeviceToHost, st), "copy counters");
    RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    float ms = 0.f;
    RT_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1), "hipEventElapsedTime");
    std::memset(stats, 0, sizeof(*stats));
    stats->primary_rays = c[kCntPrimary];
    stats->extension_rays = c[kCntExtension];
    stats->shadow_rays = c[kCntShadow];
    stats->rays = c[kCntPrimary] + c[kCntExtension] + c[kCntShadow];
    stats->node_visits = c[kCntNodes];
    stats->tri_tests = c[kCntTris];
    stats->instance_visits = c[kCntInstances];
    stats->shade_events = c[kCntShade];
    stats->anyhit_calls = c[kCntAnyHit];
    stats->tex_fetches = c[kCntTex];
    stats->light_candidates = c[kCntLights];
    stats->stack_overflows = c[kCntOverflow];
    stats->kernel_ms = ms;
    stats->shadow_node_visits = c[kCntShadowNodes];
    stats->shadow_tri_tests = c[kCntShadowTris];
    stats->shadow_instance_visits = c[kCntShadowInstances];
    stats->grad_taps = c[kCntGradTaps];
    stats->grad_atomics = c[kCntGradAtomics];
    if (c[kCntBadIndex])
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "a hit referenced an out-of-range prim info, material, index or vertex");
    return RT_OK;
}

int rt_trace_rays(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, float* accum, uint8_t* out_rgba8, uint32_t width,
                  uint32_t height, uint32_t mode, rt_trace_stats* stats) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    int rc = validate_bindings(ctx, b);
    if (rc) return rc;
    if (!accum) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "accum is NULL");
    if (width == 0 || height == 0 || uint64_t(width) * height > (1ull << 31))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "width/height out of range");
    if (mode != RT_MODE_MEGAKERNEL && mode != RT_MODE_WAVEFRONT) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "unknown mode");
    if (out_rgba8 && (reinterpret_cast<uintptr_t>(out_rgba8) & 3u))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out_rgba8 must be 4-byte aligned");
    // the HEAT mapping shows each pixel's shader clock: only the megakernel form runs a
    // pixel's whole raygen in one thread, so such frames take it (same accumulation)
    if (out_rgba8 && b->ubo->mapping == RT_MAPPING_HEAT) mode = RT_MODE_MEGAKERNEL;
    if (out_rgba8 && b->ubo->total_number_of_samples == 0)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "the display image needs total_number_of_samples > 0");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    DeviceGuard g(ctx->device);
    rc = upload_textures(ctx, b);
    if (rc) return rc;
    FrameArgs fa;
    make_frame(ctx, it->second, b, accum, width, height, &fa);
    fa.o.out = reinterpret_cast<uint32_t*>(out_rgba8);
    fa.o.bgra = ctx->out_bgra ? 1u : 0u;
    const bool any_hit = !fa.fully_opaque;
    // the counting kernels also carry the debug == 1 materials
    const bool counting = stats != nullptr || fa.debug != 0u;
    hipStream_t st = ctx->stream;
    if (stats) {
        RT_HIP(ctx, hipMemsetAsync(ctx->d_counters, 0, sizeof(unsigned long long) * kNumCounters, st), "memset");
        RT_HIP(ctx, hipEventRecord(ctx->ev0, st), "hipEventRecord");
    }
    StageTimer* tm = ctx->stage_timing ? &ctx->timer : nullptr;
    ctx->timer.used = 0;
    if (mode == RT_MODE_MEGAKERNEL) {
        rc = launch_megakernel(fa, st, any_hit, counting, tm, ctx->mega_persistent ? ctx->d_next_pixel : nullptr);
    } else {
        if (!ctx->wavefront) {
            ctx->wavefront = wavefront_create();
            if (!ctx->wavefront) return fail(ctx, RT_ERR_OUT_OF_MEMORY, "wavefront state");
        }
        std::string err;
        wavefront_set_graph(ctx->wavefront, ctx->use_graph);
        wavefront_set_refill(ctx->wavefront, ctx->refill_ext, ctx->refill_sh);
        wavefront_set_streams(ctx->wavefront, ctx->streams);
        wavefront_set_finish(ctx->wavefront, ctx->finish);
        rc = launch_wavefront(ctx->wavefront, fa, st, any_hit, counting, ctx->variant, tm, &err);
        if (rc == kWavefrontUnsupported) return fail(ctx, RT_ERR_UNSUPPORTED, "wavefront: " + err);
        if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "wavefront: " + err);
    }
    if (rc) return hip_fail(ctx, hipError_t(rc), "trace launch");
    if (stats) return read_stats(ctx, stats);
    return RT_OK;
}

int rt_trace_rays_noalloc(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, float* accum, uint8_t* out_rgba8,
                          uint32_t width, uint32_t height) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (b && b->num_textures > kNoAllocMaxTextures)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT,
                    "rt_trace_rays_noalloc takes at most " + std::to_string(kNoAllocMaxTextures) + " textures");
    int rc = validate_bindings(ctx, b);
    if (rc) return rc;
    if (!accum) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "accum is NULL");
    if (width == 0 || height == 0 || uint64_t(width) * height > (1ull << 31))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "width/height out of range");
    if (out_rgba8 && (reinterpret_cast<uintptr_t>(out_rgba8) & 3u))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out_rgba8 must be 4-byte aligned");
    if (out_rgba8 && b->ubo->total_number_of_samples == 0)
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "the display image needs total_number_of_samples > 0");
    if (b->ubo->debug == 1u)
        return fail(ctx, RT_ERR_UNSUPPORTED, "the debug == 1 materials need rt_trace_rays (counting kernels)");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    DeviceGuard g(ctx->device);
    FrameArgs fa;
    make_frame(ctx, it->second, b, accum, width, height, &fa);
    fa.sv.textures = nullptr;  // the descriptors go to the scratch array
    fa.o.out = reinterpret_cast<uint32_t*>(out_rgba8);
    fa.o.bgra = ctx->out_bgra ? 1u : 0u;
    std::string err;
    rc = launch_noalloc(ctx->device, fa, b->textures, ctx->stream, !fa.fully_opaque, &err);
    if (rc == kNoAllocUnsupported) return fail(ctx, RT_ERR_UNSUPPORTED, "noalloc: " + err);
    if (rc) return fail(ctx, RT_ERR_HIP, "noalloc: " + err + ": " + hipGetErrorString(hipError_t(rc)));
    return RT_OK;
}

int rt_trace_rays_batch(rt_ctx* ctx, const rt_batch_item* items, uint32_t count, uint32_t mode,
                        rt_trace_stats* stats) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (count && !items) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "items is NULL");
    if (mode != RT_MODE_MEGAKERNEL && mode != RT_MODE_WAVEFRONT) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "unknown mode");
    size_t ntex = 0;
    for (uint32_t i = 0; i < count; i++) {
        const rt_batch_item& it = items[i];
        std::string where = "item " + std::to_string(i) + ": ";
        int rc = validate_bindings(ctx, it.bindings);
        if (rc) return fail(ctx, rc, where + ctx->err);
        if (!it.accum) return fail(ctx, RT_ERR_INVALID_ARGUMENT, where + "accum is NULL");
        if (it.width == 0 || it.height == 0 || uint64_t(it.width) * it.height > (1ull << 31))
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, where + "width/height out of range");
        if (ctx->tlas.find(it.tlas) == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, where + "unknown TLAS handle");
        ntex += it.bindings->num_textures;
    }
    if (count == 0) return RT_OK;
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    // texture descriptors of every item, end to end
    RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    if (ntex > ctx->batch_tex_cap) {
        if (ctx->d_batch_tex) (void)hipFree(ctx->d_batch_tex);
        ctx->d_batch_tex = nullptr;
        ctx->batch_tex_cap = 0;
        hipError_t e;
        ctx->d_batch_tex = dev_alloc<rt_texture>(ntex, &e);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc textures");
        ctx->batch_tex_cap = ntex;
    }
    std::vector<FrameArgs> frames(count);
    size_t toff = 0;
    bool any_hit = false;
    bool counting = stats != nullptr;  // the counting kernels also carry the debug == 1 materials
    for (uint32_t i = 0; i < count; i++) {
        const rt_batch_item& it = items[i];
        make_frame(ctx, ctx->tlas.at(it.tlas), it.bindings, it.accum, it.width, it.height, &frames[i]);
        uint32_t n = it.bindings->num_textures;
        if (n)
            RT_HIP(ctx, hipMemcpy(ctx->d_batch_tex + toff, it.bindings->textures, sizeof(rt_texture) * n,
                                  hipMemcpyHostToDevice), "copy textures");
        frames[i].sv.textures = ctx->d_batch_tex + toff;
        toff += n;
        any_hit = any_hit || !frames[i].fully_opaque;
        counting = counting || frames[i].debug != 0u;
    }
    if (stats) {
        RT_HIP(ctx, hipMemsetAsync(ctx->d_counters, 0, sizeof(unsigned long long) * kNumCounters, st), "memset");
        RT_HIP(ctx, hipEventRecord(ctx->ev0, st), "hipEventRecord");
    }
    StageTimer* tm = ctx->stage_timing ? &ctx->timer : nullptr;
    ctx->timer.used = 0;
    int rc = 0;
    if (mode == RT_MODE_MEGAKERNEL) {
        for (uint32_t i = 0; i < count && rc == 0; i++)
            rc = launch_megakernel(frames[i], st, !frames[i].fully_opaque, counting, tm,
                                   ctx->mega_persistent ? ctx->d_next_pixel : nullptr);
        if (rc) return hip_fail(ctx, hipError_t(rc), "trace launch");
    } else {
        if (!ctx->wavefront) {
            ctx->wavefront = wavefront_create();
            if (!ctx->wavefront) return fail(ctx, RT_ERR_OUT_OF_MEMORY, "wavefront state");
        }
        // frames without samples or bounces take the megakernel (it applies the reset)
        for (uint32_t i = 0; i < count && rc == 0; i++)
            if (frames[i].samples == 0u || frames[i].bounces == 0u)
                rc = launch_megakernel(frames[i], st, !frames[i].fully_opaque, counting);
        if (rc) return hip_fail(ctx, hipError_t(rc), "trace launch");
        std::string err;
        wavefront_set_graph(ctx->wavefront, ctx->use_graph);
        wavefront_set_refill(ctx->wavefront, ctx->refill_ext, ctx->refill_sh);
        wavefront_set_streams(ctx->wavefront, ctx->streams);
        wavefront_set_finish(ctx->wavefront, ctx->finish);
        rc = launch_wavefront_batch(ctx->wavefront, frames.data(), count, st, any_hit, counting, ctx->variant, tm, &err);
        if (rc == kWavefrontUnsupported) return fail(ctx, RT_ERR_UNSUPPORTED, "wavefront: " + err);
        if (rc) return fail(ctx, rc == int(hipErrorOutOfMemory) ? RT_ERR_OUT_OF_MEMORY : RT_ERR_HIP, "wavefront: " + err);
    }
    if (stats) return read_stats(ctx, stats);
    RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
    return RT_OK;
}

int rt_trace_rays_adjoint(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, const float* d_grad_accum,
                          uint32_t width, uint32_t height, float* const* grad_textures, float* d_grad_point_intensity,
                          float* d_grad_directional_intensity, rt_trace_stats* stats) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    int rc = validate_bindings(ctx, b);
    if (rc) return rc;
    if (!d_grad_accum) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "d_grad_accum is NULL");
    if (width == 0 || height == 0 || uint64_t(width) * height > (1ull << 31))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "width/height out of range");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    if (b->ubo->number_of_bounces > kMaxAdjointBounces)
        return fail(ctx, RT_ERR_UNSUPPORTED, "the adjoint replays at most 64 bounces");
    if (b->ubo->debug == 1u)
        return fail(ctx, RT_ERR_UNSUPPORTED, "the adjoint differentiates the PBR materials, not debug == 1");
    const uint32_t nt = b->num_textures;
    if (nt && !grad_textures) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "grad_textures is NULL");
    for (uint32_t i = 0; i < nt; i++)
        if (grad_textures[i] && b->textures[i].format != RT_FMT_RGBA32F)
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "texture " + std::to_string(i) + " has a gradient buffer but is not RGBA32F");
    DeviceGuard g(ctx->device);
    hipStream_t st = ctx->stream;
    rc = upload_textures(ctx, b);
    if (rc) return rc;
    if (nt > ctx->grad_cap) {
        if (ctx->d_grad_ptrs) (void)hipFree(ctx->d_grad_ptrs);
        ctx->d_grad_ptrs = nullptr;
        ctx->grad_cap = 0;
        RT_HIP(ctx, hipMalloc(&ctx->d_grad_ptrs, sizeof(float*) * nt), "hipMalloc");
        ctx->grad_cap = nt;
    }
    if (nt) {
        // the pointer table must stay valid until the kernel has read it
        RT_HIP(ctx, hipStreamSynchronize(st), "hipStreamSynchronize");
        RT_HIP(ctx, hipMemcpy(ctx->d_grad_ptrs, grad_textures, sizeof(float*) * nt, hipMemcpyHostToDevice),
               "copy gradient pointers");
    }
    float** d_ptrs = ctx->d_grad_ptrs;
    FrameArgs fa;
    make_frame(ctx, it->second, b, nullptr, width, height, &fa);
    if (stats) {
        RT_HIP(ctx, hipMemsetAsync(ctx->d_counters, 0, sizeof(unsigned long long) * kNumCounters, st), "memset");
        RT_HIP(ctx, hipEventRecord(ctx->ev0, st), "hipEventRecord");
    }
    AdjointOut out{d_ptrs, nt, b->num_plights ? d_grad_point_intensity : nullptr,
                   b->num_dlights ? d_grad_directional_intensity : nullptr};
    rc = launch_adjoint(fa, d_grad_accum, out, kAdjointTableBounces, st, !fa.fully_opaque, stats != nullptr,
                        ctx->d_next_pixel);
    if (rc) return hip_fail(ctx, hipError_t(rc), "adjoint launch");
    if (stats) return read_stats(ctx, stats);
    return RT_OK;
}

int rt_autotune(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, uint32_t width, uint32_t height, rt_tuning* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    int rc = validate_bindings(ctx, b);
    if (rc) return rc;
    if (width == 0 || height == 0 || uint64_t(width) * height > (1ull << 31))
        return fail(ctx, RT_ERR_INVALID_ARGUMENT, "width/height out of range");
    DeviceGuard g(ctx->device);
    rt_ubo u = *b->ubo;
    u.number_of_samples = std::min(u.number_of_samples, kAutotuneSamples);
    u.total_number_of_samples = u.number_of_samples;
    rt_bindings bb = *b;
    bb.ubo = &u;
    hipError_t e;
    DevPtr scratch(dev_alloc<float4>(size_t(width) * height, &e));
    if (e != hipSuccess) return hip_fail(ctx, e, "hipMalloc");
    // start from a clear error flag (a set flag would be blamed on a setting)
    {
        uint32_t flags = 0;
        RT_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
        RT_HIP(ctx, hipMemcpy(&flags, ctx->d_err, sizeof(flags), hipMemcpyDeviceToHost), "read error flags");
        if (flags) {
            RT_HIP(ctx, hipMemset(ctx->d_err, 0, sizeof(uint32_t)), "reset error flags");
            return flags_error(ctx, flags);  // raised by earlier work: reported here, as rt_sync would
        }
    }
    const bool timing = ctx->stage_timing;
    ctx->stage_timing = false;
    const int keep_variant = ctx->variant;
    const bool keep_graph = ctx->use_graph;
    const uint32_t keep_re = ctx->refill_ext, keep_rs = ctx->refill_sh;
    const int keep_streams = ctx->streams;
    const uint32_t keep_finish = ctx->finish;
    // one probe: the mean frame time of the current settings over three timed frames
    // (after a warm-up frame), six when the three differ by more than 3 %; < 0 when a
    // traversal gave up.  The mean, not the fastest frame: with three or four slot
    // sets some settings alternate between runs of fast and of slow frames (179-184
    // against 188-192 ms on the 1080p benchmark frame), and the expected frame time is
    // what a setting should be chosen by.
    auto probe = [&](double& ms_out) -> int {
        int r = rt_trace_rays(ctx, tlas, &bb, (float*)scratch.get(), nullptr, width, height, RT_MODE_WAVEFRONT, nullptr);
        if (r) return r;
        hipError_t er = hipSuccess;
        float rep_ms[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        int reps = 3;
        for (int rep = 0; rep < reps && er == hipSuccess; rep++) {
            er = hipEventRecord(ctx->ev0, ctx->stream);
            if (er == hipSuccess)
                r = rt_trace_rays(ctx, tlas, &bb, (float*)scratch.get(), nullptr, width, height, RT_MODE_WAVEFRONT,
                                  nullptr);
            if (r) return r;
            if (er == hipSuccess) er = hipEventRecord(ctx->ev1, ctx->stream);
            if (er == hipSuccess) er = hipEventSynchronize(ctx->ev1);
            float ms = 0.f;
            if (er == hipSuccess) er = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
            rep_ms[rep] = ms;
            if (rep == 2) {
                const float lo = std::min(rep_ms[0], std::min(rep_ms[1], rep_ms[2]));
                const float hi = std::max(rep_ms[0], std::max(rep_ms[1], rep_ms[2]));
                if (hi > 1.03f * lo) reps = 6;
            }
        }
        double sum_ms = 0.0;
        for (int k = 0; k < reps; k++) sum_ms += rep_ms[k];
        uint32_t flags = 0;
        if (er == hipSuccess) er = hipMemcpy(&flags, ctx->d_err, sizeof(flags), hipMemcpyDeviceToHost);
        if (er == hipSuccess && flags) er = hipMemset(ctx->d_err, 0, sizeof(uint32_t));
        if (er != hipSuccess) return hip_fail(ctx, er, "autotune probe");
        ms_out = flags ? -1.0 : sum_ms / reps;
        return RT_OK;
    };
    // a setting replaces the incumbent only when it is faster by more than the
    // frame-to-frame noise, so repeated tunings of one scene choose alike
    double best = -1.0;
    auto consider = [&](double ms) { return ms >= 0.0 && (best < 0.0 || ms < best * (1.0 - kAutotuneMargin)); };
    ctx->use_graph = false;
    // 1. refill threshold of the traversal kernel (the default variant; the other
    //    settings are tuned under the refill chosen here), against the frame of the
    //    settings the context has
    rc = probe(best);
    if (rc == RT_OK) {
        const uint32_t refills[] = {1u, 16u, 32u, 48u, 64u};
        uint32_t chosen = ctx->refill_ext;
        for (uint32_t r : refills) {
            if (r == chosen) continue;
            ctx->refill_ext = r;
            double ms = -1.0;
            rc = probe(ms);
            if (rc) break;
            if (consider(ms)) { best = ms; chosen = r; }
        }
        ctx->refill_ext = chosen;
    }
    // 2. traversal variant (graph off): the refill probes' variant stands if its frames
    //    completed
    const int incumbent = ctx->variant;
    int best_v = best >= 0.0 ? incumbent : -1;
    for (int v = 0; v < wavefront_num_variants() && rc == RT_OK; v++) {
        if (v == incumbent && best_v >= 0) continue;
        ctx->variant = v;
        double ms = -1.0;
        rc = probe(ms);
        if (rc == RT_OK && consider(ms)) { best = ms; best_v = v; }
    }
    if (rc == RT_OK && best_v < 0) rc = fail(ctx, RT_ERR_STACK_OVERFLOW, "no traversal variant completed this frame");
    if (rc == RT_OK) ctx->variant = best_v;
    // 3-4. concurrent slot sets and the live-path count (per set) that hands the tail
    //    to the finishing launch, together: the best threshold depends on the number
    //    of sets (on the 1080p benchmark frame, 2 sets finishing below 262,144 paths
    //    beat 3 sets at every threshold, while at 65,536 paths 3 sets beat 2)
    if (rc == RT_OK) {
        // the fastest pair of the grid, which then has to beat the incumbent by the
        // margin (a margin between the pairs themselves would keep whichever of two
        // near-equal pairs came first)
        const uint32_t finish[] = {0u, 1u << 16, 1u << 17, 1u << 18};
        const int keep_n = ctx->streams;
        const uint32_t keep_f = ctx->finish;
        int grid_n = keep_n;
        uint32_t grid_f = keep_f;
        double grid_ms = -1.0;
        for (int n = 1; n <= wavefront_max_streams() && rc == RT_OK; n++) {
            for (uint32_t f : finish) {
                if (n == keep_n && f == keep_f) continue;
                ctx->streams = n;
                ctx->finish = f;
                double ms = -1.0;
                rc = probe(ms);
                if (rc) break;
                if (ms >= 0.0 && (grid_ms < 0.0 || ms < grid_ms)) { grid_ms = ms; grid_n = n; grid_f = f; }
            }
        }
        if (consider(grid_ms)) {
            best = grid_ms;
            ctx->streams = grid_n;
            ctx->finish = grid_f;
        } else {
            ctx->streams = keep_n;
            ctx->finish = keep_f;
        }
    }
    // 5. graph replay
    if (rc == RT_OK) {
        ctx->use_graph = true;
        double ms = -1.0;
        rc = probe(ms);
        if (rc == RT_OK) {
            if (consider(ms)) best = ms;
            else ctx->use_graph = false;
        }
    }
    // 6. the traversal variants again, under the settings chosen after them (the
    //    slot sets change how many rays each trace launch holds)
    if (rc == RT_OK) {
        int chosen = ctx->variant;
        for (int v = 0; v < wavefront_num_variants() && rc == RT_OK; v++) {
            if (v == chosen) continue;
            ctx->variant = v;
            double ms = -1.0;
            rc = probe(ms);
            if (rc == RT_OK && consider(ms)) { best = ms; chosen = v; }
        }
        ctx->variant = chosen;
    }
    ctx->stage_timing = timing;
    if (rc) {
        ctx->variant = keep_variant;
        ctx->use_graph = keep_graph;
        ctx->refill_ext = keep_re;
        ctx->refill_sh = keep_rs;
        ctx->streams = keep_streams;
        ctx->finish = keep_finish;
        return rc;
    }
    if (out) {
        out->traversal_variant = uint32_t(ctx->variant);
        out->graph = ctx->use_graph ? 1u : 0u;
        out->extend_refill = ctx->refill_ext;
        out->shadow_refill = ctx->refill_sh;
        out->wavefront_streams = uint32_t(ctx->streams);
        out->wavefront_finish = ctx->finish;
        out->frame_ms = best;
    }
    return RT_OK;
}

int rt_set_option(rt_ctx* ctx, uint32_t option, uint64_t value) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (option == RT_OPTION_STAGE_TIMING) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_STAGE_TIMING takes 0 or 1");
        ctx->stage_timing = value != 0;
        ctx->timer.used = 0;
        return RT_OK;
    }
    if (option == RT_OPTION_TRAVERSAL_VARIANT) {
        if (value >= uint64_t(wavefront_num_variants()))
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_TRAVERSAL_VARIANT out of range");
        ctx->variant = int(value);
        return RT_OK;
    }
    if (option == RT_OPTION_EXTEND_REFILL || option == RT_OPTION_SHADOW_REFILL) {
        if (value < 1 || value > 64) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "refill thresholds are 1 .. 64");
        (option == RT_OPTION_EXTEND_REFILL ? ctx->refill_ext : ctx->refill_sh) = uint32_t(value);
        return RT_OK;
    }
    if (option == RT_OPTION_GRAPH) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_GRAPH takes 0 or 1");
        ctx->use_graph = value != 0;
        return RT_OK;
    }
    if (option == RT_OPTION_MEGAKERNEL_PERSISTENT) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_MEGAKERNEL_PERSISTENT takes 0 or 1");
        ctx->mega_persistent = value != 0;
        return RT_OK;
    }
    if (option == RT_OPTION_BVH_COLLAPSE) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_BVH_COLLAPSE takes 0 or 1");
        ctx->collapse = uint32_t(value);
        return RT_OK;
    }
    if (option == RT_OPTION_BVH_TREELET_PASSES) {
        if (value > 8) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_BVH_TREELET_PASSES takes 0 .. 8");
        ctx->treelets = uint32_t(value);
        return RT_OK;
    }
    if (option == RT_OPTION_BLAS_LEAF_SIZE) {
        if (value < 1 || value > uint64_t(kMaxLeafPrims))
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_BLAS_LEAF_SIZE takes 1 .. 16");
        ctx->leaf_max = uint32_t(value);
        return RT_OK;
    }
    if (option == RT_OPTION_WAVEFRONT_FINISH) {
        if (value > 0xFFFFFFFFull) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_WAVEFRONT_FINISH is a 32-bit count");
        ctx->finish = uint32_t(value);
        return RT_OK;
    }
    if (option == RT_OPTION_WAVEFRONT_STREAMS) {
        if (value < 1 || value > uint64_t(wavefront_max_streams()))
            return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_WAVEFRONT_STREAMS takes 1 .. 4");
        ctx->streams = int(value);
        return RT_OK;
    }
    if (option == RT_OPTION_OUTPUT_BGRA) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_OUTPUT_BGRA takes 0 or 1");
        ctx->out_bgra = value != 0;
        return RT_OK;
    }
    if (option == RT_OPTION_FLATTEN_IDENTITY_INSTANCES) {
        if (value > 1) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_FLATTEN_IDENTITY_INSTANCES takes 0 or 1");
        ctx->flatten = value != 0;
        return RT_OK;
    }
    return fail(ctx, RT_ERR_INVALID_ARGUMENT, "unknown option");
}

int rt_get_stage_times(rt_ctx* ctx, rt_stage_times* out) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (!out) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    if (!ctx->stage_timing) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "RT_OPTION_STAGE_TIMING is off");
    DeviceGuard g(ctx->device);
    RT_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    const StageTimer& t = ctx->timer;
    double* slot[4] = {&out->extend_ms, &out->shade_ms, &out->shadow_ms, &out->resolve_ms};
    for (size_t i = 1; i < t.used; i++) {
        float ms = 0.f;
        RT_HIP(ctx, hipEventElapsedTime(&ms, t.events[i - 1], t.events[i]), "hipEventElapsedTime");
        if (t.stage[i] >= 0 && t.stage[i] < 4) *slot[t.stage[i]] += ms;
        if (t.stage[i] == 0) out->iterations++;
        out->total_ms += ms;
    }
    return RT_OK;
}

int rt_trace_ray_queries(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, const rt_ray* d_rays, uint32_t count,
                         uint32_t ray_flags, rt_hit* d_hits) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    if (count && (!d_rays || !d_hits)) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "rays/hits must be non-NULL");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    const bool any_hit = b != nullptr && !(ray_flags & RT_RAY_FLAG_OPAQUE) && it->second.any_non_opaque;
    if (b) {
        int rc = validate_bindings(ctx, b);
        if (rc) return rc;
    }
    DeviceGuard g(ctx->device);
    FrameArgs fa;
    std::memset(&fa, 0, sizeof(fa));
    fa.tlas = tlas_view(it->second, ctx->d_err);
    uint32_t seed_b = 0;
    if (b) {
        int rc = upload_textures(ctx, b);
        if (rc) return rc;
        make_frame(ctx, it->second, b, nullptr, 1, 1, &fa);
        seed_b = fa.seed_b;
    }
    int rc = launch_ray_queries(fa.tlas, fa.sv, d_rays, count, ray_flags | (any_hit ? 0u : RT_RAY_FLAG_OPAQUE), d_hits,
                                seed_b, ctx->stream, any_hit);
    if (rc) return hip_fail(ctx, hipError_t(rc), "ray query launch");
    return RT_OK;
}

int rt_trace_primary(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, uint32_t width, uint32_t height, rt_hit* d_hits) {
    if (!ctx) return fail(nullptr, RT_ERR_INVALID_ARGUMENT, "ctx is NULL");
    int rc = validate_bindings(ctx, b);
    if (rc) return rc;
    if (!d_hits || width == 0 || height == 0) return fail(ctx, RT_ERR_INVALID_ARGUMENT, "hits/size invalid");
    auto it = ctx->tlas.find(tlas);
    if (it == ctx->tlas.end()) return fail(ctx, RT_ERR_INVALID_HANDLE, "unknown TLAS handle");
    DeviceGuard g(ctx->device);
    rc = upload_textures(ctx, b);
    if (rc) return rc;
    FrameArgs fa;
    make_frame(ctx, it->second, b, nullptr, width, height, &fa);
    rc = launch_primary(fa, d_hits, ctx->stream, !fa.fully_opaque);
    if (rc) return hip_fail(ctx, hipError_t(rc), "primary launch");
    return RT_OK;
}

}  // extern "C"
