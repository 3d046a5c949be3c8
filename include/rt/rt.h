/*
 * rt.h — C ABI of the MI355X (gfx950) ray-tracing library.
 *
 * The library replaces, for the reference renderer (KaminariOS/rustracer,
 * gltf_viewer), the Vulkan ray-tracing pipeline: the acceleration-structure build
 * that vkCmdBuildAccelerationStructuresKHR performs and the raygen / closest-hit /
 * any-hit / miss shader stages that vkCmdTraceRaysKHR runs.  It consumes the
 * renderer's own buffers (layouts in rt_layout.h) and writes the same rgba32f
 * accumulation image.
 *
 * Conventions
 *  - Every call returns RT_OK (0) or a negative rt_status; rt_last_error() gives a
 *    message.  No call aborts the process.
 *  - Scene buffers and images are caller-owned HIP device allocations.  BVH memory
 *    is library-owned, behind rt_blas / rt_tlas handles freed explicitly (the
 *    reference frees acceleration structures in Drop, crates/libs/vulkan/src/
 *    ray_tracing/acceleration_structure.rs:270-277).
 *  - One context per host thread.  Calls enqueue work on the context's HIP stream
 *    and return; rt_sync() gives the fence semantics the reference gets from
 *    fence.wait (crates/libs/asset_loader/src/acceleration_structures.rs:237-243).
 */
#ifndef RT_H
#define RT_H

#include <stddef.h>
#include <stdint.h>
#include "rt_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum rt_status {
    RT_OK = 0,
    RT_ERR_INVALID_ARGUMENT = -1,
    RT_ERR_HIP = -2,
    RT_ERR_OUT_OF_MEMORY = -3,
    RT_ERR_STACK_OVERFLOW = -4,
    RT_ERR_QUEUE_OVERFLOW = -5,
    RT_ERR_UNSUPPORTED = -6,
    RT_ERR_INVALID_HANDLE = -7
} rt_status;

typedef struct rt_ctx rt_ctx;
typedef uint64_t rt_blas; /* 0 is never a valid handle */
typedef uint64_t rt_tlas;

/* ---------------------------------------------------------------- context */

/* Creates a context on HIP device `hip_device`.  `hip_stream` is the hipStream_t
 * the context enqueues on; NULL means the device's default (null) stream, so the
 * calls are ordered with whatever else the caller enqueues there. */
int rt_create(rt_ctx** out, int hip_device, void* hip_stream);
void rt_destroy(rt_ctx* ctx);
/* Message of the last failed call on this context ("" if none).  ctx may be NULL
 * for failures of rt_create. */
const char* rt_last_error(const rt_ctx* ctx);
/* Blocks until all work enqueued on the context has finished. */
int rt_sync(rt_ctx* ctx);
/* Library version string. */
const char* rt_version(void);

/* ------------------------------------------------- bottom-level (per mesh) */

enum { RT_GEOMETRY_OPAQUE = 0x1 };

/* One triangle geometry, as primitive_to_vk_geometry fills
 * VkAccelerationStructureGeometryTrianglesDataKHR + BuildRangeInfo
 * (crates/libs/asset_loader/src/acceleration_structures.rs:17-76):
 * vertex i of triangle p is at
 *   vertex_data + (first_vertex + index_data[primitive_offset/4 + 3p + i]) * vertex_stride
 * and holds three floats (R32G32B32_SFLOAT). */
typedef struct rt_blas_input {
    const void* vertex_data;   /* device address of the vertex buffer */
    uint32_t vertex_stride;    /* bytes; 128 for rt_vertex */
    uint32_t max_vertex;       /* highest index value used (bounds check) */
    const uint32_t* index_data;/* device address of the u32 index buffer */
    uint32_t first_vertex;     /* PrimInfo.v_offset */
    uint32_t primitive_offset; /* bytes into index_data: PrimInfo.i_offset * 4 */
    uint32_t primitive_count;  /* triangles */
    uint32_t geometry_flags;   /* RT_GEOMETRY_OPAQUE iff the material is OPAQUE */
} rt_blas_input;

/* Builds one BLAS per input, all in one batched build (one Morton sort over all
 * triangles), as create_as records one build per primitive into one command
 * buffer (acceleration_structures.rs:79-129,
 * crates/libs/vulkan/src/ray_tracing/acceleration_structure.rs:199-221).
 * `inputs` is a host array; out[i] receives the handle of inputs[i]. */
int rt_build_blas_batch(rt_ctx* ctx, const rt_blas_input* inputs, uint32_t count, rt_blas* out);
int rt_build_blas(rt_ctx* ctx, const rt_blas_input* input, rt_blas* out);
/* 64-bit reference to place in rt_instance.acceleration_structure_reference
 * (vkGetAccelerationStructureDeviceAddressKHR,
 * crates/libs/vulkan/src/ray_tracing/acceleration_structure.rs:160-168). */
int rt_blas_device_address(rt_ctx* ctx, rt_blas blas, uint64_t* out);
void rt_destroy_blas(rt_ctx* ctx, rt_blas blas);

/* ------------------------------------------------------ animated scenes */

/* Skinning of the vertex buffer for an animation step: the compute pass the
 * reference dispatches before rebuilding its acceleration structures
 * (crates/examples/gltf_viewer/shaders/AnimationCompute.comp:14-38, run by
 * crates/examples/gltf_viewer/src/main.rs:383-386).  d_out[i] = d_in[i], and for
 * skin_index >= 0: transform = sum of weights[k] * bones[joints[k]] of that skin;
 * pos = transform * (pos, 1); normal = normalize(transform * (normal, 0));
 * tangent = normalize(transform * (tangent.xyz, 0)) with tangent.w kept.  d_skins
 * holds num_skins * 256 column-major mat4 (the reference's Skin array).  d_in may be
 * d_out; both 16-byte aligned device buffers of `count` vertices.  Enqueued on the
 * context stream; a skin index or joint out of range leaves that vertex unskinned
 * and makes the next rt_sync return RT_ERR_INVALID_ARGUMENT. */
int rt_skin_vertices(rt_ctx* ctx, const rt_vertex* d_in, rt_vertex* d_out, uint32_t count,
                     const float* d_skins, uint32_t num_skins);

/* Refit of the BLASes of one rt_build_blas_batch call (all its handles, in its
 * order) to new vertex data with the same triangles: the reference rebuilds its
 * BLASes from the skinned vertex buffer every animation step (main.rs:388-396 ->
 * create_as); a refit keeps the tree and recomputes its boxes bottom-up, which is
 * faster and gives the same hits (a looser tree if the mesh deforms far).
 * inputs[i] must have the build's primitive_count and describe the same triangles
 * (vertex_data may be another buffer, e.g. the skinned one).  TLASes built over
 * these BLASes must be refit (rt_refit_tlas) or rebuilt (rt_build_tlas, as the
 * reference rebuilds its TLAS) afterwards.  Synchronises the context stream. */
int rt_refit_blas_batch(rt_ctx* ctx, const rt_blas_input* inputs, uint32_t count, const rt_blas* blases);

/* --------------------------------------------------- top-level (instances) */

/* Builds a TLAS over `count` instances held in device memory in the 64-byte
 * VkAccelerationStructureInstanceKHR layout (create_top_as,
 * acceleration_structures.rs:136-249).  Every acceleration_structure_reference
 * must come from rt_blas_device_address() of a live BLAS. */
int rt_build_tlas(rt_ctx* ctx, const rt_instance* d_instances, uint32_t count, rt_tlas* out);
/* Refits a TLAS to the present geometry of the BLASes it references, after
 * rt_refit_blas_batch moved their vertices (same instances, transforms and
 * topology): the flattened BVH's triangle records take their positions again from
 * their BLAS records, and the node boxes of the flattened BVH and of the TLAS over
 * the other instances are recomputed bottom-up, padded and quantised as the build
 * does.  The trees are kept, so a far deformation gives looser trees but the same
 * hits.  Replaces the create_top_as of the reference's animation step
 * (crates/examples/gltf_viewer/src/main.rs:377-404) when only the vertices moved;
 * changed transforms need rt_build_tlas.  Synchronises the stream. */
int rt_refit_tlas(rt_ctx* ctx, rt_tlas tlas);
void rt_destroy_tlas(rt_ctx* ctx, rt_tlas tlas);

/* ----------------------------------------------------------- trace (frame) */

enum { RT_FMT_RGBA8_UNORM = 0, RT_FMT_RGBA8_SRGB = 1, RT_FMT_RGBA32F = 2 };
enum { RT_WRAP_REPEAT = 0, RT_WRAP_CLAMP_TO_EDGE = 1, RT_WRAP_MIRRORED_REPEAT = 2 };
enum { RT_FILTER_NEAREST = 0, RT_FILTER_LINEAR = 1 };

/* One combined image sampler of the bindless texture array (binding 6).
 * Texel (x, y) is at data + (y * width + x) * texel_size.  Sampled at LOD 0. */
typedef struct rt_texture {
    const void* data;          /* device pointer */
    uint32_t width, height;
    uint32_t format;           /* RT_FMT_* */
    uint32_t wrap_s, wrap_t;   /* RT_WRAP_* */
    uint32_t mag_filter;       /* RT_FILTER_* (LOD 0 samples with the mag filter) */
    uint32_t min_filter;
} rt_texture;

/* The descriptor bindings of the reference pipeline
 * (crates/examples/gltf_viewer/shaders/lib/RayTracingCommons.glsl:123-136). */
typedef struct rt_bindings {
    const rt_ubo* ubo;                 /* HOST pointer, read when the call enqueues (binding 2) */
    const rt_vertex* vertices;         /* binding 3 */
    const uint32_t* indices;           /* binding 4 */
    const rt_prim_info* prim_infos;    /* binding 5, indexed by instance custom index */
    const rt_material_raw* materials;  /* binding 9 */
    const rt_light* dlights;           /* binding 10 (directional, used by miss) */
    uint32_t num_dlights;
    const rt_light* plights;           /* binding 11 (point, used by closest-hit) */
    uint32_t num_plights;
    const rt_texture* textures;        /* binding 6: HOST array of descriptors */
    uint32_t num_textures;
    uint32_t num_vertices;             /* element counts, used for bounds checks */
    uint32_t num_indices;
    uint32_t num_prim_infos;
    uint32_t num_materials;
    /* binding 12 (SKYBOX_BIND, RayTracing.rmiss:10): HOST pointer to one descriptor
     * whose data holds the six cube faces +x, -x, +y, -y, +z, -z (the order
     * cubumap.rs:19-27 sorts the face files into), each width x width texels
     * (width 1..16384), one after another (the reference's `collector` buffer).
     * Sampled at LOD 0 with mag_filter and seamless cube edges; wrap modes are
     * ignored, as for any cube image.  Read only when ubo.has_sky != 0, and
     * required then. */
    const rt_texture* skybox;
} rt_bindings;

enum { RT_MODE_MEGAKERNEL = 0, RT_MODE_WAVEFRONT = 1 };

/* Counters of one trace call (filled when a non-NULL pointer is passed; the call
 * then runs the counting variant of the kernels and synchronises). */
typedef struct rt_trace_stats {
    uint64_t rays;            /* camera + extension + shadow ray segments */
    uint64_t primary_rays;    /* camera rays */
    uint64_t extension_rays;  /* bounce rays */
    uint64_t shadow_rays;
    uint64_t node_visits;     /* wide BVH nodes fetched (TLAS + BLAS) */
    uint64_t tri_tests;       /* ray-triangle tests */
    uint64_t instance_visits; /* TLAS leaf instances entered */
    uint64_t shade_events;    /* closest-hit shader invocations */
    uint64_t anyhit_calls;    /* any-hit shader invocations */
    uint64_t tex_fetches;     /* texture() calls */
    uint64_t light_candidates;/* RIS candidate light reads */
    uint64_t stack_overflows; /* rays whose traversal stack overflowed */
    double kernel_ms;         /* device time of the trace kernels */
    uint64_t shadow_node_visits;     /* part of node_visits spent on shadow rays */
    uint64_t shadow_tri_tests;       /* part of tri_tests spent on shadow rays */
    uint64_t shadow_instance_visits; /* part of instance_visits spent on shadow rays */
    uint64_t grad_taps;       /* adjoint: gradient contributions (texel or light) */
    uint64_t grad_atomics;    /* adjoint: float atomics they became after the per-wave sums */
} rt_trace_stats;

/* ------------------------------------------------------------ batch */

/* One frame of a batch: the arguments of one rt_trace_rays call. */
typedef struct rt_batch_item {
    rt_tlas tlas;
    const rt_bindings* bindings;  /* host pointer; its UBO and texture arrays too */
    float* accum;                 /* device rgba32f image, width x height */
    uint32_t width, height;
} rt_batch_item;

/* Traces `count` frames (different scenes, resolutions, UBOs) as one set of
 * persistent wavefront queue kernels: the frames' path slots are laid end to end
 * and every kernel serves all of them, so small frames still fill the GPU.  Each
 * accumulation image equals what rt_trace_rays(mode) gives for that item alone.
 * RT_MODE_MEGAKERNEL traces the items one after another.  The reference renders
 * one scene per vkCmdTraceRaysKHR (crates/examples/gltf_viewer/src/main.rs:244-270);
 * this is the many-scene form of the same call.  Synchronises the stream. */
int rt_trace_rays_batch(rt_ctx* ctx, const rt_batch_item* items, uint32_t count, uint32_t mode,
                        rt_trace_stats* stats);

/* ------------------------------------------------------------ adjoint */

/* Path-replay adjoint of rt_trace_rays: accumulates (adds, does not clear) the
 * gradient of  sum_pixels dot(d_grad_accum, accum)  with respect to
 *   - the texels of texture i, into grad_textures[i], for every texture with a
 *     non-NULL entry: such a texture must be RT_FMT_RGBA32F, and its gradient
 *     buffer is a device array in the same layout (width * height * 4 floats).
 *     Gradients flow into the base-colour texture (R, G, B) and the
 *     metallic-roughness texture (G roughness, B metallic) of metallic-roughness
 *     materials; other uses of a texture do not contribute;
 *   - the intensity of point light i, into d_grad_point_intensity[i] (device
 *     array of b->num_plights floats, or NULL), through the NEE term
 *     BRDF * light_weight * intensity * colour (RayTracing.rchit:347-348);
 *   - the intensity of directional light i, into d_grad_directional_intensity[i]
 *     (b->num_dlights floats, or NULL), through the miss shader's
 *     -cos * colour * intensity (RayTracing.rmiss:18-26).
 * Each sample is traced twice (path replay): once for its radiance and shadow
 * results, once more to differentiate it vertex by vertex; nothing else of the
 * path is stored.  Lobe choices, directions, sampling densities (the RIS light
 * weight included) and Russian roulette are held at their forward values, so the
 * result is the derivative of the image's estimate with the sampling detached, for
 * the same bindings and UBO as the forward trace.  d_grad_accum is the device
 * array dLoss/d(accum), width * height * 4 floats (alpha ignored); pixels whose
 * gradient is zero are skipped.  At most 64 bounces, and ubo.debug == 1 (the legacy
 * materials) is not differentiated (RT_ERR_UNSUPPORTED for both).  With `stats`,
 * its ray and traversal counters count both passes, and grad_taps / grad_atomics
 * count the gradient contributions and the float atomics they became.  There is no
 * adjoint in the reference; it serves an optimisation loop around the trace that
 * record_raytracing_commands issues (crates/examples/gltf_viewer/src/main.rs:244-270). */
int rt_trace_rays_adjoint(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, const float* d_grad_accum,
                          uint32_t width, uint32_t height, float* const* grad_textures, float* d_grad_point_intensity,
                          float* d_grad_directional_intensity, rt_trace_stats* stats);

/* ------------------------------------------------------------ options */

enum {
    /* 1: record device time per pipeline stage on every trace (rt_get_stage_times).
     * Default 0. */
    RT_OPTION_STAGE_TIMING = 1,
    /* 1 (default): rt_build_tlas copies the triangles of every instance whose
     * transform is exactly the identity into one world-space BVH, traversed before
     * the TLAS over the other instances.  Hits report the same instance, primitive
     * and barycentrics as the two-level path.  Applies to TLASes built afterwards. */
    RT_OPTION_FLATTEN_IDENTITY_INSTANCES = 2,
    /* Wavefront traversal kernels: 0 (default) LDS stack ring of 32 entries,
     * registers for 3 waves/SIMD (no spills); 1: 32 entries, 4 waves; 2: 32
     * entries, 5 waves (spills); 3: 64 entries (LDS holds 2 waves/SIMD).  Results
     * are identical; speed differs by scene (a ring too shallow for the BVH depth
     * restarts traversals). */
    RT_OPTION_TRAVERSAL_VARIANT = 3,
    /* 1: the wavefront trace captures one chunk of its queue iterations (8 x
     * extend, shade, shadow, resolve) as a HIP graph and replays it until no path
     * is alive (one launch per chunk).  Same images.  Ignored while
     * RT_OPTION_STAGE_TIMING is on.  Default 0. */
    RT_OPTION_GRAPH = 4,
    /* The wavefront traversal kernel fetches rays dynamically: a wave takes new
     * rays from the queue once fewer than this many of its 64 lanes are still
     * traversing (1 = only when all have finished).  Same images.  Default 16.
     * RT_OPTION_SHADOW_REFILL (1 .. 64) is accepted and has no effect: shadow rays
     * are fetched with the extension rays. */
    RT_OPTION_EXTEND_REFILL = 5,
    RT_OPTION_SHADOW_REFILL = 6,
    /* Byte order of rt_trace_rays' display image: 0 (default) R, G, B, A; 1: B, G,
     * R, A, the reference's storage image (the swapchain's B8G8R8A8_UNORM,
     * crates/libs/vulkan/src/swapchain.rs:39-48, crates/libs/app/src/lib.rs:305-312). */
    RT_OPTION_OUTPUT_BGRA = 7,
    /* The wavefront trace deals the frame's 8x8 pixel tiles (a batch: its frames) to
     * this many slot sets (1 .. 4), each running its own chain of queue kernels on
     * its own stream, so one set's kernels fill the CUs that another set's kernel
     * tails leave idle.  Same images.  RT_OPTION_STAGE_TIMING traces one set.
     * Default 2. */
    RT_OPTION_WAVEFRONT_STREAMS = 8,
    /* Once fewer paths than this are alive in a slot set, the wavefront trace runs
     * each of them to the end of its pixel's samples in one lane of one launch
     * (the megakernel's loop from the path's state, traversals stepped as in the
     * persistent megakernel) instead of further queue iterations, whose launches
     * would last as long as their slowest ray while most of the GPU idles.
     * 0: never.  Same images.  Ignored while
     * RT_OPTION_STAGE_TIMING is on.  Default 65536. */
    RT_OPTION_WAVEFRONT_FINISH = 9,
    /* Most triangles per BVH leaf (1 .. 16) of BLASes and flattened BVHs built
     * afterwards: a subtree with at most this many triangles becomes one leaf.
     * Same hits.  Default 2. */
    RT_OPTION_BLAS_LEAF_SIZE = 10,
    /* How BVHs built afterwards (BLASes, TLASes, flattened BVHs) turn the binary
     * Morton-order radix tree into 4-wide nodes: 1 (default) the collapse of least
     * surface-area-heuristic cost, solved exactly over the binary tree (leaves of at
     * most RT_OPTION_BLAS_LEAF_SIZE triangles where the heuristic prefers them); 0
     * the greedy collapse that opens the largest child.  Same hits.  The reference
     * asks its driver for high-quality trees (PREFER_FAST_TRACE,
     * crates/libs/vulkan/src/ray_tracing/acceleration_structure.rs:105-118). */
    RT_OPTION_BVH_COLLAPSE = 12,
    /* The megakernel mode: 1 (default) a grid of as many workgroups as can be
     * resident, each lane taking the next pixel from one counter (pixels numbered
     * 8x8 tile by tile, tiles in image order) when its pixel's samples are done, so
     * no wave waits on its slowest pixel, and stepping its traversals (the wave
     * steps the lanes that trace, then shades the others together), so no lane
     * waits each bounce for the slowest traversal of its wave; 0 one lane per pixel
     * in 16x8-pixel workgroups, each traversal run to its end.  Same images. */
    RT_OPTION_MEGAKERNEL_PERSISTENT = 13,
    /* Treelet restructurings of the binary tree of BVHs built afterwards (0 .. 8),
     * before the collapse to 4-wide nodes: bottom up, each node's treelet of up to
     * five subtrees is rebuilt as the binary tree of least surface-area-heuristic
     * cost over them (Karras & Aila, HPG 2013).  0: the Morton-order tree as
     * emitted, built in less than half the time (for scenes rebuilt every frame).
     * Same hits.  Default 2. */
    RT_OPTION_BVH_TREELET_PASSES = 14
};
int rt_set_option(rt_ctx* ctx, uint32_t option, uint64_t value);

/* Device time of the stages of the last trace, when RT_OPTION_STAGE_TIMING is on
 * (the analogue of the reference's GPU timestamp query pool,
 * crates/libs/vulkan/src/query.rs:8-73).  The megakernel reports all of its time
 * as extend_ms.  Synchronises the context stream. */
/* Autotuner: traces the frame (the bindings' UBO with at most 64 samples per pixel)
 * in wavefront mode and tunes, one after another, the refill threshold, the
 * traversal variant (stack depth, register target), the number of concurrent slot
 * sets together with the live-path count that hands the tail to one finishing
 * launch (every pair of 1 .. 4 sets and 0, 65,536, 131,072, 262,144 paths), graph
 * replay, and the traversal variant once more under those settings, keeping the
 * fastest settings on this context (the RT_OPTION_* values).  Each probe takes the
 * mean of three timed frames after a warm-up frame (of six when the three differ by
 * more than 3 %), and a setting
 * replaces the one it is compared with only when it is more than 1.5 % faster.
 * A variant whose stack ring gives up on a ray of this frame (restart budget) is
 * not a candidate.  Fills *out (may be NULL).  Synchronises the stream. */
typedef struct rt_tuning {
    uint32_t traversal_variant;
    uint32_t graph;
    uint32_t extend_refill;
    uint32_t shadow_refill;
    uint32_t wavefront_streams;
    uint32_t wavefront_finish;
    double frame_ms;  /* probe frame time with the chosen settings */
} rt_tuning;
int rt_autotune(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, uint32_t width, uint32_t height, rt_tuning* out);

typedef struct rt_stage_times {
    double total_ms;
    double extend_ms;   /* traversal: each iteration's extension rays together with the
                           previous bounce's shadow rays (one kernel) */
    double shade_ms;    /* NEE radiance, closest-hit / miss shaders, next rays,
                           regeneration, compaction */
    double shadow_ms;   /* 0: shadow rays are traced with the extension rays */
    double resolve_ms;  /* path-state initialisation */
    uint32_t iterations;
    uint32_t _pad;
} rt_stage_times;
int rt_get_stage_times(rt_ctx* ctx, rt_stage_times* out);

/* Runs the whole ray-tracing pipeline for a width x height launch, as
 * record_raytracing_commands + trace_rays do
 * (crates/examples/gltf_viewer/src/main.rs:244-270, crates/libs/vulkan/src/command.rs:374-392):
 * per pixel, ubo->number_of_samples samples of up to ubo->number_of_bounces traces,
 * accumulated into `accum` (rgba32f, row-major, width*height*4 floats):
 *   accum = (number_of_samples != total_number_of_samples ? accum : 0) + sum(radiance), alpha = 0.
 * `out_rgba8` (nullable, 4-byte aligned, width*height*4 bytes) receives the display
 * image of RayTracing.rgen:132-166: accum / total_number_of_samples tone-mapped by
 * ubo->tone_mapping_mode (0 default, 1 Uncharted 2, 2 Hejl-Richard, 3 ACES, else
 * gamma 2.2 only), replaced by the DISTANCE mapping's grey level or the HEAT
 * mapping's colour ramp of the pixel's shader-clock ticks (a HEAT frame with a
 * display image runs in the megakernel form whatever `mode` says: only there does
 * one thread run a pixel's whole raygen; the accumulation is the same), stored as
 * 8-bit UNORM with alpha 255 (byte order: RT_OPTION_OUTPUT_BGRA).
 * `mode` selects the plain megakernel or the wavefront pipeline; both produce the
 * same image bit for bit. */
int rt_trace_rays(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, float* accum,
                  uint8_t* out_rgba8, uint32_t width, uint32_t height, uint32_t mode,
                  rt_trace_stats* stats);

/* No counterpart in the reference: for callers that cannot allocate device memory
 * during a trace (tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and the same image as rt_trace_rays(..., RT_MODE_PERSISTENT_MEGAKERNEL
 * or RT_MODE_WAVEFRONT, NULL), but calls no hipMalloc / hipFree: it runs the
 * persistent megakernel, whose only scratch is its pixel counter, and that counter
 * and the frame's texture descriptors live in one device array of about 41 KB per
 * GPU, reserved when the library is loaded and kept until the process ends (its
 * address is looked up by rt_create).  Limits: at most 1024 textures
 * (RT_ERR_INVALID_ARGUMENT beyond), no stats, no stage timing, no graph replay,
 * ubo->debug == 1 is RT_ERR_UNSUPPORTED.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several contexts or streams) race.  Queued
 * on the context's stream like rt_trace_rays (a copy of the descriptors and a clear
 * of the counter, then the kernel); not meant for stream capture. */
int rt_trace_rays_noalloc(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, float* accum,
                          uint8_t* out_rgba8, uint32_t width, uint32_t height);

/* ---------------------------------------------------- ray queries (parity) */

/* Ray flags (gl_RayFlags*EXT values). */
enum {
    RT_RAY_FLAG_NONE = 0x00,
    RT_RAY_FLAG_OPAQUE = 0x01,
    RT_RAY_FLAG_NO_OPAQUE = 0x02,
    RT_RAY_FLAG_TERMINATE_ON_FIRST_HIT = 0x04,
    RT_RAY_FLAG_SKIP_CLOSEST_HIT = 0x08
};

typedef struct rt_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} rt_ray;

/* Result of one traversal: primitive_id == 0xFFFFFFFF on a miss. */
typedef struct rt_hit {
    float t;
    float u, v;              /* barycentrics of vertices 1 and 2 */
    uint32_t instance_id;    /* gl_InstanceID */
    uint32_t primitive_id;   /* gl_PrimitiveID */
    uint32_t custom_index;   /* gl_InstanceCustomIndexEXT */
    uint32_t _pad[2];
} rt_hit;

/* Traverses `count` rays (device arrays) with the closest-hit rule used by
 * traceRayEXT: nearest accepted hit, ties in t broken by the smaller
 * (instance_id, primitive_id).  `b` may be NULL when every geometry is opaque or
 * RT_RAY_FLAG_OPAQUE is set; otherwise it supplies the alpha-test inputs. */
int rt_trace_ray_queries(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b,
                         const rt_ray* d_rays, uint32_t count, uint32_t ray_flags,
                         rt_hit* d_hits);

/* Primary-ray hits of the camera in `b->ubo` (sample 0 of every pixel, jitter as
 * the raygen shader draws it): one rt_hit per pixel, row-major. */
int rt_trace_primary(rt_ctx* ctx, rt_tlas tlas, const rt_bindings* b, uint32_t width,
                     uint32_t height, rt_hit* d_hits);

/* ------------------------------------------------------------- inspection */

/* Layout of a built BLAS, for validity checks.  Node records are 64 bytes (4-wide;
 * child boxes as 8-bit steps of a per-axis power-of-two grid: origin float[3],
 * exponent int8[3], pad, planes uint8[24] (lo_x[4] hi_x[4] lo_y[4] hi_y[4] lo_z[4]
 * hi_z[4], plane = origin + q * 2^exponent in fp32), child uint32[4], count uint8[4],
 * pad); triangle records are 48 bytes. */
typedef struct rt_bvh_info {
    uint64_t nodes;          /* device pointer to the node array */
    uint64_t tris;           /* device pointer to the triangle/instance records */
    uint32_t root;           /* root node index */
    uint32_t num_nodes;      /* nodes reachable from root (contiguous from root) */
    uint32_t first_prim;     /* first record index of this structure */
    uint32_t num_prims;
    float bounds[6];         /* lo xyz, hi xyz */
} rt_bvh_info;

int rt_blas_info(rt_ctx* ctx, rt_blas blas, rt_bvh_info* out);
/* The TLAS proper: num_prims = instances reached through it; `tris` points at the
 * instance records (internal 256-byte layout). */
int rt_tlas_info(rt_ctx* ctx, rt_tlas tlas, rt_bvh_info* out);
/* The flattened BVH of a TLAS's identity instances (num_prims = 0 when there is
 * none).  Triangle records are 48 bytes: v0.w = primitive index, v1.w = instance
 * record index | 0x80000000 if opaque, v2.w = source record in the BLAS pool. */
int rt_tlas_flat_info(rt_ctx* ctx, rt_tlas tlas, rt_bvh_info* out);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* RT_H */
