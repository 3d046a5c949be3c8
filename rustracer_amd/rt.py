"""Host-side mirror of the reference's ray-tracing operator surface, over the C ABI.

The reference renderer enters its ray-tracing path at two points
(crates/libs/asset_loader/src/acceleration_structures.rs, crates/examples/gltf_viewer/src/main.rs):
  create_as(context, doc, buffers, flags) -> (Vec<AccelerationStructure>, TopAS)   :79-129
  create_top_as(context, doc, blases, flags, cmd)   -> TopAS                         :136-249
  record_raytracing_commands -> trace_rays(&sbt, W, H)                          main.rs:244-270
This module offers the same three operations (same argument meaning: one BLAS per
geometry id with its OPAQUE flag, one instance per scene instance with custom index =
geometry id, a trace writing the rgba32f accumulation image) on top of
librt_gfx950.so.  Device memory comes from PyTorch (HIP allocations); the library
enqueues on PyTorch's current stream so torch copies and traces are ordered.

There is no CPU fallback: if librt_gfx950.so is missing or no GPU is visible,
loading raises.
"""
import ctypes as C
import os

import numpy as np

from . import layouts as L

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librt_gfx950.so")

RT_OK = 0
STATUS = {0: "RT_OK", -1: "RT_ERR_INVALID_ARGUMENT", -2: "RT_ERR_HIP", -3: "RT_ERR_OUT_OF_MEMORY",
          -4: "RT_ERR_STACK_OVERFLOW", -5: "RT_ERR_QUEUE_OVERFLOW", -6: "RT_ERR_UNSUPPORTED",
          -7: "RT_ERR_INVALID_HANDLE"}
MODE_MEGAKERNEL, MODE_WAVEFRONT = 0, 1
GEOMETRY_OPAQUE = 0x1
RAY_FLAG_NONE, RAY_FLAG_OPAQUE, RAY_FLAG_NO_OPAQUE = 0x0, 0x1, 0x2
RAY_FLAG_TERMINATE_ON_FIRST_HIT, RAY_FLAG_SKIP_CLOSEST_HIT = 0x4, 0x8

# every symbol include/rt/rt.h declares
EXPORTS = ["rt_create", "rt_destroy", "rt_last_error", "rt_sync", "rt_version", "rt_build_blas_batch",
           "rt_build_blas", "rt_blas_device_address", "rt_destroy_blas", "rt_build_tlas", "rt_destroy_tlas",
           "rt_trace_rays", "rt_trace_ray_queries", "rt_trace_primary", "rt_blas_info", "rt_tlas_info",
           "rt_set_option", "rt_get_stage_times", "rt_tlas_flat_info", "rt_trace_rays_adjoint",
           "rt_trace_rays_batch", "rt_autotune", "rt_skin_vertices", "rt_refit_blas_batch", "rt_refit_tlas",
           "rt_trace_rays_noalloc"]


class RtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class BlasInput(C.Structure):
    _fields_ = [("vertex_data", C.c_void_p), ("vertex_stride", C.c_uint32), ("max_vertex", C.c_uint32),
                ("index_data", C.c_void_p), ("first_vertex", C.c_uint32), ("primitive_offset", C.c_uint32),
                ("primitive_count", C.c_uint32), ("geometry_flags", C.c_uint32)]


class Texture(C.Structure):
    _fields_ = [("data", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("wrap_s", C.c_uint32), ("wrap_t", C.c_uint32), ("mag_filter", C.c_uint32),
                ("min_filter", C.c_uint32)]


class Bindings(C.Structure):
    _fields_ = [("ubo", C.c_void_p), ("vertices", C.c_void_p), ("indices", C.c_void_p),
                ("prim_infos", C.c_void_p), ("materials", C.c_void_p), ("dlights", C.c_void_p),
                ("num_dlights", C.c_uint32), ("plights", C.c_void_p), ("num_plights", C.c_uint32),
                ("textures", C.c_void_p), ("num_textures", C.c_uint32), ("num_vertices", C.c_uint32),
                ("num_indices", C.c_uint32), ("num_prim_infos", C.c_uint32), ("num_materials", C.c_uint32),
                ("skybox", C.c_void_p)]


class Tuning(C.Structure):
    _fields_ = [("traversal_variant", C.c_uint32), ("graph", C.c_uint32), ("extend_refill", C.c_uint32),
                ("shadow_refill", C.c_uint32), ("wavefront_streams", C.c_uint32), ("wavefront_finish", C.c_uint32),
                ("frame_ms", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("tlas", C.c_uint64), ("bindings", C.c_void_p), ("accum", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32)]


class TraceStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("rays", "primary_rays", "extension_rays", "shadow_rays", "node_visits",
                                            "tri_tests", "instance_visits", "shade_events", "anyhit_calls",
                                            "tex_fetches", "light_candidates", "stack_overflows")] + \
               [("kernel_ms", C.c_double)] + \
               [(n, C.c_uint64) for n in ("shadow_node_visits", "shadow_tri_tests", "shadow_instance_visits",
                                           "grad_taps", "grad_atomics")]


OPTION_STAGE_TIMING = 1
OPTION_FLATTEN_IDENTITY_INSTANCES = 2
OPTION_TRAVERSAL_VARIANT = 3
OPTION_GRAPH = 4
OPTION_EXTEND_REFILL = 5
OPTION_SHADOW_REFILL = 6
OPTION_OUTPUT_BGRA = 7
OPTION_WAVEFRONT_STREAMS = 8
OPTION_WAVEFRONT_FINISH = 9
OPTION_BLAS_LEAF_SIZE = 10
OPTION_BVH_COLLAPSE = 12
OPTION_MEGAKERNEL_PERSISTENT = 13
OPTION_BVH_TREELET_PASSES = 14

# The value every option has on a new context (rt_create).
DEFAULT_OPTIONS = {OPTION_STAGE_TIMING: 0, OPTION_FLATTEN_IDENTITY_INSTANCES: 1, OPTION_TRAVERSAL_VARIANT: 0,
                   OPTION_GRAPH: 0, OPTION_EXTEND_REFILL: 16, OPTION_SHADOW_REFILL: 1, OPTION_OUTPUT_BGRA: 0,
                   OPTION_WAVEFRONT_STREAMS: 2, OPTION_WAVEFRONT_FINISH: 65536, OPTION_BLAS_LEAF_SIZE: 2,
                   OPTION_BVH_COLLAPSE: 1, OPTION_MEGAKERNEL_PERSISTENT: 1, OPTION_BVH_TREELET_PASSES: 2}
TONE_MAP_DEFAULT, TONE_MAP_UNCHARTED, TONE_MAP_HEJL_RICHARD, TONE_MAP_ACES, TONE_MAP_SRGB = 0, 1, 2, 3, 4
NUM_TRAVERSAL_VARIANTS = 4


class StageTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("total_ms", "extend_ms", "shade_ms", "shadow_ms", "resolve_ms")] + \
               [("iterations", C.c_uint32), ("_pad", C.c_uint32)]


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_uint64), ("tris", C.c_uint64), ("root", C.c_uint32), ("num_nodes", C.c_uint32),
                ("first_prim", C.c_uint32), ("num_prims", C.c_uint32), ("bounds", C.c_float * 6)]


_lib = None


def load_library(path=LIB_PATH):
    """Loads librt_gfx950.so and declares the C signatures.  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `python -m rustracer_amd.build`")
    lib = C.CDLL(path)
    i, u, vp, u64 = C.c_int, C.c_uint32, C.c_void_p, C.c_uint64
    sig = {
        "rt_create": (i, [C.POINTER(vp), i, vp]), "rt_destroy": (None, [vp]), "rt_last_error": (C.c_char_p, [vp]),
        "rt_sync": (i, [vp]), "rt_version": (C.c_char_p, []),
        "rt_build_blas_batch": (i, [vp, vp, u, C.POINTER(u64)]), "rt_build_blas": (i, [vp, vp, C.POINTER(u64)]),
        "rt_blas_device_address": (i, [vp, u64, C.POINTER(u64)]), "rt_destroy_blas": (None, [vp, u64]),
        "rt_build_tlas": (i, [vp, vp, u, C.POINTER(u64)]), "rt_destroy_tlas": (None, [vp, u64]),
        "rt_trace_rays": (i, [vp, u64, vp, vp, vp, u, u, u, vp]),
        "rt_trace_rays_noalloc": (i, [vp, u64, vp, vp, vp, u, u]),
        "rt_skin_vertices": (i, [vp, vp, vp, u, vp, u]),
        "rt_refit_blas_batch": (i, [vp, vp, u, C.POINTER(u64)]),
        "rt_refit_tlas": (i, [vp, u64]),
        "rt_trace_ray_queries": (i, [vp, u64, vp, vp, u, u, vp]),
        "rt_trace_primary": (i, [vp, u64, vp, u, u, vp]),
        "rt_blas_info": (i, [vp, u64, vp]), "rt_tlas_info": (i, [vp, u64, vp]),
        "rt_set_option": (i, [vp, u, u64]), "rt_get_stage_times": (i, [vp, vp]),
        "rt_tlas_flat_info": (i, [vp, u64, vp]),
        "rt_trace_rays_adjoint": (i, [vp, u64, vp, vp, u, u, vp, vp, vp, vp]),
        "rt_trace_rays_batch": (i, [vp, vp, u, u, vp]),
        "rt_autotune": (i, [vp, u64, vp, u, u, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _torch():
    import torch
    return torch


def to_device(arr, device="cuda"):
    """Copies a numpy array (any dtype) into a HIP allocation; returns the torch tensor."""
    torch = _torch()
    a = np.ascontiguousarray(arr)
    t = torch.from_numpy(a.view(np.uint8).reshape(-1).copy()).to(device)
    return t


class Context:
    """rt_create / rt_destroy.  Enqueues on PyTorch's current stream of `device`."""

    def __init__(self, device=0):
        torch = _torch()
        if not torch.cuda.is_available():
            raise RuntimeError("no HIP device visible; the gfx950 library has no CPU path")
        self.lib = load_library()
        self.device = device
        with torch.cuda.device(device):
            stream = torch.cuda.current_stream(device).cuda_stream
        h = C.c_void_p()
        rc = self.lib.rt_create(C.byref(h), device, C.c_void_p(stream))
        if rc != RT_OK:
            raise RtError(rc, self.lib.rt_last_error(None).decode())
        self.handle = h

    def check(self, rc):
        if rc != RT_OK:
            raise RtError(rc, self.lib.rt_last_error(self.handle).decode())

    def sync(self):
        self.check(self.lib.rt_sync(self.handle))

    def autotune(self, tlas, buffers, ubo, width, height):
        """rt_autotune: tunes the traversal variant, refill thresholds and graph
        replay for this frame and keeps them on the context.  Returns a dict."""
        b = buffers.bindings(ubo)
        t = Tuning()
        self.check(self.lib.rt_autotune(self.handle, tlas.handle, C.byref(b), width, height, C.byref(t)))
        return {n: getattr(t, n) for n, _ in Tuning._fields_}

    def set_option(self, option, value):
        """rt_set_option (OPTION_STAGE_TIMING: event timing of every pipeline stage)."""
        self.check(self.lib.rt_set_option(self.handle, C.c_uint32(option), C.c_uint64(value)))

    def reset_options(self):
        """Every option back to its value on a new context."""
        for k, v in DEFAULT_OPTIONS.items():
            self.set_option(k, v)

    def stage_times(self):
        """rt_get_stage_times: device milliseconds per stage of the last trace."""
        st = StageTimes()
        self.check(self.lib.rt_get_stage_times(self.handle, C.byref(st)))
        return {n: getattr(st, n) for n, _ in StageTimes._fields_ if n != "_pad"}

    def close(self):
        if getattr(self, "handle", None):
            self.lib.rt_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Buffers:
    """The renderer's device buffers (globals.rs Buffers::new + textures): vertex,
    index, geometry-info, material, directional/point light buffers, textures."""

    def __init__(self, scene, device="cuda"):
        self.scene = scene
        self.vertices = to_device(scene.vertices, device)
        self.indices = to_device(scene.indices, device)
        self.prim_infos = to_device(scene.prim_infos, device)
        self.materials = to_device(scene.materials, device)
        self.dlights = to_device(scene.dlights, device) if len(scene.dlights) else None
        self.plights = to_device(scene.plights, device) if len(scene.plights) else None
        self.textures = [to_device(t.data, device) for t in scene.textures]
        self._tex = (Texture * max(1, len(scene.textures)))()
        for k, t in enumerate(scene.textures):
            h, w = t.data.shape[:2]
            self._tex[k] = Texture(self.textures[k].data_ptr(), w, h, t.format, t.wrap_s, t.wrap_t, t.mag_filter,
                                   t.min_filter)
        # skybox (binding 12): six square faces +x, -x, +y, -y, +z, -z, end to end
        self.skybox, self._sky = None, None
        sky = getattr(scene, "skybox", None)
        if sky is not None:
            if sky.data.ndim != 4 or sky.data.shape[0] != 6 or sky.data.shape[1] != sky.data.shape[2]:
                raise ValueError("the skybox needs six square faces, shape (6, w, w, 4)")
            self.skybox = to_device(sky.data, device)
            w = sky.data.shape[1]
            self._sky = Texture(self.skybox.data_ptr(), w, w, sky.format, sky.wrap_s, sky.wrap_t, sky.mag_filter,
                                sky.min_filter)

    def set_texture(self, k, tensor):
        """Points texture slot k at `tensor` (device, (h, w, 4), the slot's format:
        uint8 for RGBA8, float32 for RGBA32F): the renderer then reads the tensor's
        memory directly, e.g. a parameter texture an optimiser updates in place."""
        t = self.scene.textures[k]
        h, w = tensor.shape[:2]
        self.textures[k] = tensor
        self._tex[k] = Texture(tensor.data_ptr(), w, h, t.format, t.wrap_s, t.wrap_t, t.mag_filter, t.min_filter)

    def bindings(self, ubo, keep=None):
        """rt_bindings for one trace.  The UBO copy it points at lives in this object
        (replaced by the next call) or, when `keep` is a list, is appended to it."""
        s = self.scene
        u = np.array(ubo, copy=True)
        if keep is None:
            self._ubo = u
        else:
            keep.append(u)
        return Bindings(u.ctypes.data, self.vertices.data_ptr(), self.indices.data_ptr(),
                        self.prim_infos.data_ptr(), self.materials.data_ptr(),
                        self.dlights.data_ptr() if self.dlights is not None else None, len(s.dlights),
                        self.plights.data_ptr() if self.plights is not None else None, len(s.plights),
                        C.addressof(self._tex) if s.textures else None, len(s.textures), len(s.vertices),
                        len(s.indices), len(s.prim_infos), len(s.materials),
                        C.addressof(self._sky) if self._sky is not None else None)


class TopAS:
    def __init__(self, ctx, handle, instances):
        self.ctx, self.handle, self.instances = ctx, handle, instances

    def info(self):
        i = BvhInfo()
        self.ctx.check(self.ctx.lib.rt_tlas_info(self.ctx.handle, self.handle, C.byref(i)))
        return i

    def flat_info(self):
        """rt_tlas_flat_info: the world-space BVH over the identity instances."""
        i = BvhInfo()
        self.ctx.check(self.ctx.lib.rt_tlas_flat_info(self.ctx.handle, self.handle, C.byref(i)))
        return i

    def __del__(self):
        try:
            if self.ctx.handle:
                self.ctx.lib.rt_destroy_tlas(self.ctx.handle, self.handle)
        except Exception:
            pass


class Blas:
    def __init__(self, ctx, handle):
        self.ctx, self.handle = ctx, handle

    def address(self):
        a = C.c_uint64()
        self.ctx.check(self.ctx.lib.rt_blas_device_address(self.ctx.handle, self.handle, C.byref(a)))
        return a.value

    def info(self):
        i = BvhInfo()
        self.ctx.check(self.ctx.lib.rt_blas_info(self.ctx.handle, self.handle, C.byref(i)))
        return i

    def __del__(self):
        try:
            if self.ctx.handle:
                self.ctx.lib.rt_destroy_blas(self.ctx.handle, self.handle)
        except Exception:
            pass


def blas_inputs(scene, buffers, vertices=None):
    """primitive_to_vk_geometry (acceleration_structures.rs:17-76) for every geometry id;
    `vertices` (a device tensor in the vertex layout) replaces the vertex buffer, as the
    reference uses its animation buffer for skinned geometry (:23-27)."""
    geo = scene.geometries
    ins = (BlasInput * len(geo))()
    vb = (buffers.vertices if vertices is None else vertices).data_ptr()
    ib = buffers.indices.data_ptr()
    for g in range(len(geo)):
        first_vertex, index_offset, count, opaque = (int(x) for x in geo[g])
        idx = scene.indices[index_offset:index_offset + 3 * count]
        max_vertex = int(idx.max()) if count else 0
        ins[g] = BlasInput(vb, L.VERTEX.itemsize, max_vertex, ib, first_vertex, 4 * index_offset, count,
                           GEOMETRY_OPAQUE if opaque else 0)
    return ins


def create_as(ctx, scene, buffers):
    """create_as (acceleration_structures.rs:79-129): one BLAS per geometry id, built in
    one batch, then create_top_as.  Returns (blases, TopAS)."""
    ins = blas_inputs(scene, buffers)
    handles = (C.c_uint64 * len(ins))()
    ctx.check(ctx.lib.rt_build_blas_batch(ctx.handle, ins, len(ins), handles))
    blases = [Blas(ctx, h) for h in handles]
    return blases, create_top_as(ctx, scene, blases)


def refit_as(ctx, scene, buffers, blases, vertices=None, tlas=None):
    """rt_refit_blas_batch over the BLASes of create_as with the scene's (or the given)
    vertex data, then rt_refit_tlas on `tlas` when one is given, else create_top_as
    again (the reference rebuilds its TLAS).  Returns the TopAS."""
    ins = blas_inputs(scene, buffers, vertices)
    handles = (C.c_uint64 * len(blases))(*[b.handle for b in blases])
    ctx.check(ctx.lib.rt_refit_blas_batch(ctx.handle, ins, len(ins), handles))
    if tlas is not None:
        ctx.check(ctx.lib.rt_refit_tlas(ctx.handle, tlas.handle))
        return tlas
    return create_top_as(ctx, scene, blases)


def skin_vertices(ctx, vertices_in, vertices_out, skins):
    """rt_skin_vertices (AnimationCompute.comp): vertices_* are device tensors holding
    vertex records (128 B each; may be the same tensor), skins a float32 device tensor
    of shape (num_skins, 256, 4, 4) holding column-major mat4 (skins[s, j, c] is
    column c).  Enqueued on the context stream."""
    n = vertices_in.numel() * vertices_in.element_size() // L.VERTEX.itemsize
    if vertices_out.numel() * vertices_out.element_size() < n * L.VERTEX.itemsize:
        raise ValueError("vertices_out is smaller than vertices_in")
    if skins is not None and (skins.dtype != _torch().float32 or not skins.is_contiguous() or skins.shape[1:] != (256, 4, 4)):
        raise ValueError("skins must be a contiguous float32 tensor of shape (num_skins, 256, 4, 4)")
    ns = 0 if skins is None else skins.shape[0]
    ctx.check(ctx.lib.rt_skin_vertices(ctx.handle, C.c_void_p(vertices_in.data_ptr()), C.c_void_p(vertices_out.data_ptr()),
                                       n, C.c_void_p(skins.data_ptr()) if ns else None, ns))


def create_top_as(ctx, scene, blases):
    """create_top_as (acceleration_structures.rs:136-249): the scene's instance records
    with the BLAS reference replaced by the BLAS device address, uploaded and built."""
    inst = np.array(scene.instances, copy=True)
    addr = np.array([b.address() for b in blases], np.uint64)
    inst["acceleration_structure_reference"] = addr[scene.instances["acceleration_structure_reference"].astype(np.int64)]
    d_inst = to_device(inst)
    out = C.c_uint64()
    ctx.check(ctx.lib.rt_build_tlas(ctx.handle, C.c_void_p(d_inst.data_ptr()), len(inst), C.byref(out)))
    return TopAS(ctx, out.value, d_inst)


def new_accumulation(width, height, device="cuda"):
    """The rgba32f accumulation image (crates/libs/app/src/lib.rs:316-325)."""
    torch = _torch()
    return torch.zeros((height, width, 4), dtype=torch.float32, device=device)


def trace_rays(ctx, tlas, buffers, ubo, accum, width, height, mode=MODE_MEGAKERNEL, stats=False, out=None):
    """record_raytracing_commands + trace_rays (main.rs:244-270): runs the pipeline for
    a width x height launch into `accum` (torch float32 (h, w, 4) on the device) and,
    when `out` (torch uint8 (h, w, 4) on the device) is given, the tone-mapped display
    image (RayTracing.rgen:132-166).  Returns a dict of counters when stats=True (the
    counting kernel variant)."""
    b = buffers.bindings(ubo)
    st = TraceStats() if stats else None
    if out is not None and (out.dtype != _torch().uint8 or not out.is_contiguous() or out.numel() != width * height * 4):
        raise ValueError("out must be a contiguous uint8 tensor of width * height * 4 elements")
    ctx.check(ctx.lib.rt_trace_rays(ctx.handle, tlas.handle, C.byref(b), C.c_void_p(accum.data_ptr()),
                                    C.c_void_p(out.data_ptr()) if out is not None else None,
                                    width, height, mode, C.byref(st) if st is not None else None))
    if st is None:
        return None
    return {n: getattr(st, n) for n, _ in TraceStats._fields_}


def trace_rays_noalloc(ctx, tlas, buffers, ubo, accum, width, height, out=None):
    """rt_trace_rays_noalloc: the persistent-megakernel trace of trace_rays with no
    device allocation (its pixel counter and the texture descriptors in the library's
    per-GPU scratch array).  Same image as trace_rays in any mode."""
    b = buffers.bindings(ubo)
    if out is not None and (out.dtype != _torch().uint8 or not out.is_contiguous() or out.numel() != width * height * 4):
        raise ValueError("out must be a contiguous uint8 tensor of width * height * 4 elements")
    ctx.check(ctx.lib.rt_trace_rays_noalloc(ctx.handle, tlas.handle, C.byref(b), C.c_void_p(accum.data_ptr()),
                                            C.c_void_p(out.data_ptr()) if out is not None else None, width, height))


def trace_rays_batch(ctx, items, mode=MODE_WAVEFRONT, stats=False):
    """rt_trace_rays_batch over items = [(tlas, buffers, ubo, accum, width, height)]."""
    keep, binds = [], []
    arr = (BatchItem * max(1, len(items)))()
    for k, (tlas, buffers, ubo, accum, w, h) in enumerate(items):
        b = buffers.bindings(ubo, keep=keep)
        binds.append(b)
        arr[k] = BatchItem(tlas.handle, C.addressof(b), accum.data_ptr(), w, h)
    st = TraceStats() if stats else None
    ctx.check(ctx.lib.rt_trace_rays_batch(ctx.handle, arr, len(items), mode, C.byref(st) if st is not None else None))
    if st is None:
        return None
    return {n: getattr(st, n) for n, _ in TraceStats._fields_}


def trace_rays_adjoint(ctx, tlas, buffers, ubo, grad_accum, width, height, grads, stats=False,
                       grad_point_intensity=None, grad_directional_intensity=None):
    """rt_trace_rays_adjoint: adds to grads[k] (float32 device tensors shaped like
    RGBA32F texture k) the gradient of sum(grad_accum * accum) of the trace with the
    same buffers and UBO.  `grads` maps texture index -> tensor.  The optional
    float32 device tensors grad_point_intensity (one entry per point light) and
    grad_directional_intensity (one per directional light) receive the gradient with
    respect to the lights' intensities."""
    b = buffers.bindings(ubo)
    n = len(buffers.scene.textures)
    ptrs = (C.c_void_p * max(1, n))()
    for k, g in grads.items():
        ptrs[k] = g.data_ptr()
    for t, cnt in ((grad_point_intensity, len(buffers.scene.plights)),
                   (grad_directional_intensity, len(buffers.scene.dlights))):
        if t is not None and (t.dtype != _torch().float32 or not t.is_contiguous() or t.numel() < cnt):
            raise ValueError("light-intensity gradients need a contiguous float32 tensor with one entry per light")
    st = TraceStats() if stats else None
    ctx.check(ctx.lib.rt_trace_rays_adjoint(
        ctx.handle, tlas.handle, C.byref(b), C.c_void_p(grad_accum.data_ptr()), width, height, ptrs,
        C.c_void_p(grad_point_intensity.data_ptr()) if grad_point_intensity is not None else None,
        C.c_void_p(grad_directional_intensity.data_ptr()) if grad_directional_intensity is not None else None,
        C.byref(st) if st is not None else None))
    if st is None:
        return None
    return {n_: getattr(st, n_) for n_, _ in TraceStats._fields_}


def trace_ray_queries(ctx, tlas, rays, flags=0, buffers=None, ubo=None):
    """Traverses rays (numpy RAY records or a device tensor) -> numpy HIT records."""
    torch = _torch()
    d_rays = to_device(rays) if isinstance(rays, np.ndarray) else rays
    n = d_rays.numel() // L.RAY.itemsize
    d_hits = torch.empty(n * L.HIT.itemsize, dtype=torch.uint8, device=d_rays.device)
    b = buffers.bindings(ubo) if buffers is not None else None
    ctx.check(ctx.lib.rt_trace_ray_queries(ctx.handle, tlas.handle, C.byref(b) if b is not None else None,
                                           C.c_void_p(d_rays.data_ptr()), n, flags, C.c_void_p(d_hits.data_ptr())))
    ctx.sync()
    return d_hits.cpu().numpy().view(L.HIT)


def trace_primary(ctx, tlas, buffers, ubo, width, height):
    torch = _torch()
    d_hits = torch.empty(width * height * L.HIT.itemsize, dtype=torch.uint8, device="cuda")
    b = buffers.bindings(ubo)
    ctx.check(ctx.lib.rt_trace_primary(ctx.handle, tlas.handle, C.byref(b), width, height,
                                       C.c_void_p(d_hits.data_ptr())))
    ctx.sync()
    return d_hits.cpu().numpy().view(L.HIT)


def read_nodes(ctx, info):
    """Copies a BVH's node array (64-byte quantised nodes) to the host and decodes it
    into NODE records (fp32 child boxes exactly as traversal decodes them)."""
    ctx.sync()
    raw = np.frombuffer(read_device(info.nodes, QNODE.itemsize * info.num_nodes), QNODE)
    return decode_nodes(raw)


def decode_nodes(raw):
    """QNode -> NODE: plane = origin + q * 2^exp in fp32 (q * 2^exp is exact, so the
    single rounding matches the device's fmaf)."""
    out = np.zeros(len(raw), NODE)
    scale = np.ldexp(np.float32(1.0), raw["exp"].astype(np.int32)).astype(np.float32)  # (n, 3)
    q = raw["q"].astype(np.float32).reshape(-1, 6, 4)
    for a, (lo, hi) in enumerate((("lo_x", "hi_x"), ("lo_y", "hi_y"), ("lo_z", "hi_z"))):
        o = raw["origin"][:, a:a + 1].astype(np.float32)
        s = scale[:, a:a + 1]
        out[lo] = (o + q[:, 2 * a] * s).astype(np.float32)
        out[hi] = (o + q[:, 2 * a + 1] * s).astype(np.float32)
    out["child"] = raw["child"]
    out["count"] = raw["count"].astype(np.uint32)
    return out


QNODE = np.dtype([("origin", "<f4", 3), ("exp", "i1", 3), ("_pad0", "u1"), ("q", "u1", 24), ("child", "<u4", 4),
                  ("count", "u1", 4), ("_pad1", "<u4")])
assert QNODE.itemsize == 64


NODE = np.dtype([("lo_x", "<f4", 4), ("hi_x", "<f4", 4), ("lo_y", "<f4", 4), ("hi_y", "<f4", 4),
                 ("lo_z", "<f4", 4), ("hi_z", "<f4", 4), ("child", "<u4", 4), ("count", "<u4", 4)])


def read_device(ptr, nbytes):
    """Copies nbytes from a device pointer to a host bytes object (hipMemcpy D2H)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    buf = (C.c_uint8 * nbytes)()
    rc = hip.hipMemcpy(C.addressof(buf), C.c_void_p(ptr), nbytes, 2)
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed: {rc}")
    return bytes(buf)
