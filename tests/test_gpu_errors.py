"""Failure handling of the C ABI on the GPU: invalid arguments, handles and
references come back as status codes (the reference panics or leaves them
undefined, e.g. crates/libs/asset_loader/src/acceleration_structures.rs:33,109);
nothing faults."""
import ctypes as C

import numpy as np
import pytest
import torch

from rustracer_amd import layouts as L, rt
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu


def test_invalid_blas_inputs(gpu_ctx):
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    out = (C.c_uint64 * 1)()
    bad = rt.BlasInput(None, 128, 2, None, 0, 0, 1, 1)  # NULL buffers with triangles
    assert lib.rt_build_blas_batch(h, C.byref(bad), 1, out) == -1
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    vb, ib = ds.buffers.vertices.data_ptr(), ds.buffers.indices.data_ptr()
    assert lib.rt_build_blas_batch(h, C.byref(rt.BlasInput(vb, 6, 3, ib, 0, 0, 1, 1)), 1, out) == -1  # stride
    assert lib.rt_build_blas_batch(h, C.byref(rt.BlasInput(vb, 128, 3, ib, 0, 2, 1, 1)), 1, out) == -1  # offset
    # an index above max_vertex is reported after the build, not read
    rc = lib.rt_build_blas_batch(h, C.byref(rt.BlasInput(vb, 128, 0, ib, 0, 0, 2, 1)), 1, out)
    assert rc == -1 and b"max_vertex" in lib.rt_last_error(h)


def test_invalid_tlas_reference_is_rejected(gpu_ctx):
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    inst = np.zeros(2, L.INSTANCE)
    inst["transform"][:, 0, 0] = inst["transform"][:, 1, 1] = inst["transform"][:, 2, 2] = 1
    inst["custom_index_and_mask"] = 0xFF << 24
    inst["acceleration_structure_reference"] = 0xDEAD0000  # not a live BLAS
    d = rt.to_device(inst)
    out = C.c_uint64()
    assert lib.rt_build_tlas(h, C.c_void_p(d.data_ptr()), 2, C.byref(out)) == -1
    assert b"not a live BLAS" in lib.rt_last_error(h)


def test_destroyed_blas_address_is_no_longer_accepted(gpu_ctx):
    sc = G.host_scene("cornell_sphere")
    buf = rt.Buffers(sc)
    blases, tlas = rt.create_as(gpu_ctx, sc, buf)
    addr = blases[0].address()
    inst = np.array(sc.instances[:1], copy=True)
    inst["acceleration_structure_reference"] = addr
    gpu_ctx.lib.rt_destroy_blas(gpu_ctx.handle, blases[0].handle)
    d = rt.to_device(inst)
    out = C.c_uint64()
    assert gpu_ctx.lib.rt_build_tlas(gpu_ctx.handle, C.c_void_p(d.data_ptr()), 1, C.byref(out)) == -1
    # the TLAS built before keeps its BLAS memory alive and still traces
    ds = G.DeviceScene.__new__(G.DeviceScene)
    ds.scene, ds.buffers, ds.blases, ds.tlas = sc, buf, blases, tlas
    acc = rt.new_accumulation(sc.width, sc.height)
    rt.trace_rays(gpu_ctx, tlas, buf, sc.ubo, acc, sc.width, sc.height)
    gpu_ctx.sync()
    assert float(acc.sum()) > 0


def test_trace_argument_checks(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    b = ds.buffers.bindings(sc.ubo)
    assert lib.rt_trace_rays(h, 987654, C.byref(b), C.c_void_p(acc.data_ptr()), None, 8, 8, 0, None) == -7
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b), None, None, 8, 8, 0, None) == -1
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b), C.c_void_p(acc.data_ptr()), None, 0, 8, 0, None) == -1
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b), C.c_void_p(acc.data_ptr()), None, 8, 8, 7, None) == -1
    b2 = ds.buffers.bindings(sc.ubo)
    b2.num_materials = 0
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b2), C.c_void_p(acc.data_ptr()), None, 8, 8, 0, None) == -1
    # HasSky needs the skybox binding, with square faces
    sky = sc.ubo.copy()
    sky["has_sky"] = 1
    b3 = ds.buffers.bindings(sky)
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b3), C.c_void_p(acc.data_ptr()), None, 8, 8, 0, None) == -1
    assert b"skybox" in lib.rt_last_error(h)
    faces = rt.to_device(np.zeros((6, 2, 4, 4), np.uint8))
    not_square = rt.Texture(faces.data_ptr(), 4, 2, L.FMT_RGBA8_SRGB, 0, 0, 1, 1)
    b3.skybox = C.cast(C.pointer(not_square), C.c_void_p)
    assert lib.rt_trace_rays(h, ds.tlas.handle, C.byref(b3), C.c_void_p(acc.data_ptr()), None, 8, 8, 0, None) == -1


def test_refit_and_option_argument_checks(gpu_ctx):
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    assert lib.rt_refit_tlas(h, 987654) == -7  # unknown TLAS handle
    assert lib.rt_set_option(h, 11, 1) == -1  # no such option
    assert lib.rt_set_option(h, rt.OPTION_BVH_TREELET_PASSES, 9) == -1  # 0 .. 8
    assert b"0 .. 8" in lib.rt_last_error(h)
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    deep = sc.ubo.copy()
    deep["number_of_bounces"] = 65  # the adjoint keeps one shadow bit per bounce in 64 bits
    g = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(rt.RtError) as e:
        rt.trace_rays_adjoint(gpu_ctx, ds.tlas, ds.buffers, deep, g, sc.width, sc.height, {})
    assert e.value.code == -6 and "64 bounces" in str(e.value)


def test_out_of_range_material_is_reported_not_read(gpu_ctx):
    # a PrimInfo pointing past the material buffer: the counting trace reports it
    sc = G.host_scene("cornell_sphere")
    bad = sc.__class__(**{**sc.__dict__})
    bad.prim_infos = sc.prim_infos.copy()
    bad.prim_infos["material_id"][3] = 999
    ds = G.DeviceScene(gpu_ctx, bad)
    acc = rt.new_accumulation(sc.width, sc.height)
    with pytest.raises(rt.RtError) as e:
        rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, stats=True)
    assert e.value.code == -1
