"""Animated scenes: rt_skin_vertices against the oracle's skinning
(AnimationCompute.comp:14-38), and rt_refit_blas_batch after skinning (the
reference rebuilds its acceleration structures from the skinned vertex buffer
every animation step, main.rs:377-404): the refit BVH is valid in float, and the
image traced through it is the image traced through a fresh build over the same
skinned vertices, bit for bit, and matches the FP64 oracle; rt_refit_tlas likewise
for the flattened BVH and the TLAS over the instances."""
import numpy as np
import pytest
import torch

from oracle import binding as B
from rustracer_amd import bvhcheck, layouts as L, rt, scenes as S
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu


def _host_vertices(t):
    return t.cpu().numpy().view(L.VERTEX)


def test_skinning_matches_oracle_in_place_and_out_of_place(gpu_ctx):
    rng = np.random.default_rng(7)
    n = 4096
    v = np.zeros(n, L.VERTEX)
    v["pos"][:, :3] = rng.uniform(-5, 5, (n, 3))
    nrm = rng.normal(size=(n, 3))
    v["normal"][:, :3] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    v["tangent"] = np.concatenate([rng.normal(size=(n, 3)), rng.choice([-1.0, 1.0], (n, 1))], 1)
    v["color"] = rng.uniform(0, 1, (n, 4))
    v["uv0"], v["uv1"] = rng.uniform(0, 1, (n, 2)), rng.uniform(0, 1, (n, 2))
    v["skin_index"] = rng.integers(-1, 2, n)
    v["joints"] = rng.integers(0, 256, (n, 4))
    w = rng.uniform(0, 1, (n, 4))
    v["weights"] = w / w.sum(1, keepdims=True)
    skins = np.zeros((2, 256, 4, 4), np.float32)
    for s in range(2):
        for j in range(256):
            m = np.eye(4)
            q = np.linalg.qr(rng.normal(size=(3, 3)))[0]
            m[:3, :3] = q * rng.uniform(0.5, 1.5)
            m[:3, 3] = rng.uniform(-2, 2, 3)
            skins[s, j] = m.T
    ref, bad = B.skin_vertices(v, skins)
    assert bad == 0
    d_in = rt.to_device(v)
    d_out = torch.empty_like(d_in)
    d_sk = torch.from_numpy(skins).cuda()
    rt.skin_vertices(gpu_ctx, d_in, d_out, d_sk)
    gpu_ctx.sync()
    got = _host_vertices(d_out)
    for f in ("color", "weights", "joints", "uv0", "uv1", "skin_index"):
        assert np.array_equal(got[f], v[f]), f
    assert np.array_equal(got["tangent"][:, 3], v["tangent"][:, 3])
    assert np.allclose(got["pos"][:, :3], ref["pos"][:, :3], atol=5e-5, rtol=1e-5)
    assert np.allclose(got["normal"][:, :3], ref["normal"][:, :3], atol=2e-6)
    assert np.allclose(got["tangent"][:, :3], ref["tangent"][:, :3], atol=2e-6)
    unsk = v["skin_index"] < 0
    assert got[unsk].tobytes() == v[unsk].tobytes()
    rt.skin_vertices(gpu_ctx, d_in, d_in, d_sk)  # in place
    gpu_ctx.sync()
    assert _host_vertices(d_in).tobytes() == got.tobytes()


def test_out_of_range_skin_index_is_reported_and_left_unskinned(gpu_ctx):
    v = np.zeros(3, L.VERTEX)
    v["pos"][:, :3] = 1.0
    v["skin_index"] = (0, 1, -1)  # one skin given
    v["weights"] = (1, 0, 0, 0)
    d = rt.to_device(v)
    out = torch.empty_like(d)
    rt.skin_vertices(gpu_ctx, d, out, torch.from_numpy(S.bend_skins(0.5)).cuda())
    with pytest.raises(rt.RtError) as e:
        gpu_ctx.sync()
    assert e.value.code == -1 and "skin index" in str(e.value)  # RT_ERR_INVALID_ARGUMENT
    got = _host_vertices(out)
    assert got[1].tobytes() == v[1].tobytes()
    gpu_ctx.sync()  # the flag was cleared


@pytest.mark.parametrize("angle", [0.4, 1.2])
def test_refit_after_skinning_equals_a_fresh_build(gpu_ctx, angle):
    sc = S.skinned_cornell()
    buf = rt.Buffers(sc)
    blases, tlas = rt.create_as(gpu_ctx, sc, buf)
    skinned = torch.empty_like(buf.vertices)
    rt.skin_vertices(gpu_ctx, buf.vertices, skinned, torch.from_numpy(S.bend_skins(angle)).cuda())
    tlas2 = rt.refit_as(gpu_ctx, sc, buf, blases, vertices=skinned)
    gpu_ctx.sync()
    # the refit trees are valid for the skinned triangles
    errs = bvhcheck.check_blases(gpu_ctx, blases, [int(c) for c in sc.geometries["prim_count"]])
    assert all(not e for e in errs), errs
    assert not bvhcheck.check_tlas(gpu_ctx, tlas2, len(sc.instances))
    # the same scene built from scratch over the skinned vertices
    sc2 = S.skinned_cornell()
    sc2.vertices = _host_vertices(skinned).copy()
    assert not np.array_equal(sc2.vertices["pos"], sc.vertices["pos"])
    ds2 = G.DeviceScene(gpu_ctx, sc2)
    buf.vertices = skinned  # the trace reads the skinned vertices (normals for shading)
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        a = rt.new_accumulation(sc.width, sc.height)
        b = rt.new_accumulation(sc.width, sc.height)
        rt.trace_rays(gpu_ctx, tlas2, buf, sc.ubo, a, sc.width, sc.height, mode=mode)
        rt.trace_rays(gpu_ctx, ds2.tlas, ds2.buffers, sc2.ubo, b, sc.width, sc.height, mode=mode)
        gpu_ctx.sync()
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    o, _ = B.OracleScene(sc2).render()
    rel = G.rel_rmse(a.cpu().numpy()[..., :3].astype(np.float64), o[..., :3])
    assert rel <= 1e-3, rel


@pytest.mark.parametrize("flatten", [True, False])
def test_tlas_refit_equals_a_fresh_build(gpu_ctx, flatten):
    """rt_refit_tlas after the BLAS refit: the flattened BVH (identity instances) or
    the TLAS over the instances (flattening off) keeps its tree and takes the skinned
    geometry; its image equals the image of a fresh BLAS + TLAS build."""
    gpu_ctx.set_option(rt.OPTION_FLATTEN_IDENTITY_INSTANCES, int(flatten))
    try:
        sc = S.skinned_cornell()
        buf = rt.Buffers(sc)
        blases, tlas = rt.create_as(gpu_ctx, sc, buf)
        assert (tlas.flat_info().num_prims > 0) == flatten
        skinned = torch.empty_like(buf.vertices)
        rt.skin_vertices(gpu_ctx, buf.vertices, skinned, torch.from_numpy(S.bend_skins(1.2)).cuda())
        assert rt.refit_as(gpu_ctx, sc, buf, blases, vertices=skinned, tlas=tlas) is tlas
        gpu_ctx.sync()
        assert not bvhcheck.check_tlas(gpu_ctx, tlas, len(sc.instances))
        sc2 = S.skinned_cornell()
        sc2.vertices = _host_vertices(skinned).copy()
        ds2 = G.DeviceScene(gpu_ctx, sc2, flatten=flatten)
        buf.vertices = skinned
        for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
            a = rt.new_accumulation(sc.width, sc.height)
            b = rt.new_accumulation(sc.width, sc.height)
            rt.trace_rays(gpu_ctx, tlas, buf, sc.ubo, a, sc.width, sc.height, mode=mode)
            rt.trace_rays(gpu_ctx, ds2.tlas, ds2.buffers, sc2.ubo, b, sc.width, sc.height, mode=mode)
            gpu_ctx.sync()
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    finally:
        gpu_ctx.set_option(rt.OPTION_FLATTEN_IDENTITY_INSTANCES, 1)


def test_refit_argument_errors(gpu_ctx):
    sc = S.skinned_cornell()
    buf = rt.Buffers(sc)
    blases, _ = rt.create_as(gpu_ctx, sc, buf)
    ins = rt.blas_inputs(sc, buf)
    lib = gpu_ctx.lib
    import ctypes as C
    h = (C.c_uint64 * len(blases))(*[b.handle for b in blases])
    assert lib.rt_refit_blas_batch(gpu_ctx.handle, ins, len(ins) - 1, h) != 0  # not the whole batch
    rev = (C.c_uint64 * len(blases))(*[b.handle for b in reversed(blases)])
    assert lib.rt_refit_blas_batch(gpu_ctx.handle, ins, len(ins), rev) != 0  # order
    ins[0].primitive_count += 1
    assert lib.rt_refit_blas_batch(gpu_ctx.handle, ins, len(ins), h) != 0  # topology
    ins[0].primitive_count -= 1
    assert lib.rt_refit_blas_batch(gpu_ctx.handle, ins, len(ins), h) == 0
