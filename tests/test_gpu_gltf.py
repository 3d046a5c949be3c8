"""A glTF scene written by the test (textured floor, TRS node hierarchy, generated
flat normals, glass with KHR_materials_transmission / ior / volume, a MASK quad,
an emissive box, point, spot and directional lights), loaded with
rustracer_amd.gltf (crates/libs/asset_loader semantics) and rendered on gfx950: the
image must match the FP64 oracle (relRMSE of the accumulated radiance <= 1e-3)
and the wavefront must match the megakernel bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import gltf, rt
from tests import gltf_kit as GK
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu
REL_RMSE_GATE = 1e-3


def _render(ctx, ds, mode):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True)
    ctx.sync()
    return acc.cpu().numpy().astype(np.float64), st


def test_loaded_gltf_room_matches_oracle(gpu_ctx, tmp_path):
    path = GK.room().save(str(tmp_path / "room.glb"), glb=True)
    ds = G.device_scene(gpu_ctx, "gltf_room", lambda: gltf.load(
        path, width=320, height=240, spp=8, bounces=6, eye=(0, 1.0, 6.5), target=(0, -1, -1), fov=60))
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, sw = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    assert st["rays"] == sw["rays"]
    assert st["anyhit_calls"] > 0 and st["tex_fetches"] > 0 and st["stack_overflows"] == 0
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("gltf room relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
    assert abs(st["rays"] - ost["rays"]) <= 2e-3 * ost["rays"]


def test_skinned_gltf_renders_like_the_oracle(gpu_ctx, tmp_path):
    """A skinned glTF (two joints, the skin's joint matrices from the loader) skinned
    on the GPU (rt_skin_vertices), its acceleration structures built over the
    skinned vertices, and traced; the oracle skins the same vertices with its own
    restatement and renders them.  Lit by a point light above the box."""
    import torch
    from oracle import binding as OB
    from tests.test_gltf import _skinned_file
    from rustracer_amd import scenes as S
    path, *_ = _skinned_file(tmp_path)
    # aim the camera at the skinned box (its node's own transform does not place it)
    first = gltf.load(path)
    p = OB.skin_vertices(first.vertices, first.skins)[0]["pos"][:, :3].astype(np.float64)
    c, ext = p.mean(0), float((p.max(0) - p.min(0)).max())
    sc = gltf.load(path, width=96, height=96, spp=4, bounces=4, eye=tuple(c + ext * np.array([0.6, 0.8, 2.0])),
                   target=tuple(c), zfar=100.0 * ext)
    sc.plights = np.asarray([S.point_light(tuple(c + ext * np.array([0.8, 2.0, 1.6])), 8.0 * ext * ext)],
                            sc.plights.dtype)
    buf = rt.Buffers(sc)
    skinned = torch.empty_like(buf.vertices)
    rt.skin_vertices(gpu_ctx, buf.vertices, skinned, torch.from_numpy(sc.skins).cuda())
    gpu_ctx.sync()
    buf.vertices = skinned  # the animation buffer replaces the vertex buffer (acceleration_structures.rs:23-27)
    blases, tlas = rt.create_as(gpu_ctx, sc, buf)
    host = skinned.cpu().numpy().view(sc.vertices.dtype).reshape(-1)
    ref_v, bad = OB.skin_vertices(sc.vertices, sc.skins)
    assert bad == 0
    assert np.allclose(host["pos"], ref_v["pos"], atol=2e-5)
    acc = rt.new_accumulation(sc.width, sc.height)
    rt.trace_rays(gpu_ctx, tlas, buf, sc.ubo, acc, sc.width, sc.height, mode=rt.MODE_WAVEFRONT)
    gpu_ctx.sync()
    g = acc.cpu().numpy().astype(np.float64)
    assert g[..., :3].max() > 1.0  # the box is lit and in view
    sc_o = sc.__class__(**{**sc.__dict__})
    sc_o.vertices = host.copy()  # the oracle traces the GPU-skinned vertices (checked above)
    o, _ = OB.OracleScene(sc_o).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("skinned gltf relRMSE", rel)
    assert rel <= REL_RMSE_GATE
    del blases, tlas
