"""Textured, instanced workload: eight meshes instanced with random rotation and
scale, RGBA8 noise textures (sRGB base colour with alpha, UNORM metallic-roughness,
sRGB specular-glossiness), MASK any-hit on ~30% of the instances and
specular-glossiness on ~20%.  The gfx950 image must match the FP64 oracle (relRMSE
of the accumulated radiance <= 1e-3) and the wavefront must match the megakernel
bit for bit.  Exercises RayTracing.rahit (:44-104, MASK), RayTracing.rchit
(:166-239, textures and the spec-gloss conversion) and texture indexing
(crates/libs/asset_loader/src/texture.rs:66-78)."""
import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import rt
from rustracer_amd import scenes as S
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu
REL_RMSE_GATE = 1e-3


def _render(ctx, ds, mode, stats=True):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=stats)
    ctx.sync()
    return acc.cpu().numpy().astype(np.float64), st


def test_small_textured_scene_matches_oracle_on_every_pixel(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "textured_small", lambda: S.textured_instances(
        width=240, height=136, spp=4, bounces=6, num_instances=200, tex_size=256, tri_range=(2000, 20000)))
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, sw = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    assert st["rays"] == sw["rays"]
    assert st["anyhit_calls"] > 0 and st["tex_fetches"] > 0 and st["stack_overflows"] == 0
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("textured small relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
    assert abs(st["rays"] - ost["rays"]) <= 2e-3 * ost["rays"]


def test_full_textured_workload_matches_oracle_on_a_stratified_subsample(gpu_ctx):
    """1000 instances, 2048^2 textures, 1920x1080, 32 spp, depth 6 on the GPU; the
    oracle renders every 24th pixel of every 24th row."""
    ds = G.device_scene(gpu_ctx, "textured_instances")
    sc = ds.scene
    g, st = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert st["anyhit_calls"] > 0 and st["stack_overflows"] == 0
    ys, xs = np.meshgrid(np.arange(11, sc.height, 24), np.arange(11, sc.width, 24), indexing="ij")
    xy = np.stack([xs.reshape(-1), ys.reshape(-1)], 1)
    o, _ = B.OracleScene(sc).render_pixels(xy)
    rel = G.rel_rmse(g[xy[:, 1], xy[:, 0], :3], o[:, :3])
    print("textured full relRMSE", rel, "rays", st["rays"])
    assert rel <= REL_RMSE_GATE
