"""The skybox miss branch (RayTracing.rmiss:27-33, HasSky = 1; the cube map of
crates/libs/asset_loader/src/cubumap.rs bound at SKYBOX_BIND): images whose camera
sees three cube faces, their edges and the corner between them, rendered on gfx950
with linear (sRGB RGBA8) and nearest (RGBA32F) filtering, and a loaded glTF room
under a sky.  The gfx950 image must match the FP64 oracle (relRMSE of the
accumulated radiance <= 1e-3) and the wavefront must match the megakernel bit for
bit."""
import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import gltf, layouts as L, rt
from rustracer_amd import scenes as S
from tests import gltf_kit as GK
from tests import gpu_kit as G
from tests import scene_kit as K

pytestmark = pytest.mark.gpu
REL_RMSE_GATE = 1e-3


def _sky(fmt, filt, w=32, seed=0):
    rng = np.random.default_rng(seed)
    if fmt == L.FMT_RGBA32F:
        data = (rng.random((6, w, w, 4)) * 2.0).astype(np.float32)
    else:
        data = rng.integers(0, 256, (6, w, w, 4), dtype=np.uint8)
    return S.Texture(data, fmt, mag_filter=filt, min_filter=filt)


def _sky_scene(fmt, filt):
    sc = K.plane_scene(width=160, height=120, spp=4, bounces=2, intensity=10.0)
    view = S.look_at_rh((0, 5, 0), (1, 6, 1), (0, 1, 0))  # toward the cube corner (+x, +y, +z)
    proj = S.perspective_vk(160 / 120, 120.0, 0.1, 100.0)
    sc.ubo = S.make_ubo(view, proj, samples=4, bounces=2)
    sc.ubo["has_sky"] = 1
    sc.skybox = _sky(fmt, filt)
    return sc


def _render(ctx, ds, mode):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True)
    ctx.sync()
    return acc.cpu().numpy().astype(np.float64), st


@pytest.mark.parametrize("fmt,filt", [(L.FMT_RGBA8_SRGB, L.FILTER_LINEAR), (L.FMT_RGBA32F, L.FILTER_NEAREST)])
def test_sky_corner_view_matches_oracle(gpu_ctx, fmt, filt):
    ds = G.device_scene(gpu_ctx, f"sky_corner_{fmt}_{filt}", lambda: _sky_scene(fmt, filt))
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, _ = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("sky corner relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
    # the sky is what most camera rays see: the image is not the constant 0.01 miss
    assert np.median(g[..., :3]) > 0.05 * ds.scene.ubo["number_of_samples"]


def test_gltf_room_under_sky_matches_oracle(gpu_ctx, tmp_path):
    path = GK.room().save(str(tmp_path / "room.glb"), glb=True)
    sky = _sky(L.FMT_RGBA8_SRGB, L.FILTER_LINEAR, w=64, seed=5)
    ds = G.device_scene(gpu_ctx, "gltf_room_sky", lambda: gltf.load(
        path, width=320, height=240, spp=8, bounces=6, eye=(0, 1.0, 6.5), target=(0, -1, -1), fov=60, skybox=sky))
    assert ds.scene.ubo["has_sky"] == 1
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, _ = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("gltf room sky relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
