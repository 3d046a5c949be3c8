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
