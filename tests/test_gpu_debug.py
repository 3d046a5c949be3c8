"""ubo.debug == 1 on gfx950: the legacy Lambertian / fuzzy-metal / Schlick-dielectric
scatter of RayTracing.rchit:437-475, drawn from the LCG of Ray.RandomSeed (which the
camera's RandomInUnitDisk also draws from, aperture 0 or not).  Plane scenes must
give the closed form samples x colour x 0.01; the dielectric under an emitter and
the Cornell box must match the FP64 oracle (relRMSE <= 1e-3); the wavefront must
match the megakernel bit for bit.  The adjoint refuses debug == 1."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import rt
from rustracer_amd import scenes as S
from tests import debug_kit as D
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu
REL_RMSE_GATE = 1e-3


def _render(ctx, ds, mode, ubo=None):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, sc.ubo if ubo is None else ubo, acc, sc.width, sc.height, mode=mode,
                       stats=True)
    ctx.sync()
    return acc.cpu().numpy().astype(np.float64), st


@pytest.mark.parametrize("kind", ["lambert", "metal", "glass"])
def test_debug_planes_closed_form(gpu_ctx, kind):
    ds = G.device_scene(gpu_ctx, f"debug_{kind}", lambda: D.plane(kind))
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, _ = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    c = np.array([0.9, 0.8, 0.7]).astype(np.float32).astype(np.float64)
    assert np.allclose(g[..., :3], 8 * c * 0.01, rtol=1e-6, atol=0)
    assert st["extension_rays"] == 32 * 32 * 8


def test_debug_dielectric_under_emitter_matches_oracle(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "debug_glass_emitter", lambda: D.plane("glass", width=64, height=64, spp=16,
                                                                          emitter=True))
    g, _ = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, _ = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w)
    o, _ = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("debug glass relRMSE", rel)
    assert rel <= REL_RMSE_GATE


def test_debug_cornell_matches_oracle(gpu_ctx):
    def make():
        sc = S.cornell_sphere()
        sc.ubo["debug"] = 1
        return sc
    ds = G.device_scene(gpu_ctx, "cornell_debug", make)
    g, st = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL)
    w, sw = _render(gpu_ctx, ds, rt.MODE_WAVEFRONT)
    assert np.array_equal(g, w) and st["rays"] == sw["rays"]
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("cornell debug relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
    # the debug scatter differs from the PBR one
    pbr = ds.scene.ubo.copy()
    pbr["debug"] = 0
    p, _ = _render(gpu_ctx, ds, rt.MODE_MEGAKERNEL, ubo=pbr)
    assert G.rel_rmse(p[..., :3], g[..., :3]) > 0.1


def test_adjoint_refuses_debug(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "debug_lambert", lambda: D.plane("lambert"))
    sc = ds.scene
    grad = rt.new_accumulation(sc.width, sc.height)
    b = ds.buffers.bindings(sc.ubo)
    lib, h = gpu_ctx.lib, gpu_ctx.handle
    rc = lib.rt_trace_rays_adjoint(h, ds.tlas.handle, C.byref(b), C.c_void_p(grad.data_ptr()), sc.width, sc.height,
                                   None, None, None, None)
    assert rc == -6 and b"debug" in lib.rt_last_error(h)
