"""Transmission through a volume on gfx950 (evalIndirectCombinedBRDF's transmission
lobe with KHR_materials_volume attenuation, RayTracing.rchit:352-399 bookkeeping of
Ray.volume_dis): a glass slab seen from above.  Every pixel must be a whole number of
samples of the Beer-Lambert unit the oracle test derives from Snell's law, the image
must match the FP64 oracle and the wavefront must match the megakernel bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import rt
from tests import debug_kit as D
from tests import gpu_kit as G
from tests import volume_kit as V

pytestmark = pytest.mark.gpu


def test_glass_slab_matches_oracle_and_closed_form(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "glass_slab", V.slab)
    sc = ds.scene
    imgs = []
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        acc = rt.new_accumulation(sc.width, sc.height)
        rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode)
        gpu_ctx.sync()
        imgs.append(acc.cpu().numpy().astype(np.float64))
    assert np.array_equal(imgs[0], imgs[1])
    g = imgs[0]
    o, _ = B.OracleScene(sc).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("glass slab relRMSE", rel)
    assert rel <= 1e-3
    d = D.primary_dirs(sc)
    cos_t = np.sqrt(1 - (1 - d[..., 1] ** 2) / 1.5 ** 2)
    color = np.array([0.8, 0.5, 0.2]).astype(np.float32).astype(np.float64)
    c = np.array(V.BASE).astype(np.float32).astype(np.float64)
    unit = (0.01 * c ** 2)[None, None, :] * color[None, None, :] ** (1.0 / (2.0 * cos_t))[..., None] / 0.999 ** 2
    n = g[..., :3] / unit
    assert np.allclose(n, np.round(n), atol=2e-4)
