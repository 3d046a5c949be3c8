"""The BLEND any-hit on gfx950 (RayTracing.rahit:98-102) where the invocation order
cannot matter: one BLEND layer over an opaque emitter.  Every pixel must be a whole
number of samples of the layer's and the emitter's emission, the image must equal the
oracle's (the same stream-B draw decides each sample) and the wavefront must match
the megakernel bit for bit."""
import numpy as np
import pytest

from oracle import binding as B
from rustracer_amd import rt
from tests import gpu_kit as G
from tests import volume_kit as VK

pytestmark = pytest.mark.gpu


def test_single_blend_layer_matches_oracle(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "blend_layer", lambda: VK.blend_layer(alpha=0.3))
    sc = ds.scene
    imgs = []
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        acc = rt.new_accumulation(sc.width, sc.height)
        st = rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True)
        gpu_ctx.sync()
        imgs.append(acc.cpu().numpy().astype(np.float64))
        assert st["anyhit_calls"] == sc.width * sc.height * 16
    assert np.array_equal(imgs[0], imgs[1])
    o, _ = B.OracleScene(sc).render()
    assert G.rel_rmse(imgs[0][..., :3], o[..., :3]) <= 1e-6
    e1, e2 = np.float32(0.4) * 5.0, np.float32(0.1) * 5.0
    n1 = (imgs[0][..., 0] - 16 * e2) / (e1 - e2)
    assert np.allclose(n1, np.round(n1), atol=1e-5)


@pytest.mark.parametrize("alphas", [(0.3, 0.6)])
def test_two_blend_layers_follow_the_order_free_probabilities(gpu_ctx, alphas):
    """Two stacked BLEND layers on gfx950: the draw order inside the traversal may differ
    from the oracle's, so the image is checked against the order-free probabilities
    a1, (1-a1) a2, (1-a1)(1-a2) of ending on the near layer, the far layer or the back,
    in both execution modes, and every pixel must split into whole sample counts."""
    spp = 16
    ds = G.device_scene(gpu_ctx, "blend_stack", lambda: VK.blend_stack(alphas=alphas, spp=spp))
    sc = ds.scene
    p = VK.stack_probabilities(alphas)
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        acc = rt.new_accumulation(sc.width, sc.height)
        st = rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True)
        gpu_ctx.sync()
        n = VK.stack_counts(acc.cpu().numpy().astype(np.float64)[..., :3])
        assert np.allclose(n, np.round(n), atol=1e-5), mode
        n = np.round(n)
        assert np.all(n.sum(-1) == spp), mode
        N = sc.width * sc.height * spp
        frac = n.reshape(-1, 3).sum(0) / N
        sigma = np.sqrt(p * (1 - p) / N)
        assert np.all(np.abs(frac - p) < 5 * sigma), (mode, frac, p, sigma)
        assert N <= st["anyhit_calls"] <= 2 * N, mode
