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
