"""The oracle's BLEND any-hit (RayTracing.rahit:98-102: the candidate is ignored when
opacity <= u with u = rand(rngState)) pinned where the invocation order cannot matter:
one BLEND layer over an opaque emitter.  Each sample sees exactly one any-hit draw and
returns the layer's or the emitter's emission, so every pixel is a whole number of
samples of each, and the layer's share converges to its opacity."""
import numpy as np
import pytest

from oracle import binding as B
from tests import volume_kit as VK


@pytest.mark.parametrize("alpha", [0.3, 0.75])
def test_single_blend_layer_is_kept_with_probability_opacity(alpha):
    sc = VK.blend_layer(alpha=alpha)
    o, st = B.OracleScene(sc).render()
    x = 5.0  # exposure
    e1, e2 = np.float32(0.4) * x, np.float32(0.1) * x
    v = o[..., 0]
    n1 = (v - 16 * e2) / (e1 - e2)  # samples that kept the layer
    assert np.allclose(n1, np.round(n1), atol=1e-6)
    assert np.all(np.round(n1) >= 0) and np.all(np.round(n1) <= 16)
    N = v.size * 16
    frac = float(np.round(n1).sum()) / N
    sigma = (alpha * (1 - alpha) / N) ** 0.5
    assert abs(frac - alpha) < 5 * sigma, (frac, alpha, sigma)
    assert st["anyhit_calls"] == N  # one draw per camera ray
