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


@pytest.mark.parametrize("alphas", [(0.3, 0.6), (0.7, 0.25)])
def test_two_blend_layers_follow_the_order_free_probabilities(alphas):
    """Two stacked BLEND layers.  The order of the any-hit draws is left to the
    traversal, but each layer draws once at most, so the share of samples ending on
    the near layer, the far layer and the back is a1, (1-a1) a2, (1-a1)(1-a2) in any
    order.  Each surface emits into a colour channel of its own, so every pixel splits
    into whole sample counts that add up to the sample count."""
    spp = 16
    sc = VK.blend_stack(alphas=alphas, spp=spp)
    o, st = B.OracleScene(sc).render()
    n = VK.stack_counts(o[..., :3])
    assert np.allclose(n, np.round(n), atol=1e-5)
    n = np.round(n)
    assert np.all(n.sum(-1) == spp)
    N = n[..., 0].size * spp
    frac = n.reshape(-1, 3).sum(0) / N
    p = VK.stack_probabilities(alphas)
    sigma = np.sqrt(p * (1 - p) / N)
    assert np.all(np.abs(frac - p) < 5 * sigma), (frac, p, sigma)
    # one draw per camera ray at least, two at most
    assert N <= st["anyhit_calls"] <= 2 * N
