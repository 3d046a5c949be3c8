"""Host decode of the 64-byte quantised nodes (rt.decode_nodes) follows the device
decode fmaf(q, 2^e, origin): q * 2^e is exact in fp32, so one rounding of the sum."""
import numpy as np

from rustracer_amd import rt


def test_decode_matches_a_double_precision_reference():
    rng = np.random.default_rng(4)
    raw = np.zeros(64, rt.QNODE)
    raw["origin"] = rng.normal(scale=100.0, size=(64, 3)).astype(np.float32)
    raw["exp"] = rng.integers(-30, 10, size=(64, 3))
    raw["q"] = rng.integers(0, 256, size=(64, 24))
    raw["child"] = rng.integers(0, 1000, size=(64, 4))
    raw["count"] = rng.integers(0, 5, size=(64, 4))
    dec = rt.decode_nodes(raw)
    for a, (lo, hi) in enumerate((("lo_x", "hi_x"), ("lo_y", "hi_y"), ("lo_z", "hi_z"))):
        s = np.ldexp(1.0, raw["exp"][:, a].astype(np.int64))[:, None]
        o = raw["origin"][:, a].astype(np.float64)[:, None]
        ref_lo = (o + raw["q"][:, 8 * a:8 * a + 4].astype(np.float64) * s).astype(np.float32)
        ref_hi = (o + raw["q"][:, 8 * a + 4:8 * a + 8].astype(np.float64) * s).astype(np.float32)
        assert np.array_equal(dec[lo], ref_lo) and np.array_equal(dec[hi], ref_hi)
    assert np.array_equal(dec["child"], raw["child"]) and np.array_equal(dec["count"], raw["count"])
