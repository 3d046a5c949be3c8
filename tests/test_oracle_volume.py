"""The oracle's transmission lobe (evalIndirectCombinedBRDF, lib/PBR.glsl:700-727, and
its volume attenuation) pinned by physics that does not restate its code: Snell's
law for the refracted direction entering (eta = 1 / ior) and leaving (eta = ior) a
volume, total internal reflection past the critical angle, Beer-Lambert transmittance
attenuation_color^(distance / attenuation_distance) on the way out, and a straight
continuation without a volume."""
import numpy as np
import pytest

from oracle import binding as B
from tests import volume_kit as VK

N = np.array([0.0, 0.0, 1.0])
BASE = VK.BASE
TRANSMISSION = 3


def _view(theta):
    return np.array([np.sin(theta), 0.0, np.cos(theta)])


@pytest.mark.parametrize("theta_deg", [0.0, 10.0, 35.0, 60.0, 85.0])
def test_entering_a_volume_obeys_snell(theta_deg):
    th = np.radians(theta_deg)
    m = B.matbrdf(base_color=BASE, ior=1.5, transmission=1.0, front_face=True, volume=True,
                  attenuation_color=(0.5, 0.5, 0.5), attenuation_distance=1.0)
    ok, d, w, vd = B.sample_indirect_vd((0.3, 0.7), N, N, _view(th), m, TRANSMISSION, 2.0)
    assert ok
    assert abs(np.linalg.norm(d) - 1) < 1e-12 and d[2] < 0 and abs(d[1]) < 1e-15
    sin_t = np.hypot(d[0], d[1])
    assert abs(sin_t - np.sin(th) / 1.5) < 1e-12
    assert np.allclose(w, BASE, rtol=1e-15)  # c_diff x transmission, no attenuation on entry
    assert vd == 2.0  # the entering side leaves the distance to the integrator


@pytest.mark.parametrize("theta_deg", [0.0, 20.0, 40.0])
def test_leaving_a_volume_attenuates_by_beer_lambert(theta_deg):
    th = np.radians(theta_deg)
    color, dist, travelled = np.array([0.8, 0.5, 0.2]), 2.0, 3.0
    m = B.matbrdf(base_color=BASE, ior=1.5, transmission=0.5, front_face=False, volume=True,
                  attenuation_color=tuple(color), attenuation_distance=dist)
    ok, d, w, vd = B.sample_indirect_vd((0.3, 0.7), N, N, _view(th), m, TRANSMISSION, travelled)
    assert ok
    assert abs(np.hypot(d[0], d[1]) - 1.5 * np.sin(th)) < 1e-12 and d[2] < 0
    expect = np.array(BASE) * 0.5 * color ** (travelled / dist)
    assert np.allclose(w, expect, rtol=1e-12, atol=0)
    assert vd == -1.0  # the ray has left the volume


def test_total_internal_reflection_past_the_critical_angle():
    crit = np.arcsin(1 / 1.5)
    m = B.matbrdf(base_color=BASE, ior=1.5, transmission=1.0, front_face=False, volume=True)
    ok, _, w, _ = B.sample_indirect_vd((0.3, 0.7), N, N, _view(crit + 0.01), m, TRANSMISSION, 1.0)
    assert not ok and np.all(w == 0)
    ok2, _, _, _ = B.sample_indirect_vd((0.3, 0.7), N, N, _view(crit - 0.01), m, TRANSMISSION, 1.0)
    assert ok2


def test_thin_transmission_continues_straight():
    th = np.radians(30)
    m = B.matbrdf(base_color=BASE, ior=1.5, transmission=0.25, front_face=True, volume=False)
    ok, d, w, _ = B.sample_indirect_vd((0.3, 0.7), N, N, _view(th), m, TRANSMISSION, -1.0)
    assert ok and np.allclose(d, -_view(th), atol=1e-15)
    assert np.allclose(w, np.array(BASE) * 0.25, rtol=1e-15)


def test_slab_transmittance_follows_the_path_length_inside():
    """Each sample either dies on the 0.001 reflection lobe (weight 0) or crosses the
    slab: c_diff^2 x color^(T / (D cos theta_t)) x 0.01 / p_t^2 with p_t = 0.999 at both
    faces.  So every pixel is a whole number (0..spp) of that unit, where the
    distance inside follows from the pixel's incidence by Snell's law; the integrator's
    volume_dis bookkeeping (set on entry, grown by the hit distance, read on exit)
    is what makes the exponent."""
    from tests import debug_kit as D
    sc = VK.slab()
    o, st = B.OracleScene(sc).render()
    d = D.primary_dirs(sc)
    cos_i = np.abs(d[..., 1])
    sin_t = np.sqrt(1 - cos_i ** 2) / 1.5
    cos_t = np.sqrt(1 - sin_t ** 2)
    color = np.array([0.8, 0.5, 0.2]).astype(np.float32).astype(np.float64)
    c = np.array(BASE).astype(np.float32).astype(np.float64)
    p_t = 1 - 0.001
    unit = (0.01 * c ** 2)[None, None, :] * color[None, None, :] ** (1.0 / (2.0 * cos_t))[..., None] / p_t ** 2
    n = o[..., :3] / unit
    assert np.allclose(n, np.round(n), atol=2e-4), np.abs(n - np.round(n)).max()
    assert np.all(np.round(n) >= 0) and np.all(np.round(n) <= 16)
    assert np.all(np.round(n[..., 0]) == np.round(n[..., 2]))
    frac = float(np.round(n[..., 0]).sum()) / (48 * 48 * 16)
    assert abs(frac - p_t ** 2) < 0.002, frac
