"""The oracle's debug == 1 shading (RayTracing.rchit:437-475), pinned by closed
forms: a Lambertian, fuzzy-metal or dielectric plane seen from above under the
constant 0.01 miss returns exactly samples x colour x 0.01 per pixel (the scatter
weight is the colour, every scattered ray leaves upward or downward and misses),
where the PBR shading (debug 0) gives other values; and a dielectric plane under
an emitter reflects with Schlick's probability (lib/PBR.glsl:48-53) at each pixel's
incidence, checked in expectation over 65,536 draws."""
import numpy as np
import pytest

from oracle import binding as B
from tests import debug_kit as D


@pytest.mark.parametrize("kind", ["lambert", "metal", "glass"])
def test_plane_under_constant_miss(kind):
    sc = D.plane(kind)
    o, st = B.OracleScene(sc).render()
    c = np.array([0.9, 0.8, 0.7]).astype(np.float32).astype(np.float64)
    expect = 8 * c * 0.01
    assert np.allclose(o[..., :3], expect, rtol=1e-12, atol=0), (kind, o[0, 0], expect)
    assert st["extension_rays"] == 32 * 32 * 8
    pbr = sc.ubo.copy()
    pbr["debug"] = 0
    o2, _ = B.OracleScene(sc).render(ubo=pbr)
    assert not np.allclose(o2[..., :3], expect, rtol=1e-6)


def test_dielectric_reflects_with_schlick_probability():
    sc = D.plane("glass", width=64, height=64, spp=16, emitter=True, ior=1.5)
    o, _ = B.OracleScene(sc).render()
    c = np.array([0.9, 0.8, 0.7]).astype(np.float32).astype(np.float64)
    d = D.primary_dirs(sc)
    cos = np.abs(d[..., 1])
    ratio = 1 / 1.5
    r0 = ((1 - ratio) / (1 + ratio)) ** 2
    R = r0 + (1 - r0) * (1 - cos) ** 5
    spp = 16
    # per pixel: n_refl reflections (radiance 1) and spp - n_refl refractions (0.01)
    n_refl = (o[..., 0] / c[0] - spp * 0.01) / (1 - 0.01)
    assert np.allclose(n_refl, np.round(n_refl), atol=1e-6)  # whole draws
    assert np.allclose(o[..., 1] / c[1], o[..., 0] / c[0], rtol=1e-9)
    expect, var = float((spp * R).sum()), float((spp * R * (1 - R)).sum())
    got = float(n_refl.sum())
    print("reflections", got, "expected", expect, "sigma", var ** 0.5)
    assert abs(got - expect) < 5 * var ** 0.5
