"""Pins of three oracle functions that the rendered scenes of the other tests leave
open: the RIS light-sample weight with several lights (RayTracing.rchit:62-122,
338-350), the lobe probabilities of getBrdfProbability with a specular factor and
transmission (PBR.glsl:324-358), and the directional-light term of the miss shader
(RayTracing.rmiss:14-45).  Each check is a closed form or an expectation that holds
whatever the arithmetic inside the function, so a dropped factor, a wrong count or a
flipped sign fails it."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import binding as B
from tests import scene_kit as K

LUM = np.array([0.2126, 0.7152, 0.0722])


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


# ----------------------------------------------------------------- RIS weight
def _centre_mean(plights, frames=16, spp=1024, base=(0.5, 0.6, 0.7)):
    """Mean and standard error of the centre pixel's NEE estimate (depth 1) over
    `frames` independent frames of `spp` samples."""
    means = []
    for f in range(frames):
        sc = K.lit_plane(plights=plights, base=base, spp=spp, bounces=1, frame_count=f)
        cx, cy = sc.width // 2, sc.height // 2
        acc, st = B.OracleScene(sc).render(rect=(cx, cy, cx + 1, cy + 1))
        assert st["shadow_rays"] > 0
        means.append(acc[cy, cx, :3] / spp)
    m = np.array(means)
    return m.mean(0), m.std(0, ddof=1) / math.sqrt(frames)


def _per_light_terms(plights, base=(0.5, 0.6, 0.7)):
    """evalCombinedBRDF(N, L_l, V) * I_l * colour_l for the centre pixel, whose hit
    point is the origin and whose V is the surface normal +y."""
    m = B.matbrdf(base_color=tuple(f32(base)), metallic=0.0, roughness=1.0)
    N, V = B.vec(0, 1, 0), B.vec(0, 1, 0)
    out = np.zeros(3)
    terms = []
    for p, i, c in plights:
        Ld = np.array(p, np.float64)
        Ld /= np.linalg.norm(Ld)
        B.lib().orc_eval_combined_brdf(B.dptr(N), B.dptr(B.vec(*Ld)), B.dptr(V), C.byref(m), B.dptr(out))
        terms.append(out.copy() * f32(i) * f32(c))
    return np.array(terms)


THREE = [((1.5, 2.0, 0.5), 12.0, (1.0, 0.4, 0.2)),
         ((-2.0, 3.0, -1.0), 30.0, (0.3, 1.0, 0.5)),
         ((0.5, 1.2, -2.5), 8.0, (0.6, 0.6, 1.0))]


@pytest.mark.parametrize("plights,share", [
    (THREE, 1.0),                                        # three candidates over three lights
    (THREE[:2], 2.0 / 3.0),                              # two candidates, still divided by 3
    (THREE + [((2.5, 2.5, 2.5), 20.0, (1.0, 1.0, 1.0))], 1.0),   # three candidates over four lights
    ([((1.5, 2.0, 0.5), 0.05, (1.0, 1.0, 1.0))] + THREE[1:], 2.0 / 3.0),  # light 0 too dim: its round is skipped
])
def test_ris_estimate_converges_to_the_sum_over_lights(plights, share):
    """RIS over uniform candidates with target luminance(I * colour * attenuation):
    E[f(y) * (sum_i w_i / 3) / p(y)] = (M / 3) * sum_l f(l), M = the number of
    candidate rounds that draw (min(N, 3), less the rounds whose loop-index light is
    too dim, rchit:84-86).  f(l) = BRDF * I_l * colour_l, unattenuated (rchit:347-348).
    The per-light BRDF values come from the oracle's evalCombinedBRDF, pinned by the
    BRDF tests."""
    mean, se = _centre_mean(plights)
    expected = share * _per_light_terms(plights).sum(0)
    tol = 5 * se + 1e-9
    assert np.all(np.abs(mean - expected) <= tol), (mean, expected, se)
    assert np.all(np.abs(mean / expected - 1) < 0.02), (mean, expected)


# --------------------------------------------------------- lobe probabilities
def _probability(F0_lum_rgb, cdiff, sf, tr, cos_nv):
    """getBrdfProbability through the oracle with F0 and c_diff set directly
    (specular-glossiness workflow: F0 = the specular colour, c_diff = base * (1 - max))."""
    m = B.matbrdf(base_color=cdiff, metallic=0.0, roughness=0.5, transmission=tr, specular_factor=sf,
                  sg_specular=F0_lum_rgb)
    N = B.vec(0, 0, 1)
    V = B.vec(math.sqrt(1 - cos_nv * cos_nv), 0, cos_nv)
    out = np.zeros(3)
    B.lib().orc_brdf_probability(C.byref(m), B.dptr(V), B.dptr(N), B.dptr(out))
    F0 = np.array(m.F0[:])
    c_diff = np.array(m.c_diff[:])
    F90 = np.array(m.F90[:])
    return out, F0, c_diff, F90


@pytest.mark.parametrize("sf,tr", [(0.5, 0.3), (0.25, 0.0), (1.0, 0.6), (0.8, 1.0)])
def test_brdf_probability_head_on(sf, tr):
    """At N.V = 1 Schlick's term is F0 itself, so with s = lum(F0), c = lum(c_diff):
      spec = s * sf,  penetration = c * (1 - sf * s),
      p = clamp(spec / (spec + penetration), 0.001, 0.9),
      (p, (1 - p)(1 - tr), (1 - p) tr)."""
    out, F0, c_diff, _ = _probability((0.5, 0.45, 0.6), (0.7, 0.5, 0.3), sf, tr, 1.0)
    s, c = LUM @ F0, LUM @ c_diff
    spec = s * sf
    pen = c * (1 - sf * s)
    p = min(max(spec / (spec + pen), 0.001), 0.9)
    assert 0.001 < spec / (spec + pen) < 0.9
    exp = np.array([p, (1 - p) * (1 - tr), (1 - p) * tr])
    assert np.allclose(out, exp, rtol=0, atol=1e-12), (out, exp)


def test_brdf_probability_grazing():
    """At N.V = 0 Schlick's term is f90 = min(1, lum(F90)) (shadowedF90, PBR.glsl:305-320)."""
    sf, tr = 0.4, 0.25
    out, F0, c_diff, F90 = _probability((0.2, 0.25, 0.3), (0.6, 0.6, 0.6), sf, tr, 0.0)
    fres = min(1.0, LUM @ F90)
    c = LUM @ c_diff
    spec = fres * sf
    pen = c * (1 - sf * fres)
    p = min(max(spec / (spec + pen), 0.001), 0.9)
    exp = np.array([p, (1 - p) * (1 - tr), (1 - p) * tr])
    assert np.allclose(out, exp, rtol=0, atol=1e-12), (out, exp)


# --------------------------------------------------- directional-light miss
DLIGHTS = [((0.3, -1.0, 0.2), 3.0, (1.0, 0.8, 0.6)),    # faces every reflected ray
           ((1.0, 0.05, 0.0), 2.0, (0.2, 0.5, 1.0)),    # faces the reflected rays with d.x < -0.05 d.y
           ((0.0, 1.0, 0.0), 5.0, (1.0, 1.0, 1.0))]     # faces none


def _miss_term(d, lights=DLIGHTS):
    acc = np.zeros(d.shape[:-1] + (3,))
    for dirn, i, c in lights:
        n = np.array(dirn, np.float64)
        n /= np.linalg.norm(n)
        cos = d @ n
        acc += np.where(cos < 0, -cos, 0.0)[..., None] * f32(c) * f32(i)
    return acc + 0.01


def test_mirror_reflects_camera_rays_into_directional_lights():
    """A mirror (metallic 1, roughness 0, base 1: F = 1, G2/G1 = 1, so the sampled
    weight is exactly 1) under no point lights: each sample adds the miss emittance of
    the reflected camera ray, sum_l max(0, -cos_l) * colour_l * I_l + 0.01
    (RayTracing.rmiss:18-26,36)."""
    spp = 2
    sc = K.lit_plane(dlights=DLIGHTS, base=(1.0, 1.0, 1.0), metallic=1.0, roughness=0.0, spp=spp, bounces=2,
                     width=41, height=31)
    acc, st = B.OracleScene(sc).render()
    d = K.primary_dirs(sc)
    r = d * np.array([1.0, -1.0, 1.0])          # reflected about +y
    exp = spp * _miss_term(r)
    assert st["shadow_rays"] == 0
    assert np.allclose(acc[..., :3], exp, rtol=2e-6, atol=0), np.abs(acc[..., :3] / exp - 1).max()
    # the sign gate of the second light is crossed inside the image
    n1 = np.array(DLIGHTS[1][0]) / np.linalg.norm(DLIGHTS[1][0])
    c1 = r @ n1
    assert (c1 < 0).any() and (c1 > 0).any()


def test_primary_miss_adds_no_directional_light():
    """Ray.t == 0 on a camera ray: the directional lights are skipped and only the
    constant 0.01 remains (rmiss:18)."""
    sc = K.lit_plane(dlights=[((0.0, -1.0, 0.0), 50.0, (1.0, 1.0, 1.0))], look_away=True, spp=3, bounces=2)
    acc, _ = B.OracleScene(sc).render()
    assert np.all(acc[..., :3] == 3 * 0.01)
