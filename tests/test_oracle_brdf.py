"""Oracle BRDF library (restating crates/examples/gltf_viewer/shaders/lib/PBR.glsl,
lib/Material.glsl, lib/PunctualLight.glsl, lib/RayTracingCommons.glsl) against closed
forms, normalisation integrals and Monte Carlo expectations of its own samplers.

The sampler tests tie sampling to evaluation: the mean sampled weight of a lobe must
equal the numerically integrated BRDF*cos of the same lobe, which a dropped term, a
wrong G1/G2 or a mis-rotated frame breaks."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import binding as B


def v(*x):
    return B.vec(*x)


def integrate_hemisphere(f, nt=400, nphi=800):
    """Integral of f(L) over the +z hemisphere (midpoint rule in cos(theta), phi)."""
    mu = (np.arange(nt) + 0.5) / nt
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    M, P = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1 - M * M)
    L = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3)
    vals = np.array([f(l) for l in L])
    return vals.reshape(nt, nphi, -1).sum((0, 1)) * (1.0 / nt) * (2 * np.pi / nphi)


def stratified(n):
    k = int(math.isqrt(n))
    i, j = np.meshgrid(np.arange(k), np.arange(k), indexing="ij")
    rng = np.random.default_rng(3)
    u = (i.reshape(-1) + rng.random(k * k)) / k
    w = (j.reshape(-1) + rng.random(k * k)) / k
    return np.stack([u, w], 1)


def test_ggx_d_is_normalised(oracle_lib):
    # int D(h) (n.h) dw_h = 1 over the hemisphere
    for a2 in (0.04, 0.2, 0.5, 1.0):
        # substitute mu = 1 - s^2 to resolve the peak at mu = 1
        s = (np.arange(20000) + 0.5) / 20000
        mu = 1 - s * s
        d = np.array([oracle_lib.orc_ggx_d(a2, m) for m in mu])
        integral = np.sum(d * mu * 2 * s) / 20000 * 2 * np.pi
        assert abs(integral - 1) < 1e-4, (a2, integral)


def test_smith_terms_closed_form(oracle_lib):
    for alpha in (0.1, 0.4, 0.9):
        for nl in (0.2, 0.6, 0.95):
            for nv in (0.3, 0.8):
                lam = lambda c: (-1 + math.sqrt(1 + alpha ** 2 * (1 - c * c) / (c * c))) / 2
                g2 = 1 / (1 + lam(nl) + lam(nv))
                assert abs(oracle_lib.orc_smith_g2_height_correlated(alpha, nl, nv) - g2) < 1e-6 * max(1, g2)
                g1 = 2 / (1 + math.sqrt(1 + alpha ** 2 * (1 - nl * nl) / (nl * nl)))
                assert abs(oracle_lib.orc_smith_g1_ggx(alpha, nl, alpha * alpha, nl * nl) - g1) < 1e-12
    # limits: a smooth surface does not mask, grazing directions are fully masked
    assert abs(oracle_lib.orc_smith_g1_ggx(1e-4, 0.5, 1e-8, 0.25) - 1) < 1e-6
    assert oracle_lib.orc_smith_g1_ggx(0.5, 1e-4, 0.25, 1e-8) < 1e-3


def test_frostbite_published_form(oracle_lib):
    # Lagarde & de Rousiers 2014, listing 1 (energy-normalised Disney diffuse)
    for r in (0.0, 0.3, 1.0):
        for ndl, ndv, ldh in ((1, 1, 1), (0.3, 0.7, 0.5), (0.05, 0.9, 0.2)):
            fd90 = 0.5 * r + 2 * ldh * ldh * r
            ef = 1.0 + (1 / 1.51 - 1.0) * r
            ref = (1 + (fd90 - 1) * (1 - ndl) ** 5) * (1 + (fd90 - 1) * (1 - ndv) ** 5) * ef
            assert abs(oracle_lib.orc_frostbite_diffuse(r, ndl, ndv, ldh) - ref) < 1e-12
    assert oracle_lib.orc_frostbite_diffuse(0.0, 1.0, 1.0, 1.0) == 1.0


def test_cosine_hemisphere_moments(oracle_lib):
    u = stratified(40000)
    out = np.zeros(3)
    zs = []
    for a, b in u:
        oracle_lib.orc_sample_hemisphere(B.dptr(v(a, b)), B.dptr(out))
        assert abs(np.linalg.norm(out) - 1) < 1e-12 and out[2] >= 0
        zs.append(out[2])
    zs = np.array(zs)
    # pdf cos/pi: E[z] = 2/3, E[z^2] = 1/2
    assert abs(zs.mean() - 2 / 3) < 2e-3 and abs((zs ** 2).mean() - 0.5) < 2e-3


def test_quaternion_frame(oracle_lib):
    rng = np.random.default_rng(1)
    q, out, back = np.zeros(4), np.zeros(3), np.zeros(3)
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        oracle_lib.orc_rotation_to_z(B.dptr(n), B.dptr(q))
        oracle_lib.orc_rotate_point(B.dptr(q), B.dptr(n), B.dptr(out))
        assert np.allclose(out, (0, 0, 1), atol=1e-12)
        w = rng.normal(size=3)
        oracle_lib.orc_rotate_point(B.dptr(q), B.dptr(w), B.dptr(out))
        qi = np.array([-q[0], -q[1], -q[2], q[3]])
        oracle_lib.orc_rotate_point(B.dptr(qi), B.dptr(out), B.dptr(back))
        assert np.allclose(back, w, atol=1e-12) and abs(np.linalg.norm(out) - np.linalg.norm(w)) < 1e-12
    # the special case near -z is the 180 degree turn about x
    n = v(0, 0, -1)
    oracle_lib.orc_rotation_to_z(B.dptr(n), B.dptr(q))
    assert np.array_equal(q, [1, 0, 0, 0])
    oracle_lib.orc_rotate_point(B.dptr(q), B.dptr(n), B.dptr(out))
    assert np.allclose(out, (0, 0, 1))


def _vndf_pdf(H, V, alpha):
    # Heitz 2018: D_V(H) = G1(V) max(0, V.H) D(H) / (V.N)
    a2 = alpha * alpha
    c = H[..., 2]
    D = a2 / (np.pi * ((a2 - 1) * c * c + 1) ** 2)
    g1 = 2 / (1 + np.sqrt(1 + a2 * (1 - V[2] ** 2) / V[2] ** 2))
    return g1 * np.maximum(0, H @ V) * D / V[2]


@pytest.mark.parametrize("alpha,theta", [(0.3, 0.4), (0.6, 1.1), (0.9, 0.2)])
def test_vndf_sampling_matches_visible_normal_density(oracle_lib, alpha, theta):
    V = np.array([math.sin(theta), 0.0, math.cos(theta)])
    u = stratified(90000)
    H = np.zeros((len(u), 3))
    out = np.zeros(3)
    for i, (a, b) in enumerate(u):
        oracle_lib.orc_sample_ggx_vndf(B.dptr(V), alpha, B.dptr(v(a, b)), B.dptr(out))
        H[i] = out
    moments_mc = np.array([H[:, 0].mean(), H[:, 2].mean(), (H[:, 2] ** 2).mean()])
    moments_num = integrate_hemisphere(lambda h: _vndf_pdf(h, V, alpha) * np.array([h[0], h[2], h[2] ** 2, 1.0]),
                                       nt=300, nphi=300)
    assert abs(moments_num[3] - 1) < 5e-3  # the density is normalised
    assert np.allclose(moments_mc, moments_num[:3], atol=5e-3), (moments_mc, moments_num)


def _spec_only_mat(roughness):
    return B.matbrdf(base_color=(0.9, 0.6, 0.3), metallic=1.0, roughness=roughness)


@pytest.mark.parametrize("roughness,theta", [(0.5, 0.3), (0.7, 0.9), (0.9, 0.5)])
def test_specular_sample_weight_is_unbiased(oracle_lib, roughness, theta):
    """E[weight] of the VNDF-sampled specular lobe equals int f_spec(V,L) cos dL."""
    m = _spec_only_mat(roughness)
    V = v(math.sin(theta), 0, math.cos(theta))
    N = v(0, 0, 1)
    out = np.zeros(3)

    def f(L):
        oracle_lib.orc_eval_combined_brdf(B.dptr(N), B.dptr(np.ascontiguousarray(L)), B.dptr(V), C.byref(m),
                                          B.dptr(out))
        return out.copy()
    ref = integrate_hemisphere(f, nt=300, nphi=300)
    u = stratified(90000)
    acc = np.zeros(3)
    d, w = np.zeros(3), np.zeros(3)
    for a, b in u:
        r = oracle_lib.orc_sample_indirect(B.dptr(v(a, b)), B.dptr(N), B.dptr(N), B.dptr(V), C.byref(m), 2,
                                           B.dptr(d), B.dptr(w))
        if r:
            acc += w
    mc = acc / len(u)
    assert np.allclose(mc, ref, rtol=6e-3), (mc, ref)


def test_diffuse_sample_weight_is_unbiased(oracle_lib):
    """With F = 0 (no specular colour) the diffuse sampled weight averages to the
    integral of the evaluated Frostbite lobe."""
    m = B.matbrdf(base_color=(0.7, 0.5, 0.2), metallic=0.0, roughness=0.6, specular_color=(0, 0, 0))
    V = v(0.5, 0.1, math.sqrt(1 - 0.26))
    N = v(0, 0, 1)
    out = np.zeros(3)

    def f(L):
        oracle_lib.orc_eval_combined_brdf(B.dptr(N), B.dptr(np.ascontiguousarray(L)), B.dptr(V), C.byref(m),
                                          B.dptr(out))
        return out.copy()
    ref = integrate_hemisphere(f, nt=200, nphi=200)
    u = stratified(40000)
    acc = np.zeros(3)
    d, w = np.zeros(3), np.zeros(3)
    for a, b in u:
        if oracle_lib.orc_sample_indirect(B.dptr(v(a, b)), B.dptr(N), B.dptr(N), B.dptr(V), C.byref(m), 1,
                                          B.dptr(d), B.dptr(w)):
            acc += w
    assert np.allclose(acc / len(u), ref, rtol=3e-3), (acc / len(u), ref)


def test_sampled_direction_is_rotated_back(oracle_lib):
    # a perfectly smooth metal reflects V about the shading normal
    m = B.matbrdf(base_color=(1, 1, 1), metallic=1.0, roughness=0.0)
    rng = np.random.default_rng(5)
    d, w = np.zeros(3), np.zeros(3)
    for _ in range(50):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        V = rng.normal(size=3)
        V /= np.linalg.norm(V)
        if V @ n < 0.05:
            V = V - 2 * (V @ n) * n + 0.1 * n
            V /= np.linalg.norm(V)
        assert oracle_lib.orc_sample_indirect(B.dptr(v(0.3, 0.7)), B.dptr(n), B.dptr(n), B.dptr(np.ascontiguousarray(V)),
                                              C.byref(m), 2, B.dptr(d), B.dptr(w))
        assert np.allclose(d, 2 * (V @ n) * n - V, atol=1e-9)
        # F0 = 1 for a white metal: weight F * G2/G1 = G2/G1 <= 1
        assert np.all(w <= 1 + 1e-12) and np.all(w > 0)


def test_fresnel_endpoints_and_mat_build(oracle_lib):
    m = B.matbrdf(base_color=(0.2, 0.4, 0.6), metallic=0.25, ior=1.5, specular_factor=0.5, specular_color=(1, 0.5, 0.25))
    f0 = np.minimum(((1.5 - 1) / 2.5) ** 2 * np.array([1, 0.5, 0.25]), 1) * 0.5
    assert np.allclose(m.F0[:], f0 * 0.75 + np.array([0.2, 0.4, 0.6]) * 0.25)
    assert np.allclose(m.F90[:], np.array([1, 0.5, 0.25]) * 0.75 + 0.25)
    assert np.allclose(m.c_diff[:], np.array([0.2, 0.4, 0.6]) * 0.75)
    # ior forced to 1 (no volume): dielectric F0 vanishes (RayTracing.rchit:320)
    m1 = B.matbrdf(base_color=(0.5, 0.5, 0.5), metallic=0.0, ior=1.0)
    assert np.allclose(m1.F0[:], 0)


def test_brdf_probability_is_a_distribution(oracle_lib):
    out = np.zeros(3)
    rng = np.random.default_rng(2)
    for _ in range(200):
        m = B.matbrdf(base_color=tuple(rng.random(3)), metallic=rng.random(), roughness=rng.random(),
                      transmission=rng.choice([0.0, rng.random()]), specular_factor=rng.random())
        V = rng.normal(size=3)
        V[2] = abs(V[2])
        V /= np.linalg.norm(V)
        oracle_lib.orc_brdf_probability(C.byref(m), B.dptr(V), B.dptr(v(0, 0, 1)), B.dptr(out))
        assert abs(out.sum() - 1) < 1e-12 and np.all(out >= 0)
        assert 0.001 - 1e-12 <= out[0] <= 0.9 + 1e-12
        if m.transmission == 0:
            assert out[2] == 0


def test_convert_metallic_solves_the_quadratic(oracle_lib):
    rng = np.random.default_rng(4)
    for _ in range(100):
        d, s = rng.random(3), rng.random(3) * 0.6 + 0.05
        mx = s.max()
        m = oracle_lib.orc_convert_metallic(B.dptr(d), B.dptr(s), mx)
        pd = math.sqrt(0.299 * d[0] ** 2 + 0.587 * d[1] ** 2 + 0.114 * d[2] ** 2)
        ps = math.sqrt(0.299 * s[0] ** 2 + 0.587 * s[1] ** 2 + 0.114 * s[2] ** 2)
        if ps < 0.04:
            assert m == 0
            continue
        a, b, c = 0.04, pd * (1 - mx) / 0.96 + ps - 0.08, 0.04 - ps
        if 0 < m < 1:
            assert abs(a * m * m + b * m + c) < 1e-12
    assert oracle_lib.orc_convert_metallic(B.dptr(v(1, 1, 1)), B.dptr(v(0.01, 0.01, 0.01)), 0.01) == 0


def test_light_attenuation_closed_form(oracle_lib):
    from rustracer_amd import layouts as L
    l = np.zeros((), L.LIGHT)
    l["color"] = (1, 1, 1, 0)
    l["kind"] = L.LIGHT_POINT
    l["intensity"] = 3.0
    lp = np.ascontiguousarray(l)
    # Yuksel: 2 / (d^2 + r^2 + d sqrt(d^2 + r^2)), r = 0.5
    assert abs(oracle_lib.orc_light_intensity_luminance(lp.ctypes.data, 0.0) - 3.0 * 8.0) < 1e-9
    for d in (0.5, 2.0, 100.0):
        ref = 3.0 * 2 / (d * d + 0.25 + d * math.sqrt(d * d + 0.25))
        assert abs(oracle_lib.orc_light_intensity_luminance(lp.ctypes.data, d) - ref) < 1e-9 * ref
    # far field tends to the inverse square law
    assert abs(oracle_lib.orc_light_intensity_luminance(lp.ctypes.data, 1e4) * 1e8 / 3.0 - 1) < 1e-6
    l["kind"] = L.LIGHT_DIRECTIONAL
    lp = np.ascontiguousarray(l)
    assert oracle_lib.orc_light_intensity_luminance(lp.ctypes.data, 7.0) == pytest.approx(3.0)


def test_offset_ray_integer_ulps(oracle_lib):
    out = np.zeros(3, np.float32)
    p = np.array([1.5, -2.25, 0.01], np.float32)
    n = np.array([0.5, 0.5, -0.7], np.float32)
    oracle_lib.orc_offset_ray(p.ctypes.data, n.ctypes.data, out.ctypes.data)
    bits = p.view(np.int32).copy()
    of = (n * 256).astype(np.int32)
    # |p| >= 1/32: move the float by int(n*256) ulps away from/towards zero with n
    assert out.view(np.int32)[0] == bits[0] + of[0]
    assert out.view(np.int32)[1] == bits[1] - of[1]
    # |p| < 1/32: a fixed-point step of n / 65536
    assert out[2] == np.float32(p[2] + np.float32(1 / 65536) * n[2])
    # the offset moves the point along n
    assert out[0] > p[0] and out[1] > p[1] and out[2] < p[2]


def test_texture_sampling_rules(oracle_lib):
    from rustracer_amd import layouts as L
    img = np.zeros((2, 4, 4), np.uint8)
    img[..., 0] = np.array([[0, 64, 128, 255], [10, 20, 30, 40]])
    img[..., 3] = 255
    out = np.zeros(4)

    def sample(fmt, wrap, filt, u, w):
        t = B.RtTexture(img.ctypes.data, 4, 2, fmt, wrap, wrap, filt, filt)
        oracle_lib.orc_texture_sample(C.byref(t), u, w, B.dptr(out))
        return out.copy()
    # texel centres reproduce the texel (linear filter, UNORM)
    assert sample(L.FMT_RGBA8_UNORM, L.WRAP_REPEAT, L.FILTER_LINEAR, 1.5 / 4, 0.5 / 2)[0] == pytest.approx(64 / 255)
    # halfway between two centres averages them
    assert sample(L.FMT_RGBA8_UNORM, L.WRAP_CLAMP_TO_EDGE, L.FILTER_LINEAR, 2 / 4, 0.5 / 2)[0] == pytest.approx(96 / 255)
    # repeat: u and u + 1 agree; clamp: outside is the edge texel
    a = sample(L.FMT_RGBA8_UNORM, L.WRAP_REPEAT, L.FILTER_LINEAR, 0.3, 0.6)
    assert np.allclose(a, sample(L.FMT_RGBA8_UNORM, L.WRAP_REPEAT, L.FILTER_LINEAR, 1.3, -0.4))
    assert sample(L.FMT_RGBA8_UNORM, L.WRAP_CLAMP_TO_EDGE, L.FILTER_NEAREST, -3.0, 0.25)[0] == 0.0
    # mirrored repeat: texel -1 is texel 0, texel 4 is texel 3
    assert sample(L.FMT_RGBA8_UNORM, L.WRAP_MIRRORED_REPEAT, L.FILTER_NEAREST, -0.1, 0.25)[0] == 0.0
    assert sample(L.FMT_RGBA8_UNORM, L.WRAP_MIRRORED_REPEAT, L.FILTER_NEAREST, 1.1, 0.25)[0] == 1.0
    # sRGB decode (IEC 61966-2-1) per texel, alpha linear
    s = sample(L.FMT_RGBA8_SRGB, L.WRAP_REPEAT, L.FILTER_NEAREST, 2.5 / 4, 0.25)
    assert s[0] == pytest.approx(((128 / 255 + 0.055) / 1.055) ** 2.4) and s[3] == 1.0
    assert sample(L.FMT_RGBA8_SRGB, L.WRAP_REPEAT, L.FILTER_NEAREST, 3.5 / 4, 0.25)[0] == pytest.approx(1.0)


def test_watertight_triangle_test(oracle_lib):
    v0, v1, v2 = v(0, 0, 0), v(1, 0, 0), v(0, 1, 0)
    out = np.zeros(3)
    # a ray through the interior reports t and the barycentrics of vertices 1 and 2
    assert oracle_lib.orc_intersect_triangle(B.dptr(v(0.2, 0.3, 5)), B.dptr(v(0, 0, -1)), B.dptr(v0), B.dptr(v1),
                                             B.dptr(v2), 0.0, 1e9, B.dptr(out))
    assert np.allclose(out, (5, 0.2, 0.3))
    # a ray through the shared edge of two triangles hits at least one of them
    v3 = v(1, 1, 0)
    for x in np.linspace(0.01, 0.99, 37):
        o = v(x, 1 - x, 2)
        d = v(0.0, 0.0, -1.0)
        h1 = oracle_lib.orc_intersect_triangle(B.dptr(o), B.dptr(d), B.dptr(v0), B.dptr(v1), B.dptr(v2), 0, 1e9, B.dptr(out))
        h2 = oracle_lib.orc_intersect_triangle(B.dptr(o), B.dptr(d), B.dptr(v1), B.dptr(v3), B.dptr(v2), 0, 1e9, B.dptr(out))
        assert h1 or h2
    # outside misses; t outside [tmin, tmax] misses
    assert not oracle_lib.orc_intersect_triangle(B.dptr(v(0.8, 0.8, 1)), B.dptr(v(0, 0, -1)), B.dptr(v0), B.dptr(v1),
                                                 B.dptr(v2), 0, 1e9, B.dptr(out))
    assert not oracle_lib.orc_intersect_triangle(B.dptr(v(0.2, 0.2, 1)), B.dptr(v(0, 0, -1)), B.dptr(v0), B.dptr(v1),
                                                 B.dptr(v2), 0, 0.5, B.dptr(out))


def test_bilinear_filter_reproduces_constant_neighbourhoods_exactly(oracle_lib):
    """Texture hardware filters with fixed-point weights, so four equal texels give
    that value exactly; a metallic texture of 255 must give metallic == 1.0 (the
    diffuse lobe then has exactly zero weight).  Summing four rounded weights does
    not guarantee this; nested lerps do."""
    from rustracer_amd import layouts as L
    img = np.full((8, 8, 4), 255, np.uint8)
    img[..., 1] = 77
    out = np.zeros(4)
    t = B.RtTexture(img.ctypes.data, 8, 8, L.FMT_RGBA8_UNORM, L.WRAP_REPEAT, L.WRAP_REPEAT, L.FILTER_LINEAR,
                    L.FILTER_LINEAR)
    rng = np.random.default_rng(5)
    for u, v in rng.random((2000, 2)) * 3.0:
        oracle_lib.orc_texture_sample(C.byref(t), u, v, B.dptr(out))
        assert out[0] == 1.0 and out[2] == 1.0 and out[1] == 77 / 255
