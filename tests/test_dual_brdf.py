"""The dual-number BRDF terms of the adjoint (rustracer_amd/csrc/brdf_dual.h)
against the oracle's BRDF (oracle/oracle.cpp, which restates
crates/examples/gltf_viewer/shaders/lib/PBR.glsl): values equal the oracle's, and
derivatives equal central finite differences of the oracle's functions."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import binding as B

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def shim():
    d = tempfile.mkdtemp()
    so = os.path.join(d, "libdual.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                           os.path.join(HERE, "dual_shim.cpp"), "-o", so])
    lib = C.CDLL(so)
    dp = C.POINTER(C.c_double)
    lib.dual_eval.argtypes = [dp] * 6
    lib.dual_weight.argtypes = [C.c_int, dp, dp, dp, dp, C.c_double, C.c_int, dp, dp]
    return lib


def _p(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _mat_vec(base, metallic, roughness, spec=1.0):
    return np.array([*base, metallic, roughness, 1.0, 0.0, spec, 1.0, 1.0, 1.0])


def _unit(rng, n=3):
    v = rng.normal(size=n)
    return v / np.linalg.norm(v)


def _hemi(rng, N):
    while True:
        v = _unit(rng)
        if v @ N > 0.05:
            return v


def _dual_eval(shim, N, L, V, mv):
    val, der = np.zeros(3), np.zeros(15)
    args = [_p(x) for x in (N, L, V, mv)]
    shim.dual_eval(*[a[1] for a in args], _p(val)[1], _p(der)[1])
    return val, der.reshape(3, 5)


def _oracle_eval(N, L, V, mv):
    m = B.matbrdf(base_color=tuple(mv[:3]), metallic=mv[3], roughness=mv[4], specular_factor=mv[7])
    out = np.zeros(3)
    B.lib().orc_eval_combined_brdf(_p(N)[1], _p(L)[1], _p(V)[1], C.byref(m), _p(out)[1])
    return out


def _fd(fun, mv, k, h=1e-6):
    a, b = mv.copy(), mv.copy()
    idx = [0, 1, 2, 4, 3][k]  # parameter order base r g b, roughness, metallic
    a[idx] += h
    b[idx] -= h
    return (fun(a) - fun(b)) / (2 * h)


def test_eval_brdf_value_and_derivatives_match_the_oracle(shim):
    rng = np.random.default_rng(7)
    for _ in range(200):
        N = _unit(rng)
        L, V = _hemi(rng, N), _hemi(rng, N)
        mv = _mat_vec(rng.uniform(0.05, 0.95, 3), rng.uniform(0.05, 0.95), rng.uniform(0.2, 0.95))
        val, der = _dual_eval(shim, N, L, V, mv)
        ref = _oracle_eval(N, L, V, mv)
        assert np.allclose(val, ref, rtol=1e-12, atol=1e-14)
        for k in range(5):
            fd = _fd(lambda m: _oracle_eval(N, L, V, m), mv, k)
            assert np.allclose(der[:, k], fd, rtol=1e-5, atol=1e-8), (k, der[:, k], fd)


def _oracle_sample(u, N, V, mv, lobe):
    m = B.matbrdf(base_color=tuple(mv[:3]), metallic=mv[3], roughness=mv[4], specular_factor=mv[7])
    d, w = np.zeros(3), np.zeros(3)
    ok = B.lib().orc_sample_indirect(_p(u)[1], _p(N)[1], _p(N)[1], _p(V)[1], C.byref(m), lobe, _p(d)[1], _p(w)[1])
    return ok, d, w


def _dual_weight(shim, lobe, N, L, V, mv, p_lobe):
    val, der = np.zeros(3), np.zeros(15)
    args = [_p(x) for x in (N, L, V, mv)]
    shim.dual_weight(lobe, *[a[1] for a in args], p_lobe, 0, _p(val)[1], _p(der)[1])
    return val, der.reshape(3, 5)


@pytest.mark.parametrize("lobe", [1, 2])
def test_sample_weight_value_matches_the_oracle(shim, lobe):
    """The value of the detached weight is the shader's sampled weight (divided by
    the lobe probability) for the direction the oracle samples."""
    rng = np.random.default_rng(11 + lobe)
    n = 0
    while n < 200:
        N = _unit(rng)
        V = _hemi(rng, N)
        mv = _mat_vec(rng.uniform(0.05, 0.95, 3), rng.uniform(0.05, 0.95), rng.uniform(0.2, 0.95))
        ok, L, w = _oracle_sample(rng.random(2), N, V, mv, lobe)
        if not ok or L @ N <= 1e-3:
            continue
        val, _ = _dual_weight(shim, lobe, N, L, V, mv, 0.25)
        assert np.allclose(val * 0.25, w, rtol=1e-9, atol=1e-12), (val * 0.25, w)
        n += 1


def test_diffuse_weight_derivatives_match_finite_differences(shim):
    """Cosine sampling does not depend on the parameters, so the detached diffuse
    weight's derivative is the oracle weight's derivative at fixed random numbers."""
    rng = np.random.default_rng(21)
    n = 0
    while n < 100:
        N = _unit(rng)
        V = _hemi(rng, N)
        u = rng.random(2)
        mv = _mat_vec(rng.uniform(0.05, 0.95, 3), rng.uniform(0.05, 0.95), rng.uniform(0.2, 0.95))
        ok, L, _ = _oracle_sample(u, N, V, mv, 1)
        if not ok:
            continue
        _, der = _dual_weight(shim, 1, N, L, V, mv, 1.0)
        for k in range(5):
            fd = _fd(lambda m: _oracle_sample(u, N, V, m, 1)[2], mv, k)
            assert np.allclose(der[:, k], fd, rtol=1e-5, atol=1e-8), (k, der[:, k], fd)
        n += 1


def test_specular_weight_derivatives_match_finite_differences(shim):
    """Specular: F D G2 / (G1 D) with the density's G1 D detached.  Finite
    differences of the oracle's GGX D, height-correlated G2 and Schlick Fresnel
    (through its BRDF evaluation with the diffuse part removed) at a fixed
    direction."""
    rng = np.random.default_rng(31)
    for _ in range(100):
        N = _unit(rng)
        V, L = _hemi(rng, N), _hemi(rng, N)
        mv = _mat_vec(rng.uniform(0.05, 0.95, 3), rng.uniform(0.05, 0.95), rng.uniform(0.2, 0.95))
        val, der = _dual_weight(shim, 2, N, L, V, mv, 1.0)

        def spec_eval(m):
            # evalCombinedBRDF minus its Frostbite diffuse term: F D G2 / (4 n.v)
            m_nodiff = m.copy()
            full = _oracle_eval(N, L, V, m)
            base = np.array(m[:3])
            H = (L + V) / np.linalg.norm(L + V)
            ndl, ndv, ldh = np.clip(N @ L, 1e-5, 1), np.clip(N @ V, 1e-5, 1), np.clip(L @ H, 0, 1)
            r = m[4]
            fd90m1 = 0.5 * r + 2 * ldh * ldh * r - 1
            fb = (1 + fd90m1 * (1 - ndl) ** 5) * (1 + fd90m1 * (1 - ndv) ** 5) * ((1 - r) + r / 1.51)
            diff = base * (1 - m[3]) * fb * ndl / 3.141592653589
            del m_nodiff
            return (full - diff) * 4 * ndv
        # value: F D G2 / (G1 D) = (4 n.v spec_eval) / (G1(v) D)|p0
        a2 = (mv[4] ** 2) ** 2
        ndv = np.clip(N @ V, 1e-5, 1)
        g1v = 2 / (np.sqrt((a2 * (1 - ndv * ndv) + ndv * ndv) / (ndv * ndv)) + 1)
        H = (L + V) / np.linalg.norm(L + V)
        ndh = np.clip(N @ H, 0, 1)
        D0 = a2 / (3.141592653589 * ((a2 - 1) * ndh * ndh + 1) ** 2)
        assert np.allclose(val, spec_eval(mv) / (g1v * D0), rtol=1e-8)
        for k in range(5):
            fd = _fd(lambda m: spec_eval(m) / (g1v * D0), mv, k)
            assert np.allclose(der[:, k], fd, rtol=1e-5, atol=1e-8), (k, der[:, k], fd)
