"""Oracle RNG (oracle/oracle.cpp) against the committed known-answer vectors and an
independent numpy restatement (tests/golden/rng_numpy.py).  Restates
crates/examples/gltf_viewer/shaders/lib/Random.glsl."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from tests.golden import rng_numpy as R

HERE = os.path.dirname(os.path.abspath(__file__))
KAT = json.load(open(os.path.join(HERE, "golden", "rng_kat.json")))


def orc_pcg4d(lib, s):
    a = (C.c_uint32 * 4)(*s)
    b = (C.c_uint32 * 4)()
    lib.orc_pcg4d(a, b)
    return list(b)


def orc_rand(lib, s, n):
    a = (C.c_uint32 * 4)(*s)
    out = np.zeros(n, np.float32)
    w = lib.orc_rand_sequence(a, n, out.ctypes.data)
    return out, w


def test_pcg4d_known_answers(oracle_lib):
    for e in KAT["pcg4d"]:
        assert orc_pcg4d(oracle_lib, e["in"]) == e["out"]


def test_rand_known_answers(oracle_lib):
    for e in KAT["rand"]:
        vals, w = orc_rand(oracle_lib, e["state"], len(e["values_hex"]))
        assert [float(v).hex() for v in vals] == e["values_hex"]
        assert w == e["final_w"]


def test_tea_and_lcg_known_answers(oracle_lib):
    for e in KAT["tea16"]:
        assert oracle_lib.orc_tea16(*e["in"]) == e["out"]
    for e in KAT["lcg"]:
        out = np.zeros(len(e["values_hex"]), np.float32)
        final = oracle_lib.orc_lcg_sequence(e["seed"], len(out), out.ctypes.data)
        assert [float(v).hex() for v in out] == e["values_hex"] and final == e["final"]


def test_rand_matches_numpy_restatement_on_a_grid(oracle_lib):
    rng = np.random.default_rng(7)
    for _ in range(64):
        s = [int(v) for v in rng.integers(0, 2 ** 32, 4, dtype=np.uint64)]
        vals, w = orc_rand(oracle_lib, s, 5)
        ref, rw = R.rand_sequence(s, 5)
        assert np.array_equal(vals, ref) and w == rw


def test_rand_float_conversion_is_exact():
    # uintToFloat keeps the 23 high bits: the value is k / 2^23 exactly, in [0, 1)
    vals, _ = R.rand_sequence((5, 6, 7, 8), 200)
    k = vals.astype(np.float64) * 2 ** 23
    assert np.all(k == np.round(k)) and vals.min() >= 0 and vals.max() < 1


def test_pcg4d_is_a_bijection_sample():
    # pcg4d is invertible on uint32^4; distinct inputs must give distinct outputs
    x = np.arange(4096, dtype=np.uint64)
    out = np.stack(R.pcg4d(x, x * 3, x * 7, x * 11), 1)
    assert len(np.unique(out, axis=0)) == 4096


def test_stream_b_seed_differs_from_stream_a(oracle_lib):
    # stream B replaces clockARB(); it must not coincide with stream A (frame_count)
    for seed in range(4):
        for frame in range(4):
            assert oracle_lib.orc_stream_b_seed(seed, frame) == R.tea16(seed, frame)
            assert R.tea16(seed, frame) != frame


@pytest.mark.parametrize("mutate", ["drop_xor", "wrong_shift"])
def test_planted_mistake_is_caught(mutate):
    # a plausible restatement error changes the known answers
    def bad(x, y, z, w):
        x, y, z, w = (R._u(v) for v in (x, y, z, w))
        a, c = np.uint64(1664525), np.uint64(1013904223)
        x, y, z, w = ((v * a + c) & R.M32 for v in (x, y, z, w))
        x = (x + y * w) & R.M32
        y = (y + z * x) & R.M32
        z = (z + x * y) & R.M32
        w = (w + y * z) & R.M32
        sh = np.uint64(15 if mutate == "wrong_shift" else 16)
        if mutate != "drop_xor":
            x, y, z, w = (v ^ (v >> sh) for v in (x, y, z, w))
        x = (x + y * w) & R.M32
        return int(x)
    e = KAT["pcg4d"][1]
    assert bad(*e["in"]) != e["out"][0]
