"""Independent numpy (uint32) restatement of the reference RNG, used to pin the
oracle's C++ version bit for bit.

Follows crates/examples/gltf_viewer/shaders/lib/Random.glsl: pcg4d (:81-98, from
Jarzynski & Olano, "Hash Functions for GPU Rendering", JCGT 9(3) 2020), rand (:102-105,
counter pre-increment, 23 high bits as mantissa), InitRandomSeed / TEA with 16
rounds (:12-25), RandomInt / RandomFloat (LCG with Numerical Recipes constants, :27-42).
"""
import numpy as np

M32 = np.uint64(0xFFFFFFFF)


def _u(x):
    return np.asarray(x, np.uint64) & M32


def pcg4d(x, y, z, w):
    x, y, z, w = (_u(v) for v in (x, y, z, w))
    a, c = np.uint64(1664525), np.uint64(1013904223)
    x, y, z, w = ((v * a + c) & M32 for v in (x, y, z, w))
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    x, y, z, w = (v ^ (v >> np.uint64(16)) for v in (x, y, z, w))
    x = (x + y * w) & M32
    y = (y + z * x) & M32
    z = (z + x * y) & M32
    w = (w + y * z) & M32
    return x, y, z, w


def rand_sequence(state, n):
    """n successive rand() values from state (x, y, z, w); returns (floats, final w)."""
    x, y, z, w = (int(v) for v in state)
    out = np.zeros(n, np.float32)
    for i in range(n):
        w = (w + 1) & 0xFFFFFFFF
        px = int(pcg4d(x, y, z, w)[0])
        out[i] = np.float32((px >> 9) / float(1 << 23))
    return out, w


def tea16(v0, v1):
    v0, v1, s0 = int(v0), int(v1), 0
    for _ in range(16):
        s0 = (s0 + 0x9E3779B9) & 0xFFFFFFFF
        v0 = (v0 + ((((v1 << 4) + 0xA341316C) & 0xFFFFFFFF) ^ ((v1 + s0) & 0xFFFFFFFF)
                    ^ (((v1 >> 5) + 0xC8013EA4) & 0xFFFFFFFF))) & 0xFFFFFFFF
        v1 = (v1 + ((((v0 << 4) + 0xAD90777D) & 0xFFFFFFFF) ^ ((v0 + s0) & 0xFFFFFFFF)
                    ^ (((v0 >> 5) + 0x7E95761E) & 0xFFFFFFFF))) & 0xFFFFFFFF
    return v0


def lcg_sequence(seed, n):
    seed = int(seed)
    out = np.zeros(n, np.float32)
    for i in range(n):
        seed = (1664525 * seed + 1013904223) & 0xFFFFFFFF
        out[i] = np.float32((seed & 0xFFFFFF) / float(1 << 24))
    return out, seed
