"""Pins of the oracle's skinning (AnimationCompute.comp:14-38) by closed forms:
the identity, a rigid bone, a blend of translations, linearity in the weights,
unskinned vertices and out-of-range indices."""
import numpy as np

from oracle import binding as B
from rustracer_amd import layouts as L


def _verts(n, rng):
    v = np.zeros(n, L.VERTEX)
    v["pos"][:, :3] = rng.uniform(-5, 5, (n, 3))
    nrm = rng.normal(size=(n, 3))
    v["normal"][:, :3] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    tan = rng.normal(size=(n, 3))
    v["tangent"][:, :3] = tan / np.linalg.norm(tan, axis=1, keepdims=True)
    v["tangent"][:, 3] = rng.choice([-1.0, 1.0], n)
    v["color"] = rng.uniform(0, 1, (n, 4))
    v["uv0"] = rng.uniform(0, 1, (n, 2))
    v["skin_index"] = 0
    v["joints"] = (0, 0, 0, 0)
    v["weights"] = (1, 0, 0, 0)
    return v


def _skins(mats):
    sk = np.zeros((1, 256, 4, 4), np.float32)
    sk[0, :] = np.eye(4, dtype=np.float32)
    for j, m in mats.items():
        sk[0, j] = np.asarray(m, np.float64).T.astype(np.float32)
    return sk


def _rot(axis, ang):
    a = np.asarray(axis, float) / np.linalg.norm(axis)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K


def test_identity_and_unskinned_vertices_are_copied():
    rng = np.random.default_rng(1)
    v = _verts(50, rng)
    v["skin_index"][::2] = -1
    out, bad = B.skin_vertices(v, _skins({}))
    assert bad == 0
    assert out[::2].tobytes() == v[::2].tobytes()
    assert np.array_equal(out["pos"], v["pos"]) and np.array_equal(out["color"], v["color"])
    assert np.allclose(out["normal"][:, :3], v["normal"][:, :3], atol=1e-7)


def test_rigid_bone_rotates_and_translates():
    rng = np.random.default_rng(2)
    v = _verts(40, rng)
    R, t = _rot((1, 2, 3), 0.7), np.array([0.5, -1.0, 2.0])
    m = np.eye(4)
    m[:3, :3], m[:3, 3] = R, t
    v["joints"] = (3, 0, 0, 0)
    out, _ = B.skin_vertices(v, _skins({3: m}))
    p = v["pos"][:, :3].astype(np.float64) @ R.T + t
    assert np.allclose(out["pos"][:, :3], p, atol=1e-5)
    assert np.allclose(out["normal"][:, :3], v["normal"][:, :3] @ R.T, atol=1e-6)
    assert np.allclose(out["tangent"][:, :3], v["tangent"][:, :3] @ R.T, atol=1e-6)
    assert np.array_equal(out["tangent"][:, 3], v["tangent"][:, 3])  # handedness kept


def test_blend_of_translations_is_the_weighted_translation_and_linear():
    rng = np.random.default_rng(3)
    v = _verts(30, rng)
    t1, t2 = np.array([1.0, 0, 0]), np.array([0, 2.0, -1.0])
    m1, m2 = np.eye(4), np.eye(4)
    m1[:3, 3], m2[:3, 3] = t1, t2
    w = rng.uniform(0, 1, 30)
    v["joints"] = (1, 2, 0, 0)
    v["weights"] = np.stack([w, 1 - w, 0 * w, 0 * w], 1)
    out, _ = B.skin_vertices(v, _skins({1: m1, 2: m2}))
    expect = v["pos"][:, :3] + w[:, None] * t1 + (1 - w)[:, None] * t2
    assert np.allclose(out["pos"][:, :3], expect, atol=1e-5)
    # a uniform scale of the weights scales the transform (no renormalisation of weights)
    v["weights"] = np.stack([2 * w, 2 * (1 - w), 0 * w, 0 * w], 1)
    out2, _ = B.skin_vertices(v, _skins({1: m1, 2: m2}))
    assert np.allclose(out2["pos"][:, :3], 2 * expect, atol=2e-5)
    assert np.allclose(out2["normal"][:, :3], out["normal"][:, :3], atol=1e-6)  # normalised


def test_out_of_range_skin_or_joint_is_counted_and_copied():
    rng = np.random.default_rng(4)
    v = _verts(4, rng)
    v["skin_index"][1] = 1       # one skin only
    v["joints"][2] = (0, 256, 0, 0)
    out, bad = B.skin_vertices(v, _skins({}))
    assert bad == 2
    assert out[1].tobytes() == v[1].tobytes() and out[2].tobytes() == v[2].tobytes()
