"""The host BVH validity walk (rustracer_amd.bvhcheck) on hand-built trees: a
valid tree passes, and each planted defect (a triangle outside a box on its
path, a record in two leaves, a node reached twice, an out-of-range child, a
missing record) is reported."""
import numpy as np

from rustracer_amd import bvhcheck as V
from rustracer_amd import rt

E = V.EMPTY


def _tris(n):
    t = np.zeros(n, V.TRI)
    for i in range(n):  # unit triangles along x, one per unit cell
        t["v0"][i] = (i, 0, 0)
        t["v1"][i] = (i + 0.5, 0.5, 0, 0)
        t["v2"][i] = (i, 0.5, 0.5, 0)
        t["prim"][i] = i
    return t


def _node(children):
    """children: list of (lo, hi, child, count)."""
    n = np.zeros(1, rt.NODE)[0]
    n["child"][:] = E
    for k, (lo, hi, c, cnt) in enumerate(children):
        n["lo_x"][k], n["lo_y"][k], n["lo_z"][k] = lo
        n["hi_x"][k], n["hi_y"][k], n["hi_z"][k] = hi
        n["child"][k], n["count"][k] = c, cnt
    return n


def _tree():
    """root -> node 1 (leaves [0,2), [2,3)) and a leaf [3,5); 5 triangles."""
    nodes = np.zeros(2, rt.NODE)
    nodes[0] = _node([((0, 0, 0), (3, 1, 1), 1, 0), ((3, 0, 0), (5, 1, 1), 3, 2)])
    nodes[1] = _node([((0, 0, 0), (2, 1, 1), 0, 2), ((2, 0, 0), (3, 1, 1), 2, 1)])
    return nodes, _tris(5)


def _chk(nodes, *args):
    return V._check(V._Pool(nodes), *args)


def test_valid_tree_passes():
    nodes, tris = _tree()
    assert _chk(nodes, 0, 0, 5, tris, "triangle") == []


def test_triangle_outside_an_ancestor_box_is_reported():
    nodes, tris = _tree()
    # the leaf box still contains triangle 2, the root's box of node 1 no longer does
    nodes[0]["hi_x"][0] = 2.4
    assert any("not inside every box" in e for e in _chk(nodes, 0, 0, 5, tris, "triangle"))


def test_record_in_two_leaves_and_missing_record():
    nodes, tris = _tree()
    nodes[1]["child"][1] = 1  # leaf [1,2) instead of [2,3): record 1 twice, record 2 never
    errs = _chk(nodes, 0, 0, 5, tris, "triangle")
    assert any("not in exactly one leaf" in e for e in errs)


def test_node_reached_twice_and_out_of_range_child():
    nodes, tris = _tree()
    nodes[0]["child"][1], nodes[0]["count"][1] = 1, 0
    assert any("reached twice" in e for e in _chk(nodes, 0, 0, 5, tris, "triangle"))
    nodes, tris = _tree()
    nodes[0]["child"][0] = 7
    assert any("out of range" in e for e in _chk(nodes, 0, 0, 5, tris, "triangle"))


def test_inverted_box_and_foreign_leaf():
    nodes, tris = _tree()
    nodes[1]["lo_y"][0] = 2.0
    assert any("inverted" in e for e in _chk(nodes, 0, 0, 5, tris, "triangle"))
    nodes, tris = _tree()
    assert any("leaves outside" in e for e in _chk(nodes, 0, 1, 4, tris, "triangle"))


def test_walks_of_one_pool_do_not_see_each_other():
    """Two BLASes in one node pool (as a batched build leaves them), walked in turn
    with the pool decoded once: a node that the first walk reached twice leaves no
    visit count or path box behind for the second walk or for a repeat of the first."""
    a, ta = _tree()
    b, tb = _tree()
    a[0]["child"][1], a[0]["count"][1] = 1, 0  # first BLAS: node 1 reached twice
    real = b["child"] != E
    b["child"][real & (b["count"] == 0)] += 2   # second BLAS: nodes 2, 3
    b["child"][real & (b["count"] > 0)] += 5    # and records 5..9
    nodes, tris = np.concatenate([a, b]), np.concatenate([ta, tb])
    tris["prim"][5:] = np.arange(5)
    pool = V._Pool(nodes)
    first = V._check(pool, 0, 0, 5, tris, "triangle")
    assert any("reached twice" in e for e in first)
    assert V._check(pool, 2, 5, 5, tris, "triangle") == []
    assert V._check(pool, 0, 0, 5, tris, "triangle") == first
    assert not pool.visits.any() and np.all(pool.path_lo == -np.inf) and np.all(pool.path_hi == np.inf)
