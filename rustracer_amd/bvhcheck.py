"""Host-side validity checks of a built BVH (used by tests and the smoke run).

A BLAS is valid when, walking from its root: every child reference is in range,
no node is reached twice, every triangle record of the BLAS lies in exactly one
leaf, and every triangle lies inside every box on its path from the root, as the
traversal decodes the quantised boxes (in fp32, no tolerance)."""
import numpy as np

from . import rt

TRI = np.dtype([("v0", "<f4", 3), ("prim", "<u4"), ("v1", "<f4", 4), ("v2", "<f4", 4)])
EMPTY = 0xFFFFFFFF


def check_blas(ctx, blas, expect_prims=None):
    """Errors of one BLAS (empty list: valid)."""
    return check_blases(ctx, [blas], None if expect_prims is None else [expect_prims])[0]


def check_blases(ctx, blases, expect_prims=None):
    """Errors of each BLAS.  The BLASes of one batched build share a node pool and a
    triangle pool, which are read and decoded once."""
    infos = [b.info() for b in blases]
    pools = {}
    for k, info in enumerate(infos):
        pools.setdefault((info.nodes, info.num_nodes, info.tris), []).append(k)
    out = [None] * len(blases)
    for members in pools.values():
        info0 = infos[members[0]]
        pool = _Pool(rt.read_nodes(ctx, info0))
        n_total = max(infos[k].first_prim + infos[k].num_prims for k in members)
        tris = np.frombuffer(rt.read_device(info0.tris, 48 * max(1, n_total)), TRI)[:n_total] if n_total else np.zeros(0, TRI)
        for k in members:
            info = infos[k]
            out[k] = _check(pool, info.root, info.first_prim, info.num_prims, tris, "triangle",
                            None if expect_prims is None else expect_prims[k])
    return out


class _Pool:
    """The decoded nodes of one node pool, and the per-node scratch of a walk (visit
    counts, path boxes), which each walk resets on the nodes it reached so that the
    BLASes of one pool are walked without re-decoding or re-allocating the pool."""

    def __init__(self, nodes):
        self.n = n = len(nodes)
        self.child = nodes["child"].astype(np.int64)
        self.count = nodes["count"].astype(np.int64)
        self.lo = np.stack([nodes["lo_x"], nodes["lo_y"], nodes["lo_z"]], -1)  # (n, 4, 3)
        self.hi = np.stack([nodes["hi_x"], nodes["hi_y"], nodes["hi_z"]], -1)
        self.visits = np.zeros(n, np.int64)
        # intersection of the decoded boxes on the path to each reached node
        self.path_lo = np.full((n, 3), -np.inf, np.float32)
        self.path_hi = np.full((n, 3), np.inf, np.float32)

    def reset(self, reached):
        self.visits[reached] = 0
        self.path_lo[reached] = -np.inf
        self.path_hi[reached] = np.inf


def _check(pool, root, first_prim, num_prims, tris, kind, expect_prims=None):
    """Level-by-level walk from `root` over all nodes of a level at once."""
    n = pool.n
    if root >= n:
        return [f"node index {root} out of range {n}"]
    reached = [np.array([root], np.int64)]
    try:
        return _walk(pool, root, first_prim, num_prims, tris, kind, expect_prims, reached)
    finally:
        pool.reset(np.concatenate(reached))


def _walk(pool, root, first_prim, num_prims, tris, kind, expect_prims, reached):
    errors = []
    n, child, count, lo, hi = pool.n, pool.child, pool.count, pool.lo, pool.hi
    visits, path_lo, path_hi = pool.visits, pool.path_lo, pool.path_hi
    visits[root] = 1
    frontier = np.array([root], np.int64)
    leaf_first, leaf_cnt, leaf_lo, leaf_hi = [], [], [], []
    while frontier.size:
        c, cnt = child[frontier], count[frontier]
        valid = c != EMPTY
        if np.any(valid & np.any(lo[frontier] > hi[frontier], -1)):
            errors.append("a child box is inverted")
        blo = np.maximum(lo[frontier], path_lo[frontier][:, None, :])
        bhi = np.minimum(hi[frontier], path_hi[frontier][:, None, :])
        internal = valid & (cnt == 0)
        ci = c[internal]
        if np.any(ci >= n):
            errors.append(f"node index {int(ci.max())} out of range {n}")
            keep = ci < n
            ci, ilo, ihi = ci[keep], blo[internal][keep], bhi[internal][keep]
        else:
            ilo, ihi = blo[internal], bhi[internal]
        reached.append(ci)
        np.add.at(visits, ci, 1)
        if np.any(visits[ci] > 1):
            errors.append(f"{int(np.sum(visits[ci] > 1))} nodes reached twice")
        first_time = visits[ci] == 1
        ci, ilo, ihi = ci[first_time], ilo[first_time], ihi[first_time]
        path_lo[ci] = ilo
        path_hi[ci] = ihi
        leaf = valid & (cnt > 0)
        leaf_first.append(c[leaf])
        leaf_cnt.append(cnt[leaf])
        leaf_lo.append(blo[leaf])
        leaf_hi.append(bhi[leaf])
        frontier = np.unique(ci)
    lf, lc = np.concatenate(leaf_first), np.concatenate(leaf_cnt)
    llo, lhi = np.concatenate(leaf_lo), np.concatenate(leaf_hi)
    bad = (lf < first_prim) | (lf + lc > first_prim + num_prims)
    if np.any(bad):
        errors.append(f"{int(bad.sum())} leaves outside [{first_prim},{first_prim + num_prims})")
        lf, lc, llo, lhi = lf[~bad], lc[~bad], llo[~bad], lhi[~bad]
    # every record of every leaf, with its leaf's path box
    rec = np.repeat(lf, lc) + (np.arange(int(lc.sum())) - np.repeat(np.cumsum(lc) - lc, lc))
    covered = np.bincount(rec - first_prim, minlength=num_prims)[:num_prims] if num_prims else np.zeros(0, np.int64)
    if np.any(covered != 1):
        badr = np.nonzero(covered != 1)[0]
        errors.append(f"{len(badr)} records not in exactly one leaf (first {badr[:5] + first_prim}, counts {covered[badr[:5]]})")
    if kind in ("triangle", "flat") and len(rec):
        rlo, rhi = np.repeat(llo, lc, axis=0), np.repeat(lhi, lc, axis=0)
        t = tris[rec]
        for v in (t["v0"], t["v1"][:, :3], t["v2"][:, :3]):
            out = np.any(v < rlo, -1) | np.any(v > rhi, -1)
            if np.any(out):
                errors.append(f"{int(out.sum())} triangles not inside every box on their path")
                break
    if kind == "triangle" and num_prims:
        ids = np.sort(tris["prim"][first_prim:first_prim + num_prims])
        if not np.array_equal(ids, np.arange(num_prims, dtype=np.uint32)):
            errors.append("primitive ids of the BLAS are not a permutation of 0..n-1")
    if expect_prims is not None and expect_prims != num_prims:
        errors.append(f"BLAS has {num_prims} primitives, expected {expect_prims}")
    return errors


INST = np.dtype([("w2o_d", "<f8", (3, 4)), ("nodes", "<u8"), ("tris", "<u8"), ("root", "<u4"), ("mask", "<u4"),
                 ("flags", "<u4"), ("instance_id", "<u4"), ("o2w", "<f4", (3, 4)), ("blas", "<u8"),
                 ("custom_index", "<u4"), ("_pad", "<u4", 17)])
FLAT_OPAQUE = 0x80000000


def check_tlas(ctx, tlas, num_instances):
    """The TLAS proper, the instance records, and the flattened BVH over identity
    instances: every instance is reached exactly once, through one or the other."""
    info = tlas.info()
    nodes = rt.read_nodes(ctx, info)
    errs = _check(_Pool(nodes), info.root, 0, info.num_prims, None, "instance")
    inst = np.frombuffer(rt.read_device(info.tris, INST.itemsize * max(1, num_instances)), INST)[:num_instances]
    if num_instances and not np.array_equal(np.sort(inst["instance_id"]), np.arange(num_instances)):
        errs.append("instance ids are not a permutation")
    fi = tlas.flat_info()
    nt = info.num_prims
    if fi.num_prims == 0:
        if nt != num_instances:
            errs.append(f"TLAS reaches {nt} of {num_instances} instances and there is no flattened BVH")
        return errs
    fnodes = rt.read_nodes(ctx, fi)
    ftris = np.frombuffer(rt.read_device(fi.tris, 48 * fi.num_prims), TRI)
    errs += _check(_Pool(fnodes), fi.root, 0, fi.num_prims, ftris, "flat")
    bits = ftris["v1"][:, 3].copy().view("<u4")
    ii = bits & ~np.uint32(FLAT_OPAQUE)
    if np.any(ii < nt) or np.any(ii >= num_instances):
        errs.append("flattened records reference instances outside the flattened range")
        return errs
    counts = np.bincount(ii - nt, minlength=num_instances - nt)
    if np.any(counts == 0):
        errs.append(f"{int(np.sum(counts == 0))} instance records after the TLAS ones have no flattened triangles")
    order = np.lexsort((ftris["prim"], ii))
    pi, pp = ii[order], ftris["prim"][order]
    starts = np.concatenate([[0], np.cumsum(counts)[:-1]])
    expect = np.arange(len(pp)) - np.repeat(starts, counts)
    if not np.array_equal(pp, expect.astype(np.uint32)):
        errs.append("primitive ids of a flattened instance are not a permutation of 0..n-1")
    if np.any(np.abs(inst["w2o_d"][nt:] - np.eye(3, 4)) > 0):
        errs.append("a flattened instance does not have the identity transform")
    return errs
