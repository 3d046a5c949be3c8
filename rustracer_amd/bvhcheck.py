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
    info = blas.info()
    nodes = rt.read_nodes(ctx, info)
    n_total = info.first_prim + info.num_prims
    tris = np.frombuffer(rt.read_device(info.tris, 48 * max(1, n_total)), TRI)[:n_total] if n_total else np.zeros(0, TRI)
    return _check(nodes, info.root, info.first_prim, info.num_prims, tris, "triangle", expect_prims)


def _check(nodes, root, first_prim, num_prims, tris, kind, expect_prims=None):
    errors = []
    seen_nodes = np.zeros(len(nodes), bool)
    covered = np.zeros(num_prims, np.int32)
    stack = [(root, None)]  # (node, intersection of the decoded boxes on the path)
    while stack:
        ni, path_box = stack.pop()
        if ni >= len(nodes):
            errors.append(f"node index {ni} out of range {len(nodes)}")
            continue
        if seen_nodes[ni]:
            errors.append(f"node {ni} reached twice")
            continue
        seen_nodes[ni] = True
        nd = nodes[ni]
        for k in range(4):
            c, cnt = int(nd["child"][k]), int(nd["count"][k])
            if c == EMPTY:
                continue
            lo = np.array([nd["lo_x"][k], nd["lo_y"][k], nd["lo_z"][k]], np.float32)
            hi = np.array([nd["hi_x"][k], nd["hi_y"][k], nd["hi_z"][k]], np.float32)
            if np.any(lo > hi):
                errors.append(f"node {ni} child {k} box is inverted")
            box = (lo, hi) if path_box is None else (np.maximum(lo, path_box[0]), np.minimum(hi, path_box[1]))
            if cnt == 0:
                stack.append((c, box))
                continue
            if c < first_prim or c + cnt > first_prim + num_prims:
                errors.append(f"leaf [{c},{c + cnt}) outside [{first_prim},{first_prim + num_prims})")
                continue
            covered[c - first_prim:c - first_prim + cnt] += 1
            if kind in ("triangle", "flat"):
                t = tris[c:c + cnt]
                pts = np.concatenate([t["v0"], t["v1"][:, :3], t["v2"][:, :3]])
                if np.any(pts < box[0]) or np.any(pts > box[1]):
                    errors.append(f"leaf [{c},{c + cnt}) triangles not inside every box on their path")
    if np.any(covered != 1):
        bad = np.nonzero(covered != 1)[0]
        errors.append(f"{len(bad)} records not in exactly one leaf (first {bad[:5] + first_prim}, counts {covered[bad[:5]]})")
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
    errs = _check(nodes, info.root, 0, info.num_prims, None, "instance")
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
    errs += _check(fnodes, fi.root, 0, fi.num_prims, ftris, "flat")
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
