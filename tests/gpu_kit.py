"""Shared set-up of the GPU tests: device scenes (buffers + acceleration structures)
cached per process, and the near-edge tie explanation used by the ID parity checks."""
import functools

import numpy as np

from rustracer_amd import layouts as L
from rustracer_amd import rt
from rustracer_amd import scenes as S

MISS = 0xFFFFFFFF


class DeviceScene:
    def __init__(self, ctx, scene, flatten=True):
        import torch
        self.scene = scene
        self.buffers = rt.Buffers(scene)
        torch.cuda.synchronize()
        ctx.set_option(rt.OPTION_FLATTEN_IDENTITY_INSTANCES, int(flatten))
        try:
            self.blases, self.tlas = rt.create_as(ctx, scene, self.buffers)
        finally:
            ctx.set_option(rt.OPTION_FLATTEN_IDENTITY_INSTANCES, 1)
        ctx.sync()


@functools.lru_cache(maxsize=None)
def host_scene(name):
    return S.WORKLOADS[name]()


_dev = {}


def device_scene(ctx, name, factory=None, flatten=True):
    """Buffers + BLASes + TLAS of a named scene, cached per context.  flatten=False
    builds the TLAS with every instance entered through it (two-level traversal)."""
    key = (id(ctx), name, flatten)
    if key not in _dev:
        sc = factory() if factory is not None else host_scene(name)
        _dev[key] = DeviceScene(ctx, sc, flatten)
    return _dev[key]


def triangle_world(scene, inst_id, prim):
    """World-space vertices (float64) of triangle `prim` of instance `inst_id`."""
    inst = scene.instances[inst_id]
    geo = scene.geometries[int(inst["acceleration_structure_reference"])]
    idx = scene.indices[int(geo["index_offset"]) + 3 * prim: int(geo["index_offset"]) + 3 * prim + 3]
    p = scene.vertices["pos"][int(geo["first_vertex"]) + idx.astype(np.int64), :3].astype(np.float64)
    m = inst["transform"].astype(np.float64)
    return p @ m[:, :3].T + m[:, 3]


def edge_distance(o, d, tri):
    """Smallest barycentric coordinate (normalised) of the line o + t d on the
    triangle's plane, and the line's t there: ~0 means the ray grazes an edge."""
    v0, v1, v2 = tri
    e1, e2 = v1 - v0, v2 - v0
    p = np.cross(d, e2)
    det = e1 @ p
    if abs(det) < 1e-300:
        return 0.0, np.inf
    s = o - v0
    u = (s @ p) / det
    q = np.cross(s, e1)
    v = (d @ q) / det
    t = (e2 @ q) / det
    return min(1 - u - v, u, v), t


def explain_mismatches(scene, rays_o, rays_d, gpu, orc, orc_t, tol_bary=2e-5, tol_t=1e-5):
    """For rays where GPU and oracle report different (instance, primitive), check the
    difference is a documented sub-ulp tie: both candidate triangles are hit within
    tol_bary of an edge (in barycentric terms), or their t values agree to tol_t
    (coplanar/coincident surfaces).  Returns the indices that are NOT explained."""
    bad = []
    for i in np.nonzero((gpu["instance_id"] != orc["instance_id"]) | (gpu["primitive_id"] != orc["primitive_id"]))[0]:
        o, d = rays_o[i].astype(np.float64), rays_d[i].astype(np.float64)
        cands = []
        for h in (gpu[i], orc[i]):
            if h["primitive_id"] == MISS:
                cands.append(None)
                continue
            cands.append(edge_distance(o, d, triangle_world(scene, int(h["instance_id"]), int(h["primitive_id"]))))
        ok = False
        if cands[0] is not None and cands[1] is not None:
            (b0, t0), (b1, t1) = cands
            ok = abs(t0 - t1) <= tol_t * max(1.0, abs(t1)) or (b0 > -tol_bary and abs(b0) < tol_bary) or \
                abs(b1) < tol_bary
        else:
            c = cands[0] if cands[0] is not None else cands[1]
            ok = abs(c[0]) < tol_bary  # a silhouette edge: hit on one side, miss on the other
        if not ok:
            bad.append(i)
    return np.array(bad, np.int64)


def rel_rmse(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def unexplained_primary_mismatches(gpu, orc, orc_tuv, tol_bary=2e-5, tol_t=1e-5):
    """Camera-ray hits where GPU and oracle name different (instance, primitive) and
    the difference is not a documented sub-ulp tie.  A tie is: both candidates hit
    within tol_bary of one of their triangle's edges (barycentrics of each side's own
    hit record), or their t values agree to tol_t (coincident surfaces), or one side
    misses and the other hits within tol_bary of an edge (a silhouette).  Returns the
    pixel indices left unexplained."""
    diff = np.nonzero((gpu["instance_id"] != orc["instance_id"]) | (gpu["primitive_id"] != orc["primitive_id"]))[0]
    bad = []
    for i in diff:
        g_hit, o_hit = gpu["primitive_id"][i] != MISS, orc["primitive_id"][i] != MISS
        gb = min(1.0 - float(gpu["u"][i]) - float(gpu["v"][i]), float(gpu["u"][i]), float(gpu["v"][i])) if g_hit else None
        ob = min(1.0 - orc_tuv[i, 1] - orc_tuv[i, 2], orc_tuv[i, 1], orc_tuv[i, 2]) if o_hit else None
        if g_hit and o_hit:
            ok = (abs(float(gpu["t"][i]) - orc_tuv[i, 0]) <= tol_t * max(1.0, abs(orc_tuv[i, 0]))
                  or (abs(gb) < tol_bary and abs(ob) < tol_bary))
        else:
            ok = abs(gb if g_hit else ob) < tol_bary
        if not ok:
            bad.append(int(i))
    return bad
