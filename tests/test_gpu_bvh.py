"""BVH validity of the gfx950 LBVH builder (bit-exact structural checks): every
triangle in exactly one leaf, every box conservative in float, node references in
range, primitive ids a permutation; TLAS over instances likewise.  The builder
replaces the driver build behind vkCmdBuildAccelerationStructuresKHR
(crates/libs/vulkan/src/command.rs:354-372) for create_as / create_top_as
(crates/libs/asset_loader/src/acceleration_structures.rs:79-249)."""
import numpy as np
import pytest

from rustracer_amd import bvhcheck, layouts as L, rt, scenes as S
from tests import gpu_kit as G
from tests import scene_kit as K

pytestmark = pytest.mark.gpu


def _check_all(ctx, ds, every=1):
    sc = ds.scene
    idx = list(range(0, len(ds.blases), every))
    all_errs = bvhcheck.check_blases(ctx, [ds.blases[g] for g in idx],
                                     [int(sc.geometries["prim_count"][g]) for g in idx])
    for g, errs in zip(idx, all_errs):
        assert not errs, (g, errs[:3])
    errs = bvhcheck.check_tlas(ctx, ds.tlas, len(sc.instances))
    assert not errs, errs[:3]


def test_cornell_bvh_valid(gpu_ctx):
    _check_all(gpu_ctx, G.device_scene(gpu_ctx, "cornell_sphere"))


def test_displaced_sphere_all_1951_blas_valid(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "displaced_sphere")
    assert len(ds.blases) == 1951 and ds.scene.num_triangles == 1997120
    _check_all(gpu_ctx, ds)


def test_single_large_blas_valid(gpu_ctx):
    # one BLAS over a 327,680-triangle sphere: deep tree, full radix passes
    def make():
        b = S.Builder()
        mid = b.add_material(L.default_material())
        p, t = S.geodesic_sphere(128)
        g = b.add_primitive(S.make_vertices(p, p), t.reshape(-1), mid)
        return b.finish("big", S.make_instances([(None, g)]), [], [L.default_dlight()], K._ubo(8, 8), 8, 8)
    ds = G.device_scene(gpu_ctx, "big_sphere", make)
    _check_all(gpu_ctx, ds)


def test_degenerate_and_duplicate_triangles(gpu_ctx):
    # zero-area triangles and identical centroids (equal Morton keys)
    def make():
        b = S.Builder()
        mid = b.add_material(L.default_material())
        p = np.zeros((300, 3))
        p[:150] = [1.0, 2.0, 3.0]                      # 50 fully degenerate triangles at one point
        p[150:] = np.repeat(np.random.default_rng(0).normal(size=(50, 3)), 3, axis=0)
        p[150::3] += 0.1
        v = S.make_vertices(p, np.tile([0.0, 0.0, 1.0], (300, 1)))
        g = b.add_primitive(v, np.arange(300, dtype=np.uint32), mid)
        return b.finish("degen", S.make_instances([(None, g)]), [], [], K._ubo(8, 8), 8, 8)
    ds = G.device_scene(gpu_ctx, "degenerate", make)
    _check_all(gpu_ctx, ds)


def test_tiny_and_empty_blas_in_one_batch(gpu_ctx):
    # geometries of 0, 1, 2, 3, 4, 5 and 17 triangles built in one batch
    def make():
        b = S.Builder()
        mid = b.add_material(L.default_material())
        rng = np.random.default_rng(1)
        for n in (0, 1, 2, 3, 4, 5, 17):
            p = rng.normal(size=(3 * max(n, 1), 3))
            v = S.make_vertices(p, np.tile([0.0, 0.0, 1.0], (len(p), 1)))
            b.add_primitive(v, np.arange(3 * n, dtype=np.uint32), mid)
        inst = S.make_instances([(None, g) for g in range(7)])
        return b.finish("tiny", inst, [], [], K._ubo(8, 8), 8, 8)
    ds = G.device_scene(gpu_ctx, "tiny", make)
    _check_all(gpu_ctx, ds)
    assert ds.blases[0].info().num_prims == 0


def test_tlas_over_1000_instances(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "inst1000", lambda: K.instanced_spheres(n_inst=1000, seed=5, level=1))
    _check_all(gpu_ctx, ds)


def test_build_is_deterministic(gpu_ctx):
    sc = S.cornell_sphere()
    a = G.DeviceScene(gpu_ctx, sc)
    b = G.DeviceScene(gpu_ctx, sc)
    for x, y in zip(a.blases, b.blases):
        nx = rt.read_nodes(gpu_ctx, x.info())
        ny = rt.read_nodes(gpu_ctx, y.info())
        assert nx.tobytes() == ny.tobytes()
