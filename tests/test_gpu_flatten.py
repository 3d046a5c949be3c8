"""Identity-instance flattening (RT_OPTION_FLATTEN_IDENTITY_INSTANCES): the
world-space BVH over the triangles of identity instances must give the same hits
(instance, primitive, t, barycentrics) and therefore the same images as entering
each instance through the TLAS, as the driver's two-level structure does for the
reference (crates/libs/asset_loader/src/acceleration_structures.rs:136-249)."""
import numpy as np
import pytest

from rustracer_amd import bvhcheck, rt
from tests import gpu_kit as G
from tests import scene_kit as K

pytestmark = pytest.mark.gpu


def _render(ctx, ds, mode, stats=True):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=stats)
    ctx.sync()
    return acc.cpu().numpy(), st


@pytest.mark.parametrize("name,n_flat", [("cornell_sphere", 3), ("displaced_sphere", 1951)])
def test_flattened_structure_is_valid(gpu_ctx, name, n_flat):
    ds = G.device_scene(gpu_ctx, name)
    n = len(ds.scene.instances)
    errs = bvhcheck.check_tlas(gpu_ctx, ds.tlas, n)
    assert not errs, errs[:3]
    assert ds.tlas.info().num_prims == n - n_flat
    fi = ds.tlas.flat_info()
    flat_ids = [i for i in range(n) if np.array_equal(ds.scene.instances[i]["transform"], np.eye(3, 4, dtype=np.float32))]
    expect = sum(int(ds.scene.geometries["prim_count"][int(ds.scene.instances[i]["acceleration_structure_reference"])])
                 for i in flat_ids)
    assert len(flat_ids) == n_flat and fi.num_prims == expect
    # the two-level build of the same scene has no flattened part
    ds2 = G.device_scene(gpu_ctx, name, flatten=False)
    assert ds2.tlas.flat_info().num_prims == 0 and ds2.tlas.info().num_prims == n
    assert not bvhcheck.check_tlas(gpu_ctx, ds2.tlas, n)


@pytest.mark.parametrize("name", ["cornell_sphere", "displaced_sphere"])
@pytest.mark.parametrize("mode", [rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT])
def test_flattened_and_two_level_render_the_same_image(gpu_ctx, name, mode):
    flat = G.device_scene(gpu_ctx, name)
    two = G.device_scene(gpu_ctx, name, flatten=False)
    a, sa = _render(gpu_ctx, flat, mode)
    b, sb = _render(gpu_ctx, two, mode)
    assert np.array_equal(a, b)
    for k in ("rays", "primary_rays", "extension_rays", "shadow_rays", "shade_events"):
        assert sa[k] == sb[k], k
    assert sb["instance_visits"] > 0
    if name == "displaced_sphere":
        assert sa["instance_visits"] == 0  # every instance is flattened


def test_flattened_and_two_level_primary_ids_match(gpu_ctx):
    flat = G.device_scene(gpu_ctx, "displaced_sphere")
    two = G.device_scene(gpu_ctx, "displaced_sphere", flatten=False)
    sc = flat.scene
    a = rt.trace_primary(gpu_ctx, flat.tlas, flat.buffers, sc.ubo, sc.width, sc.height)
    b = rt.trace_primary(gpu_ctx, two.tlas, two.buffers, sc.ubo, sc.width, sc.height)
    assert a.tobytes() == b.tobytes()


def test_flattened_alpha_mask_matches_two_level(gpu_ctx):
    """MASK any-hit decisions do not depend on traversal order, so the images match."""
    def make():
        sc = K.triangle_soup(400, seed=31, opaque=False)
        rng = np.random.default_rng(32)
        sc.vertices["color"][:, 3] = rng.random(len(sc.vertices))
        sc.ubo["fully_opaque"] = 0
        sc.ubo["number_of_bounces"] = 5
        sc.ubo["number_of_samples"] = sc.ubo["total_number_of_samples"] = 3
        from rustracer_amd import scenes as S
        sc.plights = np.array([S.point_light((0.0, 3.0, 4.0), 30.0)])
        return sc
    flat = G.device_scene(gpu_ctx, "mask_soup_lit", make)
    two = G.device_scene(gpu_ctx, "mask_soup_lit", make, flatten=False)
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        a, sa = _render(gpu_ctx, flat, mode)
        b, sb = _render(gpu_ctx, two, mode)
        assert sa["anyhit_calls"] > 0
        assert np.array_equal(a, b)


def test_mixed_identity_and_transformed_instances(gpu_ctx):
    """Instances of one BLAS, some at the identity and some transformed: ray queries
    agree between the flattened and two-level builds, hit for hit."""
    def make():
        sc = K.instanced_spheres(n_inst=40, seed=8, level=2)
        sc.instances["transform"][::3] = np.eye(3, 4, dtype=np.float32)
        return sc
    flat = G.device_scene(gpu_ctx, "mixed_inst", make)
    two = G.device_scene(gpu_ctx, "mixed_inst", make, flatten=False)
    assert flat.tlas.flat_info().num_prims > 0
    assert not bvhcheck.check_tlas(gpu_ctx, flat.tlas, len(flat.scene.instances))
    rays = K.random_rays(40000, [-8, -8, -8], [8, 8, 8], seed=21, toward=(0, 0, 0))
    for flags in (rt.RAY_FLAG_OPAQUE, rt.RAY_FLAG_OPAQUE | rt.RAY_FLAG_TERMINATE_ON_FIRST_HIT):
        a = rt.trace_ray_queries(gpu_ctx, flat.tlas, rays, flags)
        b = rt.trace_ray_queries(gpu_ctx, two.tlas, rays, flags)
        if flags & rt.RAY_FLAG_TERMINATE_ON_FIRST_HIT:
            assert np.array_equal(a["primitive_id"] == G.MISS, b["primitive_id"] == G.MISS)
        else:
            assert a.tobytes() == b.tobytes()
