"""Image parity of the gfx950 trace (megakernel) against the FP64 oracle of the same
integrator with the same counter-based RNG streams (relative RMSE of the
accumulated radiance <= 1e-3), plus the analytic images, the accumulation rule
and determinism.  Restates crates/examples/gltf_viewer/shaders/RayTracing.rgen
(:25-167), RayTracing.rchit (:136-477) and RayTracing.rmiss (:14-45)."""
import numpy as np
import pytest
import torch

from oracle import binding as B
from rustracer_amd import rt
from rustracer_amd import scenes as S
from tests import gpu_kit as G
from tests import scene_kit as K

pytestmark = pytest.mark.gpu
REL_RMSE_GATE = 1e-3


def _render(ctx, ds, ubo=None, mode=rt.MODE_MEGAKERNEL, accum=None, stats=False):
    sc = ds.scene
    ubo = sc.ubo if ubo is None else ubo
    acc = rt.new_accumulation(sc.width, sc.height) if accum is None else accum
    st = rt.trace_rays(ctx, ds.tlas, ds.buffers, ubo, acc, sc.width, sc.height, mode=mode, stats=stats)
    ctx.sync()
    return acc, st


def test_cornell_matches_fp64_oracle(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    acc, st = _render(gpu_ctx, ds, stats=True)
    g = acc.cpu().numpy().astype(np.float64)
    o, ost = B.OracleScene(ds.scene).render()
    rel = G.rel_rmse(g[..., :3], o[..., :3])
    print("cornell relRMSE", rel, "rays", st["rays"], ost["rays"])
    assert rel <= REL_RMSE_GATE
    assert np.all(g[..., 3] == 0)
    # same path structure: ray counts agree up to rare decision flips
    assert abs(st["rays"] - ost["rays"]) <= 1e-3 * ost["rays"]
    assert st["primary_rays"] == ost["primary_rays"] == 256 * 256 * 4
    assert st["stack_overflows"] == 0


def test_displaced_sphere_matches_fp64_oracle_on_a_stratified_subsample(gpu_ctx):
    """Full 1920x1080, 64 spp, depth 8 frame on the GPU; the oracle renders every
    16th pixel of every 16th row (8100 pixels x 64 spp)."""
    ds = G.device_scene(gpu_ctx, "displaced_sphere")
    sc = ds.scene
    acc, st = _render(gpu_ctx, ds, stats=True)
    g = acc.cpu().numpy().astype(np.float64)
    ys, xs = np.meshgrid(np.arange(7, sc.height, 16), np.arange(7, sc.width, 16), indexing="ij")
    xy = np.stack([xs.reshape(-1), ys.reshape(-1)], 1)
    o, ost = B.OracleScene(sc).render_pixels(xy)
    gs = g[xy[:, 1], xy[:, 0], :3]
    rel = G.rel_rmse(gs, o[:, :3])
    print("displaced relRMSE", rel, "gpu rays", st["rays"])
    assert rel <= REL_RMSE_GATE
    assert st["stack_overflows"] == 0


def test_analytic_miss_constant_and_emissive(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "plane_away", lambda: K.plane_scene(look_away=True, spp=3, bounces=4))
    acc, _ = _render(gpu_ctx, ds)
    assert np.all(acc.cpu().numpy()[..., :3] == np.float32(3 * np.float32(0.01)))
    ds2 = G.device_scene(gpu_ctx, "plane_emissive", lambda: K.plane_scene(
        base=(0, 0, 0), emissive=(0.2, 0.5, 1.0), intensity=0.0, dlights=False, bounces=3, spp=2))
    acc2, _ = _render(gpu_ctx, ds2)
    exp = 2 * np.array([np.float32(0.2) * 5, np.float32(0.5) * 5, 5.0], np.float32)
    assert np.allclose(acc2.cpu().numpy()[..., :3], exp, rtol=1e-6)


def test_single_light_diffuse_closed_form(gpu_ctx):
    base = np.array([0.5, 0.6, 0.7])
    ds = G.device_scene(gpu_ctx, "plane_lit", lambda: K.plane_scene(base=tuple(base), light_h=2.0, intensity=10.0))
    acc, _ = _render(gpu_ctx, ds)
    sc = ds.scene
    c = acc.cpu().numpy()[sc.height // 2, sc.width // 2, :3].astype(np.float64)
    expected = base.astype(np.float32).astype(np.float64) / (1.51 * 3.141592653589) * 10.0 / 3.0
    assert np.allclose(c, expected, rtol=2e-5), (c, expected)


def test_accumulation_across_frames(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_small", lambda: __import__("rustracer_amd.scenes", fromlist=["x"])
                        .cornell_sphere(width=64, height=48, spp=2))
    sc = ds.scene
    first, _ = _render(gpu_ctx, ds)
    f1 = first.clone()
    ubo = sc.ubo.copy()
    ubo["total_number_of_samples"] = 4
    ubo["frame_count"] = 1
    _render(gpu_ctx, ds, ubo=ubo, accum=first)
    fresh = sc.ubo.copy()
    fresh["frame_count"] = 1
    second, _ = _render(gpu_ctx, ds, ubo=fresh)
    assert np.array_equal(first.cpu().numpy()[..., :3], (f1 + second).cpu().numpy()[..., :3])
    # reset: NumberOfSamples == TotalNumberOfSamples overwrites whatever was there
    junk = first.clone().fill_(123.0)
    _render(gpu_ctx, ds, accum=junk)
    assert np.array_equal(junk.cpu().numpy(), f1.cpu().numpy())


def test_deterministic_bit_for_bit(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    a, _ = _render(gpu_ctx, ds)
    b, _ = _render(gpu_ctx, ds)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    # the counting variant renders the same image
    c, _ = _render(gpu_ctx, ds, stats=True)
    assert np.array_equal(a.cpu().numpy(), c.cpu().numpy())


@pytest.mark.parametrize("name", ["cornell_sphere", "displaced_sphere"])
def test_wavefront_matches_megakernel_bit_for_bit(gpu_ctx, name):
    ds = G.device_scene(gpu_ctx, name)
    a, sa = _render(gpu_ctx, ds, mode=rt.MODE_MEGAKERNEL, stats=True)
    b, sb = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT, stats=True)
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    for k in ("rays", "primary_rays", "extension_rays", "shadow_rays", "shade_events", "light_candidates"):
        assert sa[k] == sb[k], k
    c, _ = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT)
    assert np.array_equal(b.cpu().numpy(), c.cpu().numpy())


@pytest.mark.parametrize("w,h", [(37, 21), (1, 1), (1, 65), (9, 1)])
def test_ragged_frames_match_the_oracle_in_both_modes(gpu_ctx, w, h):
    """Frames whose sides are not multiples of the 8x8 slot tile (partial tiles at the
    right and bottom edges, one-pixel frames): every pixel is written (the image starts
    as NaN), the two modes agree bit for bit and the image matches the oracle's."""
    ds = G.device_scene(gpu_ctx, f"cornell_{w}x{h}", lambda: S.cornell_sphere(width=w, height=h, spp=3, bounces=4))
    imgs = []
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        acc = rt.new_accumulation(w, h).fill_(float("nan"))
        _render(gpu_ctx, ds, mode=mode, accum=acc)
        imgs.append(acc.cpu().numpy().astype(np.float64))
    assert np.all(np.isfinite(imgs[0])), "a pixel was not written"
    assert np.array_equal(imgs[0], imgs[1])
    o, _ = B.OracleScene(ds.scene).render()
    assert G.rel_rmse(imgs[0][..., :3], o[..., :3]) <= REL_RMSE_GATE


def test_wavefront_alpha_mask_and_accumulation(gpu_ctx):
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
    ds = G.device_scene(gpu_ctx, "mask_soup_lit", make)
    a, sa = _render(gpu_ctx, ds, mode=rt.MODE_MEGAKERNEL, stats=True)
    b, sb = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT, stats=True)
    assert sa["anyhit_calls"] > 0 and sa["rays"] == sb["rays"]
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    o, _ = B.OracleScene(ds.scene).render()
    assert G.rel_rmse(a.cpu().numpy()[..., :3].astype(np.float64), o[..., :3]) <= REL_RMSE_GATE
    # accumulation onto an existing image
    ubo = ds.scene.ubo.copy()
    ubo["total_number_of_samples"] = 6
    ubo["frame_count"] = 1
    a2, b2 = a.clone(), b.clone()
    _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_MEGAKERNEL, accum=a2)
    _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_WAVEFRONT, accum=b2)
    assert np.array_equal(a2.cpu().numpy(), b2.cpu().numpy())


@pytest.mark.parametrize("mode", [rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT])
def test_stage_timing_and_split_counters(gpu_ctx, mode):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    a, st = _render(gpu_ctx, ds, mode=mode, stats=True)
    gpu_ctx.set_option(rt.OPTION_STAGE_TIMING, 1)
    try:
        b, _ = _render(gpu_ctx, ds, mode=mode)
        t = gpu_ctx.stage_times()
    finally:
        gpu_ctx.set_option(rt.OPTION_STAGE_TIMING, 0)
    # timing changes nothing in the image
    assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())
    parts = t["extend_ms"] + t["shade_ms"] + t["shadow_ms"] + t["resolve_ms"]
    assert t["total_ms"] > 0 and abs(parts - t["total_ms"]) <= 1e-3 * t["total_ms"] + 1e-3
    if mode == rt.MODE_WAVEFRONT:
        assert t["iterations"] >= int(ds.scene.ubo["number_of_bounces"])
        # one trace kernel serves extension and shadow rays; resolve_ms is the init
        assert min(t["extend_ms"], t["shade_ms"], t["resolve_ms"]) > 0 and t["shadow_ms"] == 0
    else:
        assert t["iterations"] == 1 and t["total_ms"] == t["extend_ms"]
    # shadow traversal is a part of the totals, and present (one point light)
    assert 0 < st["shadow_node_visits"] < st["node_visits"]
    assert 0 < st["shadow_tri_tests"] < st["tri_tests"]
    assert st["shadow_instance_visits"] <= st["instance_visits"]
    with pytest.raises(rt.RtError):
        gpu_ctx.stage_times()  # option off


@pytest.mark.parametrize("name", ["cornell_sphere", "mask_soup"])
def test_traversal_variants_render_the_same_image(gpu_ctx, name):
    """Every wavefront traversal variant (stack depth, register target) gives the
    megakernel's image bit for bit."""
    if name == "mask_soup":
        def make():
            sc = K.triangle_soup(400, seed=31, opaque=False)
            rng = np.random.default_rng(32)
            sc.vertices["color"][:, 3] = rng.random(len(sc.vertices))
            sc.ubo["fully_opaque"] = 0
            sc.ubo["number_of_bounces"] = 5
            return sc
        ds = G.device_scene(gpu_ctx, "mask_soup_variants", make)
    else:
        ds = G.device_scene(gpu_ctx, name)
    ref, _ = _render(gpu_ctx, ds, mode=rt.MODE_MEGAKERNEL)
    try:
        for v in range(rt.NUM_TRAVERSAL_VARIANTS):
            gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, v)
            img, _ = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT)
            assert np.array_equal(ref.cpu().numpy(), img.cpu().numpy()), v
        gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, 0)
        for re, rs in ((1, 1), (64, 64), (17, 33)):
            gpu_ctx.set_option(rt.OPTION_EXTEND_REFILL, re)
            gpu_ctx.set_option(rt.OPTION_SHADOW_REFILL, rs)
            img, _ = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT)
            assert np.array_equal(ref.cpu().numpy(), img.cpu().numpy()), (re, rs)
    finally:
        gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, 0)
        gpu_ctx.set_option(rt.OPTION_EXTEND_REFILL, 16)
        gpu_ctx.set_option(rt.OPTION_SHADOW_REFILL, 1)
    with pytest.raises(rt.RtError):
        gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, rt.NUM_TRAVERSAL_VARIANTS)


def test_graph_replay_matches_eager(gpu_ctx):
    """RT_OPTION_GRAPH: the captured chunk is replayed many times per frame, across
    frames with different UBOs (exec update), with eager work in between; every
    image equals the eager one."""
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    eager = []
    for fc in range(3):
        ubo = sc.ubo.copy()
        ubo["frame_count"] = fc
        a, _ = _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_WAVEFRONT)
        eager.append(a.cpu().numpy())
    gpu_ctx.set_option(rt.OPTION_GRAPH, 1)
    try:
        for rep in range(2):
            for fc in range(3):
                ubo = sc.ubo.copy()
                ubo["frame_count"] = fc
                a, _ = _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_WAVEFRONT)
                assert np.array_equal(a.cpu().numpy(), eager[fc]), (rep, fc)
                # eager work between replays
                _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_MEGAKERNEL)
    finally:
        gpu_ctx.set_option(rt.OPTION_GRAPH, 0)


def test_autotune_keeps_the_image(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    ref, _ = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT)
    try:
        t = gpu_ctx.autotune(ds.tlas, ds.buffers, sc.ubo, sc.width, sc.height)
        assert 0 <= t["traversal_variant"] < rt.NUM_TRAVERSAL_VARIANTS and t["graph"] in (0, 1) and t["frame_ms"] > 0
        assert 1 <= t["extend_refill"] <= 64 and 1 <= t["shadow_refill"] <= 64
        a, _ = _render(gpu_ctx, ds, mode=rt.MODE_WAVEFRONT)
        assert np.array_equal(a.cpu().numpy(), ref.cpu().numpy())
    finally:
        gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, 0)
        gpu_ctx.set_option(rt.OPTION_GRAPH, 0)
        gpu_ctx.set_option(rt.OPTION_EXTEND_REFILL, 16)
        gpu_ctx.set_option(rt.OPTION_SHADOW_REFILL, 1)


def test_autotune_on_a_deep_bvh_leaves_no_error(gpu_ctx):
    """Shallow-stack variants can give up on rays of a deep BVH: the autotuner must
    drop them and leave the context usable, with the image unchanged."""
    ds = G.device_scene(gpu_ctx, "displaced_sphere")
    sc = ds.scene
    ubo = sc.ubo.copy()
    ubo["number_of_samples"] = ubo["total_number_of_samples"] = 2
    ref, _ = _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_WAVEFRONT)
    try:
        gpu_ctx.autotune(ds.tlas, ds.buffers, ubo, sc.width, sc.height)
        gpu_ctx.sync()
        a, st = _render(gpu_ctx, ds, ubo=ubo, mode=rt.MODE_WAVEFRONT, stats=True)
        assert np.array_equal(a.cpu().numpy(), ref.cpu().numpy())
    finally:
        gpu_ctx.set_option(rt.OPTION_TRAVERSAL_VARIANT, 0)
        gpu_ctx.set_option(rt.OPTION_GRAPH, 0)
        gpu_ctx.set_option(rt.OPTION_EXTEND_REFILL, 16)
        gpu_ctx.set_option(rt.OPTION_SHADOW_REFILL, 1)


@pytest.mark.parametrize("name", ["cornell_sphere", "textured_small"])
def test_persistent_megakernel_equals_one_lane_per_pixel(gpu_ctx, name):
    """RT_OPTION_MEGAKERNEL_PERSISTENT: lanes that fetch their next pixel give the
    same images (and display images) as one lane per pixel; HEAT frames too keep
    the accumulation."""
    if name == "textured_small":
        ds = G.device_scene(gpu_ctx, "textured_small_mk", lambda: S.textured_instances(
            width=100, height=60, spp=2, bounces=4, num_instances=60, tex_size=64, tri_range=(2000, 20000)))
    else:
        ds = G.device_scene(gpu_ctx, name)
    sc = ds.scene
    out = []
    for persistent in (0, 1):
        gpu_ctx.set_option(rt.OPTION_MEGAKERNEL_PERSISTENT, persistent)
        acc = rt.new_accumulation(sc.width, sc.height)
        disp = torch.zeros((sc.height, sc.width, 4), dtype=torch.uint8, device="cuda")
        st = rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=rt.MODE_MEGAKERNEL,
                           stats=True, out=disp)
        acc2 = rt.new_accumulation(sc.width, sc.height)
        rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc2, sc.width, sc.height, mode=rt.MODE_MEGAKERNEL)
        gpu_ctx.sync()
        assert torch.equal(acc, acc2)
        out.append((acc.cpu().numpy(), disp.cpu().numpy(), st["rays"]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


@pytest.mark.parametrize("name", ["cornell_sphere", "textured_small"])
@pytest.mark.parametrize("streams", [1, 3])
def test_finishing_launch_over_most_of_the_frame_equals_one_lane_per_pixel(gpu_ctx, name, streams):
    """A finishing threshold above the pixel count hands each slot set to k_wf_finish
    after its first 8 iterations, so its stepped lanes take paths with and without a
    pending shadow ray and retiring pixels, in opaque and alpha-tested scenes: the
    image, its display image and the ray counts equal those of one lane per pixel
    running each traversal to its end."""
    if name == "textured_small":
        ds = G.device_scene(gpu_ctx, "textured_small_mk", lambda: S.textured_instances(
            width=100, height=60, spp=2, bounces=4, num_instances=60, tex_size=64, tri_range=(2000, 20000)))
    else:
        ds = G.device_scene(gpu_ctx, name)
    sc = ds.scene
    out = []
    for mode in (rt.MODE_MEGAKERNEL, rt.MODE_WAVEFRONT):
        gpu_ctx.set_option(rt.OPTION_MEGAKERNEL_PERSISTENT, 0)
        gpu_ctx.set_option(rt.OPTION_WAVEFRONT_STREAMS, streams)
        gpu_ctx.set_option(rt.OPTION_WAVEFRONT_FINISH, 1 << 30)
        acc = rt.new_accumulation(sc.width, sc.height)
        disp = torch.zeros((sc.height, sc.width, 4), dtype=torch.uint8, device="cuda")
        st = rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True,
                           out=disp)
        gpu_ctx.sync()
        out.append((acc.cpu().numpy(), disp.cpu().numpy(), st))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    for k in ("rays", "primary_rays", "extension_rays", "shadow_rays", "shade_events"):
        assert out[0][2][k] == out[1][2][k], k
