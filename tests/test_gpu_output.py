"""The display image of rt_trace_rays (RayTracing.rgen:132-166: tone mapping,
HEAT / DISTANCE mappings, 8-bit UNORM store) against the oracle's restatement of
the same stage, applied to the GPU's own accumulation image.  The 8-bit values
may differ by one step where powf of the two libraries rounds differently at a
rounding boundary."""
import numpy as np
import pytest
import torch

from oracle import binding as B
from rustracer_amd import rt
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu


def _trace(ctx, ds, ubo, mode=rt.MODE_MEGAKERNEL):
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    out = torch.zeros((sc.height, sc.width, 4), dtype=torch.uint8, device="cuda")
    rt.trace_rays(ctx, ds.tlas, ds.buffers, ubo, acc, sc.width, sc.height, mode=mode, out=out)
    ctx.sync()
    return acc.cpu().numpy(), out.cpu().numpy()


def _close(gpu, ref, frac=1e-3):
    d = np.abs(gpu.astype(np.int32) - ref.astype(np.int32))
    assert d.max() <= 1, f"max difference {d.max()}"
    assert np.count_nonzero(d) <= frac * d.size, f"{np.count_nonzero(d)} values differ by one step"


@pytest.mark.parametrize("tone", [0, 1, 2, 3, 4])
def test_tone_mapped_image_matches_oracle_and_modes_agree(gpu_ctx, tone):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    ubo = ds.scene.ubo.copy()
    ubo["tone_mapping_mode"] = tone
    acc, out = _trace(gpu_ctx, ds, ubo)
    assert np.all(out[..., 3] == 255)
    ref = B.output_image(acc, ubo)
    _close(out, ref)
    assert out[..., :3].max() > 0
    acc_w, out_w = _trace(gpu_ctx, ds, ubo, rt.MODE_WAVEFRONT)
    assert np.array_equal(acc_w, acc) and np.array_equal(out_w, out)


def test_accumulating_launch_tone_maps_the_running_sum(gpu_ctx):
    """Second launch with NumberOfSamples != TotalNumberOfSamples: the image shows
    (acc_prev + sum) / TotalNumberOfSamples."""
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    ubo = sc.ubo.copy()
    ubo["number_of_samples"] = 2
    ubo["total_number_of_samples"] = 2
    acc = rt.new_accumulation(sc.width, sc.height)
    out = torch.zeros((sc.height, sc.width, 4), dtype=torch.uint8, device="cuda")
    rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, ubo, acc, sc.width, sc.height, out=out)
    ubo["total_number_of_samples"] = 4
    ubo["frame_count"] = int(ubo["frame_count"]) + 1
    rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, ubo, acc, sc.width, sc.height, mode=rt.MODE_WAVEFRONT, out=out)
    gpu_ctx.sync()
    _close(out.cpu().numpy(), B.output_image(acc.cpu().numpy(), ubo))


def test_bgra_option_swaps_red_and_blue(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    ubo = ds.scene.ubo.copy()
    _, rgba = _trace(gpu_ctx, ds, ubo)
    gpu_ctx.set_option(rt.OPTION_OUTPUT_BGRA, 1)
    try:
        _, bgra = _trace(gpu_ctx, ds, ubo, rt.MODE_WAVEFRONT)
    finally:
        gpu_ctx.set_option(rt.OPTION_OUTPUT_BGRA, 0)
    assert np.array_equal(bgra[..., [2, 1, 0, 3]], rgba)


def test_distance_mapping_uses_the_last_trace_t(gpu_ctx):
    """DISTANCE: grey min((S - max(t, tMin)) / S, 1) of the payload t after the pixel's
    last trace; the oracle renders a pixel subsample and reports that t."""
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    ubo = sc.ubo.copy()
    ubo["mapping"] = 4  # DISTANCE
    ubo["heatmap_scale"] = 8.0
    acc, out = _trace(gpu_ctx, ds, ubo)
    _, out_w = _trace(gpu_ctx, ds, ubo, rt.MODE_WAVEFRONT)
    assert np.array_equal(out_w, out)
    ys, xs = np.meshgrid(np.arange(3, sc.height, 8), np.arange(3, sc.width, 8), indexing="ij")
    xy = np.stack([xs.reshape(-1), ys.reshape(-1)], 1)
    o, _ = B.OracleScene(sc).render_pixels(xy, ubo=ubo)
    ref = B.output_image(acc[xy[:, 1], xy[:, 0]], ubo, last_t=o[:, 3].astype(np.float32))
    got = out[xy[:, 1], xy[:, 0]]
    assert len(np.unique(got[:, 0])) > 10  # the grey level varies with distance
    # a path whose last bounce differs between fp32 and fp64 decisions may end elsewhere
    d = np.abs(got.astype(np.int32) - ref.astype(np.int32)).max(axis=1)
    assert np.count_nonzero(d > 1) <= 0.002 * len(d)


def test_heat_mapping_draws_from_the_colour_ramp(gpu_ctx):
    """HEAT: each pixel is the ramp's colour at its clock ticks; every output lies on
    the oracle's ramp.  A wavefront-mode call takes the megakernel form for such a
    frame (the same accumulation)."""
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    ubo = sc.ubo.copy()
    ubo["mapping"] = 1  # HEAT
    ubo["heatmap_scale"] = 0.5
    _, out = _trace(gpu_ctx, ds, ubo)
    ramp = np.array([np.rint(B.heatmap(t) * 255) for t in np.linspace(0, 1, 20001)], np.int32)
    px = np.unique(out.reshape(-1, 4)[:, :3].astype(np.int32), axis=0)
    dist = np.abs(px[:, None, :] - ramp[None, :, :]).max(axis=2).min(axis=1)
    assert dist.max() <= 1
    acc = rt.new_accumulation(sc.width, sc.height)
    o = torch.zeros((sc.height, sc.width, 4), dtype=torch.uint8, device="cuda")
    rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, ubo, acc, sc.width, sc.height, mode=rt.MODE_WAVEFRONT, out=o)
    ref = rt.new_accumulation(sc.width, sc.height)
    rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, ubo, ref, sc.width, sc.height, mode=rt.MODE_WAVEFRONT)
    gpu_ctx.sync()
    assert torch.equal(acc, ref)
    pw = np.unique(o.cpu().numpy().reshape(-1, 4)[:, :3].astype(np.int32), axis=0)
    assert np.abs(pw[:, None, :] - ramp[None, :, :]).max(axis=2).min(axis=1).max() <= 1
