"""Many-scene batches (rt_trace_rays_batch): every frame of a batch - different
scenes, resolutions, UBOs, opaque and alpha-masked - comes out bit for bit as its
own rt_trace_rays gives it, in both modes; counters add up."""
import numpy as np
import torch
import pytest

from rustracer_amd import rt
from rustracer_amd import scenes as S
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu


def _scenes():
    out = []
    for seed, side in zip(range(5), (40, 64, 96, 33, 128)):
        sc = S.batch_scene(seed, spp=3, bounces=5, tri_range=(500, 40_000), resolutions=(side,))
        out.append(sc)
    out.append(S.textured_instances(width=72, height=40, spp=2, bounces=6, num_instances=60, tex_size=64,
                                    tri_range=(500, 5000)))
    out.append(S.cornell_sphere(width=50, height=30, spp=2, bounces=0))  # no bounces: megakernel reset path
    return out


@pytest.mark.parametrize("mode,graph", [(rt.MODE_WAVEFRONT, 0), (rt.MODE_WAVEFRONT, 1), (rt.MODE_MEGAKERNEL, 0)])
def test_batch_frames_equal_single_frames(gpu_ctx, mode, graph):
    dss = [G.device_scene(gpu_ctx, f"batch_case_{k}", (lambda s=s: s)) for k, s in enumerate(_scenes())]
    singles, single_rays = [], 0
    for ds in dss:
        sc = ds.scene
        acc = rt.new_accumulation(sc.width, sc.height)
        st = rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height, mode=mode, stats=True)
        gpu_ctx.sync()
        singles.append(acc.cpu().numpy())
        single_rays += st["rays"]
    accs = [rt.new_accumulation(ds.scene.width, ds.scene.height).fill_(7.0) for ds in dss]
    items = [(ds.tlas, ds.buffers, ds.scene.ubo, a, ds.scene.width, ds.scene.height) for ds, a in zip(dss, accs)]
    gpu_ctx.set_option(rt.OPTION_GRAPH, graph)
    try:
        st = rt.trace_rays_batch(gpu_ctx, items, mode=mode, stats=True)
    finally:
        gpu_ctx.set_option(rt.OPTION_GRAPH, 0)
    for k, (a, ref) in enumerate(zip(accs, singles)):
        assert np.array_equal(a.cpu().numpy(), ref), k
    assert st["rays"] == single_rays
    # a second batch call reuses the state and gives the same images
    accs2 = [rt.new_accumulation(ds.scene.width, ds.scene.height) for ds in dss]
    rt.trace_rays_batch(gpu_ctx, [(i[0], i[1], i[2], a, i[4], i[5]) for i, a in zip(items, accs2)], mode=mode)
    for a, ref in zip(accs2, singles):
        assert np.array_equal(a.cpu().numpy(), ref)


def test_batch_argument_errors(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_sphere")
    sc = ds.scene
    acc = rt.new_accumulation(sc.width, sc.height)
    bad = rt.TopAS(gpu_ctx, 987654, None)
    with pytest.raises(rt.RtError):
        rt.trace_rays_batch(gpu_ctx, [(ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height),
                                      (bad, ds.buffers, sc.ubo, acc, sc.width, sc.height)])
    assert rt.trace_rays_batch(gpu_ctx, []) is None


def test_full_batch_workload_matches_oracle(gpu_ctx):
    """The many-scene workload at its full size: 64 scenes (seeds 0..63, 1k-500k
    triangles, 128^2-1024^2, 16 spp, depth 6) in one rt_trace_rays_batch with graph
    replay on.  Every frame against the FP64 oracle on a stratified subsample (every
    (side/32)-th pixel of every (side/32)-th row, all 16 samples: 1024 pixels per
    frame), relRMSE <= 1e-3 per frame; the smallest and largest frames also equal
    their own single-frame traces bit for bit."""
    from oracle import binding as B
    scs = S.batch_scenes()
    dss = [G.DeviceScene(gpu_ctx, sc) for sc in scs]
    accs = [rt.new_accumulation(sc.width, sc.height) for sc in scs]
    items = [(ds.tlas, ds.buffers, ds.scene.ubo, a, ds.scene.width, ds.scene.height) for ds, a in zip(dss, accs)]
    gpu_ctx.set_option(rt.OPTION_GRAPH, 1)
    rt.trace_rays_batch(gpu_ctx, items, mode=rt.MODE_WAVEFRONT)  # the timed kernels, graph replay
    gpu_ctx.sync()  # raises if a traversal gave up
    worst = 0.0
    for k, (sc, a) in enumerate(zip(scs, accs)):
        s = max(1, sc.width // 32)
        ys, xs = np.meshgrid(np.arange(s // 2, sc.height, s), np.arange(s // 2, sc.width, s), indexing="ij")
        xy = np.stack([xs.reshape(-1), ys.reshape(-1)], 1)
        o = B.OracleScene(sc)
        ref, _ = o.render_pixels(xy)
        o.close()
        g = a.cpu().numpy().astype(np.float64)[xy[:, 1], xy[:, 0], :3]
        rel = G.rel_rmse(g, ref[:, :3])
        worst = max(worst, rel)
        assert rel <= 1e-3, (k, sc.width, sc.num_triangles, rel)
    print("64-scene batch: worst per-frame relRMSE vs the FP64 oracle", worst)
    sizes = [sc.width * sc.height for sc in scs]
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        ds = dss[k]
        one = rt.new_accumulation(ds.scene.width, ds.scene.height)
        rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, ds.scene.ubo, one, ds.scene.width, ds.scene.height,
                      mode=rt.MODE_WAVEFRONT)
        gpu_ctx.sync()
        assert torch.equal(one, accs[k]), k
