"""Path-replay adjoint (rt_trace_rays_adjoint) against central finite differences
of the FP64 oracle's detached evaluation: the oracle renders with the sampling
(lobe choices, directions, densities, RIS light choice and weight, Russian roulette)
of the unperturbed parameters and the BRDF / light values of the perturbed ones,
which is the function whose derivative the adjoint computes.  Scene: the Cornell box
with RGBA32F base-colour and metallic-roughness parameter textures on the white
walls (restates the closest-hit shading of
crates/examples/gltf_viewer/shaders/RayTracing.rchit :136-416, NEE :338-350, the
miss shader RayTracing.rmiss :14-45 and the raygen loop of RayTracing.rgen :85-128).

Per-entry bar: |adjoint - FD| <= 1e-3 |FD| + 1e-6 max|FD| (the second term is the
noise floor of float32 gradient sums)."""
import numpy as np
import pytest
import torch

from oracle import binding as B
from rustracer_amd import diff, rt
from rustracer_amd import scenes as S
from tests import gpu_kit as G

pytestmark = pytest.mark.gpu
H_FD = 1e-4
REL, FLOOR = 1e-3, 1e-6


def _small():
    return S.cornell_params(width=48, height=48, spp=4, bounces=4, tex_size=12)


def _config4_reduced():
    # the differentiable workload's image and texture size, 2 samples per pixel
    return S.cornell_params(width=512, height=512, spp=2, bounces=4, tex_size=512)


def _adjoint(ctx, ds, g, stats=False):
    sc = ds.scene
    grads = {k: torch.zeros(sc.textures[k].data.shape, dtype=torch.float32, device="cuda") for k in (1, 2)}
    gp = torch.zeros(len(sc.plights), dtype=torch.float32, device="cuda")
    gd = torch.zeros(len(sc.dlights), dtype=torch.float32, device="cuda")
    st = rt.trace_rays_adjoint(ctx, ds.tlas, ds.buffers, sc.ubo, g, sc.width, sc.height, grads, stats=stats,
                               grad_point_intensity=gp, grad_directional_intensity=gd)
    ctx.sync()
    out = {k: v.cpu().numpy().astype(np.float64) for k, v in grads.items()}
    return out, gp.cpu().numpy().astype(np.float64), gd.cpu().numpy().astype(np.float64), st


def _check(a, fd, what):
    a, fd = np.asarray(a, np.float64), np.asarray(fd, np.float64)
    scale = np.abs(fd).max()
    err = np.abs(a - fd)
    bad = err > REL * np.abs(fd) + FLOOR * scale
    print(f"{what}: max |fd| {scale:.4g}, max abs err {err.max():.3g}, max rel err above the floor "
          f"{(err / np.maximum(np.abs(fd), FLOOR * scale)).max():.3g}, entries {len(fd)}")
    assert scale > 0 and not bad.any(), (np.nonzero(bad)[0], a[bad], fd[bad])


def _light_fd(o, sc, g64, kind):
    """Central FD of sum(g * image) in each light's eval intensity (detached)."""
    lights = sc.plights if kind == "point" else sc.dlights
    base = lights["intensity"].astype(np.float32)
    fd = np.zeros(len(lights))
    for i in range(len(lights)):
        vals, steps = [], []
        for sgn in (1, -1):
            e = base.copy()
            e[i] = np.float32(float(base[i]) + sgn * H_FD * max(1.0, abs(float(base[i]))))
            steps.append(float(e[i]) - float(base[i]))
            o.set_eval_lights(e, None) if kind == "point" else o.set_eval_lights(None, e)
            img, _ = o.render()
            vals.append(float((img[..., :3] * g64[..., :3]).sum()))
        fd[i] = (vals[0] - vals[1]) / (steps[0] - steps[1])
    o.set_eval_lights(None, None)
    return fd


def _random_entries(rng, size, n):
    """n random (texture, y, x, channel): base colour R, G, B and roughness, metallic."""
    out = []
    for _ in range(n):
        k = int(rng.integers(1, 3))
        ch = int(rng.integers(0, 3)) if k == 1 else int(rng.integers(1, 3))
        out.append((k, int(rng.integers(size)), int(rng.integers(size)), ch))
    return out


def _texel_fd(o, base, g64, entries, xy=None, footprints=None):
    """Central FD of sum(g * image) in each entry's texel channel (detached).  With
    footprints, only the pixels that read the texel are rendered again (the others
    do not depend on it)."""
    fd = np.zeros(len(entries))
    for i, (k, ty, tx, ch) in enumerate(entries):
        if footprints is not None and len(footprints[i]) == 0:
            continue
        vals, steps = [], []
        for sgn in (1, -1):
            ev = [b.copy() for b in base]
            v0 = ev[k][ty, tx, ch]
            ev[k][ty, tx, ch] = np.float32(float(v0) + sgn * H_FD)  # the texture is fp32: use the step it holds
            steps.append(float(ev[k][ty, tx, ch]) - float(v0))
            o.set_eval_textures(ev)
            if footprints is None:
                img, _ = o.render()
                vals.append(float((img[..., :3] * g64[..., :3]).sum()))
            else:
                p = xy[footprints[i]]
                img, _ = o.render_pixels(p)
                vals.append(float((img[:, :3] * g64[p[:, 1], p[:, 0], :3]).sum()))
        fd[i] = (vals[0] - vals[1]) / (steps[0] - steps[1])
    o.set_eval_textures(None)
    return fd


def test_adjoint_matches_finite_differences_small(gpu_ctx):
    """48x48, 4 spp, 12x12 textures: 256 texel channels, the point light and the
    directional light, every pixel rendered for each FD."""
    ds = G.device_scene(gpu_ctx, "cornell_params_small", _small)
    sc = ds.scene
    rng = np.random.default_rng(3)
    g = rng.normal(size=(sc.height, sc.width, 4)).astype(np.float32)
    g[..., 3] = 0
    adj, gp, gd, _ = _adjoint(gpu_ctx, ds, torch.from_numpy(g).cuda())
    # forward images agree (no path flips between GPU and oracle on this scene)
    acc = rt.new_accumulation(sc.width, sc.height)
    rt.trace_rays(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, acc, sc.width, sc.height)
    gpu_ctx.sync()
    o = B.OracleScene(sc)
    ref, _ = o.render()
    assert G.rel_rmse(acc.cpu().numpy()[..., :3].astype(np.float64), ref[..., :3]) < 1e-6
    base = [t.data.astype(np.float32) for t in sc.textures]
    g64 = g.astype(np.float64)
    entries = _random_entries(rng, sc.textures[1].data.shape[0], 256)
    fd = _texel_fd(o, base, g64, entries)
    a = np.array([adj[k][ty, tx, ch] for k, ty, tx, ch in entries])
    assert (np.abs(fd) > 1e-6 * np.abs(fd).max()).sum() > 200
    _check(a, fd, "texels")
    _check(gp, _light_fd(o, sc, g64, "point"), "point-light intensity")
    # the default directional light (intensity 0) still has a gradient: escaped rays
    _check(gd, _light_fd(o, sc, g64, "directional"), "directional-light intensity")
    # channels the adjoint does not own stay zero (base alpha; metallic-roughness R, A)
    assert np.all(adj[1][..., 3] == 0) and np.all(adj[2][..., 0] == 0) and np.all(adj[2][..., 3] == 0)


def test_adjoint_matches_finite_differences_at_full_texture_size(gpu_ctx):
    """The differentiable workload's sizes (512x512 image, 512x512 textures, depth 4)
    at 2 spp: 256 texel channels (half with the largest adjoint gradients, half
    random) and the light intensities.  Each texel FD renders again only the pixels
    whose detached evaluation reads that texel (orc_texel_footprints)."""
    ds = G.device_scene(gpu_ctx, "cornell_params_512_2spp", _config4_reduced)
    sc = ds.scene
    rng = np.random.default_rng(11)
    g = rng.normal(size=(sc.height, sc.width, 4)).astype(np.float32)
    g[..., 3] = 0
    adj, gp, gd, st = _adjoint(gpu_ctx, ds, torch.from_numpy(g).cuda(), stats=True)
    print("adjoint counters", {k: st[k] for k in ("rays", "grad_taps", "grad_atomics")})
    assert 0 < st["grad_atomics"] < st["grad_taps"] * 3
    size = sc.textures[1].data.shape[0]
    # entries: 128 of the largest |adjoint| channels, 128 random
    flat = []
    for k, chans in ((1, (0, 1, 2)), (2, (1, 2))):
        for ch in chans:
            v = np.abs(adj[k][..., ch]).reshape(-1)
            top = np.argpartition(v, -64)[-64:]
            flat += [(float(v[t]), k, int(t // size), int(t % size), ch) for t in top]
    flat.sort(reverse=True)
    entries = [(k, ty, tx, ch) for _, k, ty, tx, ch in flat[:128]] + _random_entries(rng, size, 128)
    o = B.OracleScene(sc)
    base = [t.data.astype(np.float32) for t in sc.textures]
    o.set_eval_textures(base)
    ys, xs = np.meshgrid(np.arange(sc.height), np.arange(sc.width), indexing="ij")
    xy = np.stack([xs.reshape(-1), ys.reshape(-1)], 1)
    fps = o.texel_footprints(xy, [(k, ty * size + tx) for k, ty, tx, _ in entries])
    g64 = g.astype(np.float64)
    fd = _texel_fd(o, base, g64, entries, xy, fps)
    a = np.array([adj[k][ty, tx, ch] for k, ty, tx, ch in entries])
    print("entries with a non-empty footprint:", sum(len(f) > 0 for f in fps))
    _check(a, fd, "texels (512^2)")
    _check(gp, _light_fd(o, sc, g64, "point"), "point-light intensity (512^2)")


def test_adjoint_rejects_gradient_buffers_for_non_float_textures(gpu_ctx):
    ds = G.device_scene(gpu_ctx, "cornell_params_small", _small)
    sc = ds.scene
    g = torch.zeros((sc.height, sc.width, 4), dtype=torch.float32, device="cuda")
    bad = {0: torch.zeros((1, 1, 4), dtype=torch.float32, device="cuda")}  # slot 0 is RGBA8
    with pytest.raises(rt.RtError):
        rt.trace_rays_adjoint(gpu_ctx, ds.tlas, ds.buffers, sc.ubo, g, sc.width, sc.height, bad)


def test_pre_reduction_does_not_change_the_gradient(gpu_ctx):
    """The adjoint's timed and counting kernels (same per-wave sums) agree, and a
    zero pixel gradient gives zero everywhere."""
    ds = G.device_scene(gpu_ctx, "cornell_params_small", _small)
    sc = ds.scene
    rng = np.random.default_rng(8)
    g = torch.from_numpy(rng.normal(size=(sc.height, sc.width, 4)).astype(np.float32)).cuda()
    a0, p0, d0, _ = _adjoint(gpu_ctx, ds, g)
    a1, p1, d1, st = _adjoint(gpu_ctx, ds, g, stats=True)
    for k in a0:
        assert np.allclose(a0[k], a1[k], rtol=1e-5, atol=1e-6 * np.abs(a0[k]).max())
    assert np.allclose(p0, p1, rtol=1e-5) and np.allclose(d0, d1, rtol=1e-5, atol=1e-9)
    z, zp, zd, zs = _adjoint(gpu_ctx, ds, torch.zeros_like(g), stats=True)
    assert all(np.all(v == 0) for v in z.values()) and np.all(zp == 0) and np.all(zd == 0)
    assert zs["rays"] == 0


def test_autograd_function_and_an_optimisation_step(gpu_ctx):
    """loss = sum((accum - target)^2): torch.autograd gives the adjoint's gradient for
    g = 2 (accum - target), textures and light intensity, and gradient steps on the
    parameters lower the loss."""
    sc = S.cornell_params(width=64, height=64, spp=4, bounces=4, tex_size=16)
    tgt_scene = S.cornell_params(width=64, height=64, spp=4, bounces=4, tex_size=16, target=True)
    tds = G.DeviceScene(gpu_ctx, tgt_scene)
    target = rt.new_accumulation(64, 64)
    rt.trace_rays(gpu_ctx, tds.tlas, tds.buffers, tgt_scene.ubo, target, 64, 64)
    gpu_ctx.sync()
    dscene = diff.DifferentiableScene(gpu_ctx, sc, [1, 2], lights=True)
    img = dscene.render()
    # the megakernel forward gives the same image
    mscene = diff.DifferentiableScene(gpu_ctx, sc, [1, 2], lights=True, mode=rt.MODE_MEGAKERNEL)
    assert torch.equal(mscene.render().detach(), img.detach())
    loss = ((img - target) ** 2).sum()
    loss.backward()
    g = (2 * (img - target)).detach().contiguous()
    ref = {k: torch.zeros_like(p) for k, p in zip([1, 2], dscene.params)}
    gi = torch.zeros_like(dscene.intensity)
    rt.trace_rays_adjoint(gpu_ctx, dscene.tlas, dscene.buffers, sc.ubo, g, 64, 64, ref, grad_point_intensity=gi)
    gpu_ctx.sync()
    for k, p in zip([1, 2], dscene.params):
        assert torch.allclose(p.grad, ref[k], rtol=1e-5, atol=1e-6 * float(ref[k].abs().max()))
    assert torch.allclose(dscene.intensity.grad, gi, rtol=1e-5)
    losses = [float(loss.detach())]
    opt = torch.optim.Adam(dscene.params, lr=0.02)
    for _ in range(6):
        opt.zero_grad()
        img = dscene.render()
        loss = ((img - target) ** 2).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            for p in dscene.params:
                p.clamp_(0.02, 1.0)
        losses.append(float(loss.detach()))
    print("losses", losses)
    assert losses[-1] < 0.9 * losses[0]


def test_two_views_before_one_backward(gpu_ctx):
    """Two forwards (two cameras) and one backward of the summed loss: each view's
    backward replays its own camera, so the gradients equal the sum of the two
    views' adjoints taken one at a time."""
    sc = S.cornell_params(width=40, height=40, spp=2, bounces=4, tex_size=10)
    dscene = diff.DifferentiableScene(gpu_ctx, sc, [1, 2], lights=True)
    ubo_b = sc.ubo.copy()
    view = S.look_at_rh((0.6, 1.0, 3.2), (0.0, 0.9, 0.0), (0, 1, 0))
    ubo_b["model_view"] = np.asarray(view, np.float64).T.reshape(16).astype(np.float32)
    ubo_b["model_view_inverse"] = np.linalg.inv(view).T.reshape(16).astype(np.float32)
    rng = np.random.default_rng(4)
    wa = torch.from_numpy(rng.normal(size=(40, 40, 4)).astype(np.float32)).cuda()
    wb = torch.from_numpy(rng.normal(size=(40, 40, 4)).astype(np.float32)).cuda()
    loss = (dscene.render() * wa).sum() + (dscene.render(ubo_b) * wb).sum()
    loss.backward()
    want = {k: torch.zeros_like(p) for k, p in zip([1, 2], dscene.params)}
    gi = torch.zeros_like(dscene.intensity)
    for u, w in ((sc.ubo, wa), (ubo_b, wb)):
        rt.trace_rays_adjoint(gpu_ctx, dscene.tlas, dscene.buffers, u, w.contiguous(), 40, 40, want,
                              grad_point_intensity=gi)
    gpu_ctx.sync()
    for k, p in zip([1, 2], dscene.params):
        assert float(want[k].abs().max()) > 0
        assert torch.allclose(p.grad, want[k], rtol=1e-5, atol=1e-6 * float(want[k].abs().max()))
    assert torch.allclose(dscene.intensity.grad, gi, rtol=1e-5)
