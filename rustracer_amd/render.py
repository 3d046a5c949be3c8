"""Renders one of the named synthetic workloads on the GPU and prints one JSON line
with its frame time and Mrays/s.  It is the command-line counterpart of the viewer's
render loop (crates/examples/gltf_viewer/src/main.rs:244-270: one trace_rays per
frame) for the generated scenes in rustracer_amd.scenes.WORKLOADS.

    python -m rustracer_amd.render textured_instances --mode wavefront --frames 3

Ray counts come from one counting frame before the timed frames.  --autotune runs
rt_autotune first; --ppm writes the last frame's accumulation image, divided by the
sample count and clamped to [0, 1], as a binary PPM.  --adjoint also times the
path-replay adjoint of the frame (rt_trace_rays_adjoint) for a seeded random image
gradient, with gradients for every RGBA32F texture and the point lights'
intensities.
"""
import argparse
import json
import sys
import time


def main(argv=None):
    from rustracer_amd import scenes
    ap = argparse.ArgumentParser(prog="python -m rustracer_amd.render")
    ap.add_argument("workload", choices=sorted(scenes.WORKLOADS))
    ap.add_argument("--mode", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--ppm", default=None)
    ap.add_argument("--adjoint", action="store_true")
    args = ap.parse_args(argv)
    if args.frames < 1 or args.warmup < 0:
        ap.error("--frames must be at least 1 and --warmup at least 0")

    import numpy as np
    from rustracer_amd import rt

    sc = scenes.WORKLOADS[args.workload]()
    W, H = sc.width, sc.height
    ctx = rt.Context(0)
    buf = rt.Buffers(sc)
    blases, tlas = rt.create_as(ctx, sc, buf)
    mode = rt.MODE_WAVEFRONT if args.mode == "wavefront" else rt.MODE_MEGAKERNEL
    tuned = ctx.autotune(tlas, buf, sc.ubo, W, H) if args.autotune else None
    acc = rt.new_accumulation(W, H)
    st = rt.trace_rays(ctx, tlas, buf, sc.ubo, acc, W, H, mode=mode, stats=True)
    ctx.sync()
    for _ in range(args.warmup):
        rt.trace_rays(ctx, tlas, buf, sc.ubo, acc, W, H, mode=mode)
    ctx.sync()
    times = []
    for _ in range(args.frames):
        t0 = time.perf_counter()
        rt.trace_rays(ctx, tlas, buf, sc.ubo, acc, W, H, mode=mode)
        ctx.sync()
        times.append((time.perf_counter() - t0) * 1e3)
    ms = min(times)
    line = {"workload": args.workload, "mode": args.mode, "width": W, "height": H,
            "samples": int(sc.ubo["number_of_samples"]), "bounces": int(sc.ubo["number_of_bounces"]),
            "rays": int(st["rays"]), "ms_per_frame": round(ms, 3), "mrays_per_s": round(st["rays"] / ms / 1e3, 1),
            "autotune": tuned}
    if args.adjoint:
        line["adjoint"] = _adjoint(ctx, sc, buf, tlas, args.frames, args.warmup)
    print(json.dumps(line), flush=True)
    if args.ppm:
        img = acc.cpu().numpy()[..., :3] / float(sc.ubo["total_number_of_samples"])
        px = (np.clip(img, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)
        with open(args.ppm, "wb") as f:
            f.write(b"P6\n%d %d\n255\n" % (W, H))
            f.write(px.tobytes())
    del blases, tlas, buf
    return 0


def _adjoint(ctx, sc, buf, tlas, frames, warmup):
    import numpy as np
    import torch
    from rustracer_amd import layouts as L
    from rustracer_amd import rt
    W, H = sc.width, sc.height
    g = torch.from_numpy(np.random.default_rng(0).normal(size=(H, W, 4)).astype(np.float32)).cuda()
    grads = {k: torch.zeros(t.data.shape, dtype=torch.float32, device="cuda")
             for k, t in enumerate(sc.textures) if t.format == L.FMT_RGBA32F}
    gp = torch.zeros(max(1, len(sc.plights)), dtype=torch.float32, device="cuda")
    st = rt.trace_rays_adjoint(ctx, tlas, buf, sc.ubo, g, W, H, grads, stats=True, grad_point_intensity=gp)
    for _ in range(warmup):
        rt.trace_rays_adjoint(ctx, tlas, buf, sc.ubo, g, W, H, grads, grad_point_intensity=gp)
    ctx.sync()
    times = []
    for _ in range(frames):
        t0 = time.perf_counter()
        rt.trace_rays_adjoint(ctx, tlas, buf, sc.ubo, g, W, H, grads, grad_point_intensity=gp)
        ctx.sync()
        times.append((time.perf_counter() - t0) * 1e3)
    return {"ms": round(min(times), 3), "rays": int(st["rays"]), "grad_contributions": int(st["grad_taps"]),
            "grad_atomics": int(st["grad_atomics"])}


if __name__ == "__main__":
    sys.exit(main())
