"""Renders one of the named synthetic workloads on the GPU and prints one JSON line
with its frame time and Mrays/s.  It is the command-line counterpart of the viewer's
render loop (crates/examples/gltf_viewer/src/main.rs:244-270: one trace_rays per
frame) for the generated scenes in rustracer_amd.scenes.WORKLOADS.

    python -m rustracer_amd.render textured_instances --mode wavefront --frames 3

Ray counts come from one counting frame before the timed frames.  --autotune runs
rt_autotune first; --ppm writes the last frame's accumulation image, divided by the
sample count and clamped to [0, 1], as a binary PPM.
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
    print(json.dumps(line), flush=True)
    if args.ppm:
        img = acc.cpu().numpy()[..., :3] / float(sc.ubo["total_number_of_samples"])
        px = (np.clip(img, 0.0, 1.0) * 255.0 + 0.5).astype(np.uint8)
        with open(args.ppm, "wb") as f:
            f.write(b"P6\n%d %d\n255\n" % (W, H))
            f.write(px.tobytes())
    del blases, tlas, buf
    return 0


if __name__ == "__main__":
    sys.exit(main())
