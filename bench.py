"""Benchmark: forward + backward pixel-samples/s and Adam steps/s on one MI355X.

Workload (the headline configuration): 1024x1024 canvas, 1000 closed cubic
paths of 3-8 segments (centre uniform, control points centre + U(-64,64)^2),
solid RGBA fills with alpha in [0.3, 1), white background, 2x2 jittered
samples, box filter r = 0.5, non-prefiltered (boundary sampling on), seed 0.
d_render = 2 (image - target), the target rendered from the same scene with
path points jittered by N(0, 2 px).

One step = weights + forward + resolve + weights + interior backward +
boundary sampling over all 4,194,304 pixel-samples, scene resident in HBM.
`value` = pixel-samples / s of that step.  Adam steps/s (pack on the host,
upload, forward, loss gradient, backward, Adam update) is reported beside it.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diffvg-rs_amd"))

METRIC = "fwd+bwd pixel-samples/s and Adam steps/s @1024², 1k paths, grad err ≤1e-3"
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--adam-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from pydvg import scenes
    from pydvg.optim import AdamLoop
    from pydvg.renderer import DeviceScene, PreparedScene, Renderer

    torch.cuda.init()
    dev = torch.device("cuda:0")
    # one stream (default priority) shared by torch and libdvg
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    r = Renderer(0)
    r.set_stream(stream.cuda_stream)

    sc, opts = scenes.headline(seed=0)
    W, H = sc.width, sc.height
    spp = opts.samples_x * opts.samples_y
    n_samples = W * H * spp
    log(f"bench: headline {W}x{H}, {len(sc.shapes)} paths, spp {spp}")
    prep = PreparedScene(sc, opts)
    ds = DeviceScene(r, prep)
    tgt_prep = PreparedScene(scenes.jittered(sc, 2.0, seed=1), opts)
    tgt_ds = DeviceScene(r, tgt_prep)
    img = torch.empty(W * H * 4, dtype=torch.float32, device=dev)
    tgt = torch.empty_like(img)
    tgt_ds.render(tgt)
    bufs, sizes = ds.grad_buffers(torch)
    d_render = torch.empty_like(img)

    def step():
        ds.render(img)
        torch.sub(img, tgt, out=d_render)
        d_render.mul_(2.0)
        ds.render_backward(d_render, bufs, sizes)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log("bench: warmup done")
    # per-phase timing with events on the stream the kernels run on
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ev[0].record(stream)
        ds.render(img)
        torch.sub(img, tgt, out=d_render)
        d_render.mul_(2.0)
        ev[1].record(stream)
        ds.render_backward(d_render, bufs, sizes)
        ev[2].record(stream)
        ev[2].synchronize()
        fwd_ms += ev[0].elapsed_time(ev[1])
        bwd_ms += ev[1].elapsed_time(ev[2])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1000.0 * wall / args.steps
    fwd_ms /= args.steps
    bwd_ms /= args.steps
    value = n_samples / (wall / args.steps)
    log(f"bench: {ms_step:.3f} ms/step (forward {fwd_ms:.3f} ms, backward {bwd_ms:.3f} ms)")

    # Adam steps/s: host packing + upload + forward + loss + backward + update
    loop = AdamLoop(r, sc, opts, tgt, lr_points=0.5, lr_colors=1e-2)
    loop.step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    losses = [loop.step() for _ in range(args.adam_steps)]
    adam_s = (time.perf_counter() - t1) / args.adam_steps
    log(f"bench: Adam {1.0 / adam_s:.2f} steps/s, losses {losses}")

    roofline = None
    flops_path = os.path.join(ROOT, "profiles", "flops_headline.json")
    if os.path.exists(flops_path):
        with open(flops_path) as f:
            fl = json.load(f)
        # dominant kernel: interior backward + boundary = the backward call
        per_launch = fl["bwd_flops_per_sample"] * n_samples
        achieved = per_launch / (bwd_ms * 1e-3) / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": fl.get("bwd_hbm_bytes"),
                    "kernel": "render_backward (k_backward_interior + k_boundary)",
                    "note": "fp32 VALU peak; algorithmic flops counted by the oracle"}

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        from pydvg.scene import SceneDesc
        dr_host = d_render.cpu().numpy()
        cpu = oracle_py.time_slice(SceneDesc(sc), opts, dr_host, args.cpu_seconds)
        cpu = {"value": cpu["samples_per_s"], "unit": "pixel-samples/s", "cores": 1, "kind": "port",
               "sample": cpu["sample"]}

    out = {
        "metric": METRIC, "value": value, "unit": "pixel-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "headline: 1024x1024, 1000 closed cubic paths (3-8 segments), 2x2 jittered spp, "
                               "box r=0.5, non-prefiltered fwd + interior bwd + boundary sampling",
                   "global_batch": 1, "samples_per_step": n_samples, "parallelism": "single-gpu"},
        "forward_ms": fwd_ms, "backward_ms": bwd_ms, "adam_steps_per_s": 1.0 / adam_s,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
