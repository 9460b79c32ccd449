"""Benchmark: forward + backward pixel-samples/s and Adam steps/s on one MI355X.

Default workload (the headline configuration): 1024x1024 canvas, 1000 closed
cubic paths of 3-8 segments (centre uniform, control points centre +
U(-64,64)^2), solid RGBA fills with alpha in [0.3, 1), white background, 2x2
jittered samples, box filter r = 0.5, non-prefiltered (boundary sampling on),
seed 0.  d_render = 2 (image - target), the target rendered from the same
scene with path points jittered by N(0, 2 px).

One step = weights + forward + resolve + weights + interior backward +
boundary sampling over all 4,194,304 pixel-samples, scene resident in HBM.
`value` = pixel-samples / s of that step.  Adam steps/s (re-pack on the
device, forward, loss gradient, backward, Adam update, all device resident)
is reported beside it; for c5 the Adam loop runs against the procedural
Gaussian target and the line adds its loss curve and device memory in use.

--workload picks another configuration for the same measurement (one JSON
line each): c1 (256x256, 64 cubic paths, 4x4 spp), c2 (512x512, 512 mixed
shapes, tent filter), c3 (1024x1024 prefiltered: 2000 stroked cubic paths +
500 polygons, 1 spp), c5 (2048x2048, 10k cubic paths, 2x2 spp).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diffvg-rs_amd"))

METRIC = "fwd+bwd pixel-samples/s and Adam steps/s @1024², 1k paths, grad err ≤1e-3"
# name -> (scenes function, description)
WORKLOADS = {
    "headline": ("headline", "headline: 1024x1024, 1000 closed cubic paths (3-8 segments), 2x2 jittered spp, "
                             "box r=0.5, non-prefiltered fwd + interior bwd + boundary sampling"),
    "c1": ("config1", "256x256, 64 closed cubic paths, 4x4 jittered spp, box r=0.5, non-prefiltered"),
    "c2": ("config2", "512x512, 512 mixed shapes (circles/ellipses/rects/paths, 30% stroked, gradients), "
                      "2x2 spp, tent r=1, non-prefiltered"),
    "c3": ("config3", "1024x1024 prefiltered (analytic coverage), 2000 stroked open cubic paths + 500 filled "
                      "polygons, 1 spp"),
    "c5": ("config5", "2048x2048, 10000 closed cubic paths, 2x2 jittered spp, box r=0.5, non-prefiltered"),
    "c4": ("config4", "64 scenes of uneven size (16-4096 mixed shapes log-uniform, 128^2-1024^2), 2x2 spp, "
                      "batched fwd + L2 loss + bwd in one persistent launch per phase, HIP graph, autotuned shapes"),
}
FP32_PEAK_TFLOPS = 157.3


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--adam-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", default="headline", choices=sorted(WORKLOADS))
    ap.add_argument("--batch-scenes", type=int, default=4,
                    help="headline-size scenes in the batched measurement of the default line (0: none)")
    ap.add_argument("--unchecked-boundary", action="store_true",
                    help="the unchecked f32 boundary pass instead of the default checked one (single-scene workloads)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import numpy as np
    import torch
    from pydvg import scenes
    from pydvg.optim import DeviceAdamLoop
    from pydvg.renderer import DeviceScene, PreparedScene, Renderer

    torch.cuda.init()
    dev = torch.device("cuda:0")
    # one stream (default priority) shared by torch and libdvg
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    r = Renderer(0)
    r.set_stream(stream.cuda_stream)
    if args.unchecked_boundary:
        r.set_exact_boundary(False)

    if args.workload == "c4":
        return bench_batch(args, r, stream, dev)
    fn_name, workload_desc = WORKLOADS[args.workload]
    headline = args.workload == "headline"
    sc, opts = getattr(scenes, fn_name)(seed=0)
    W, H = sc.width, sc.height
    spp = opts.samples_x * opts.samples_y
    n_samples = W * H * spp
    log(f"bench: {args.workload} {W}x{H}, {len(sc.shapes)} shapes, spp {spp}")
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

    # a plain copy of known size (64 MiB read + 64 MiB written), the scale
    # reference for FETCH_SIZE / WRITE_SIZE counter runs
    cal_src = torch.ones(16 * 1024 * 1024, dtype=torch.float32, device=dev)
    cal_dst = torch.empty_like(cal_src)
    cal_dst.copy_(cal_src)
    torch.cuda.synchronize()

    # Adam steps/s: device re-packing + forward + loss + backward + Adam update,
    # parameters and optimiser state resident on the device
    adam_s = None
    adam = None
    if (headline or args.workload == "c5") and args.adam_steps > 0:
        if args.workload == "c5":
            tgt_adam = torch.tensor(scenes.gaussian_target(W, H, 256, seed=0), device=dev)
        else:
            tgt_adam = tgt
        # the step captured in a HIP graph (its first two steps are the eager
        # sizing step and the capture) and, beside it, the same loop eager
        timed_s = {}
        for mode in ("eager", "graph"):
            loop = DeviceAdamLoop(r, sc, opts, tgt_adam, lr_points=0.5, lr_colors=1e-2, graph=mode == "graph")
            losses = [loop.step(want_loss=False), loop.step(want_loss=False)]
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for i in range(args.adam_steps):
                losses.append(loop.step(want_loss=False))
                if (i + 1) % 50 == 0:
                    log(f"bench: Adam ({mode}) step {i + 1}, loss {float(losses[-1]):.1f}")
            torch.cuda.synchronize()
            timed_s[mode] = (time.perf_counter() - t1) / args.adam_steps
            del loop
        adam_s = timed_s["graph"]
        free, total = torch.cuda.mem_get_info()
        curve = [float(x) for x in losses]
        adam = {"steps": args.adam_steps, "steps_per_s": 1.0 / adam_s, "hip_graph": True,
                "eager_steps_per_s": 1.0 / timed_s["eager"],
                "loss_first": curve[:6], "loss_every_10": curve[::10], "loss_last": curve[-1],
                "target": "256 anisotropic Gaussians" if args.workload == "c5" else "scene with points jittered N(0, 2 px)",
                "device_used_bytes": total - free, "torch_peak_bytes": torch.cuda.max_memory_allocated()}
        log(f"bench: Adam {1.0 / adam_s:.2f} steps/s, losses {curve[:6]} ... {curve[-1]:.1f}, "
            f"device memory in use {(total - free) / 2**30:.2f} GiB")

    # per-kernel durations from HIP events recorded on the kernels' stream
    r.set_timing(True)
    r.kernel_times(reset=True)
    n_prof = max(3, min(args.steps, 10))
    for _ in range(n_prof):
        step()
    kt = r.kernel_times(reset=True)
    r.set_timing(False)
    per_kernel_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in kt.items()}
    # one launch of each kernel per step except weights/zero (twice)
    log("bench: per-kernel ms " + json.dumps({k: round(v, 4) for k, v in per_kernel_ms.items()}))

    roofline = None
    flops_path = os.path.join(ROOT, "profiles", f"flops_{args.workload}.json")
    if os.path.exists(flops_path):
        with open(flops_path) as f:
            fl = json.load(f)
        alg = {"boundary_eval": fl["bwd_boundary_flops_per_sample"] * n_samples,
               "backward_interior": fl["bwd_interior_flops_per_sample"] * n_samples,
               "forward": fl["fwd_flops_per_sample"] * n_samples}
        dom = max(alg, key=lambda k: per_kernel_ms.get(k, 0.0))
        ms = per_kernel_ms[dom]
        achieved = alg[dom] / (ms * 1e-3) / 1e12
        traffic = None
        tr_path = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
        if os.path.exists(tr_path):
            with open(tr_path) as f:
                traffic = json.load(f).get("traffic_bytes", {}).get(dom)
        valu = None
        va_path = os.path.join(ROOT, "profiles", f"valu_{args.workload}.json")
        if os.path.exists(va_path):
            with open(va_path) as f:
                valu = json.load(f).get("kernels", {}).get(dom)
        roofline = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(achieved / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                    "kernel": {"boundary_eval": "k_bnd_eval_chk" if not args.unchecked_boundary else "k_bnd_eval",
                               "backward_interior": "k_backward_prefiltered" if opts.use_prefiltering
                               else "k_backward_interior", "forward": "k_forward"}[dom], "kernel_ms": ms,
                    "counters": valu,
                    "per_kernel_tflops": {k: round(alg[k] / (per_kernel_ms[k] * 1e-3) / 1e12, 3)
                                          for k in alg if per_kernel_ms.get(k)},
                    "note": "VALU-bound fp32 scalar geometry (no MFMA, not HBM); peak is MI355X's published "
                            "fp32 vector 157.3 TFLOP/s, which counts packed fp32 (this build compiles without "
                            "packed fp32 ops: its own ceiling is half that); algorithmic flops per sample counted "
                            f"by the oracle's op-counting scalar type, profiles/flops_{args.workload}.json; "
                            f"traffic and counters from profiles/{args.workload}_hbm_traffic.md and _valu.md"}

    batch = None
    if headline and args.batch_scenes > 0:
        batch = bench_headline_batch(args, r, dev, args.batch_scenes)

    cpu = None
    if headline and not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        from pydvg.scene import SceneDesc
        dr_host = d_render.cpu().numpy()
        cpu = oracle_py.time_slice(SceneDesc(sc), opts, dr_host, args.cpu_seconds)
        cpu = {"value": cpu["samples_per_s"], "unit": "pixel-samples/s", "cores": 1, "kind": "port",
               "sample": cpu["sample"]}

    out = {
        "metric": METRIC if headline else f"fwd+bwd pixel-samples/s ({args.workload})", "value": value, "unit": "pixel-samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": workload_desc, "global_batch": 1, "samples_per_step": n_samples, "parallelism": "single-gpu",
                   "boundary_pass": "unchecked f32" if args.unchecked_boundary else "checked, uncertain samples in f64"},
        "forward_ms": fwd_ms, "backward_ms": bwd_ms, "adam_steps_per_s": (1.0 / adam_s) if adam_s else None, "adam": adam,
        "kernel_ms": {k: round(v, 4) for k, v in per_kernel_ms.items()},
        "roofline": roofline, "cpu_baseline": cpu, "batch": batch,
    }
    print(json.dumps(out))


def bench_headline_batch(args, r, dev, n):
    """n headline-size scenes (seeds 0..n-1) stepped together by the batch
    (one persistent launch per phase, replayed from a HIP graph) against the
    plain per-scene loop over the same device scenes.  Both run the unchecked
    boundary pass, the one the batch has."""
    import torch
    from pydvg import scenes
    from pydvg.batch import Batch
    from pydvg.renderer import DeviceScene, PreparedScene

    exact = True
    r.set_exact_boundary(False)
    try:
        items = [scenes.headline(seed=k) for k in range(n)]
        n_samples = sum(sc.width * sc.height * o.samples_x * o.samples_y for sc, o in items)
        targets, singles = [], []
        for sc, opts in items:
            tds = DeviceScene(r, PreparedScene(scenes.jittered(sc, 2.0, seed=1), opts))
            tgt = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device=dev)
            tds.render(tgt)
            targets.append(tgt)
            ds = DeviceScene(r, PreparedScene(sc, opts))
            bufs, sizes = ds.grad_buffers(torch)
            singles.append((ds, tgt, torch.empty_like(tgt), torch.empty_like(tgt), bufs, sizes))
        torch.cuda.synchronize()

        def plain_step():
            for ds, tgt, img, dr, bufs, sizes in singles:
                ds.render(img)
                torch.sub(img, tgt, out=dr)
                dr.mul_(2.0)
                ds.render_backward(dr, bufs, sizes)

        def timed(fn, k):
            for _ in range(max(args.warmup, 1)):
                fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(k):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / k

        k = max(3, min(args.steps, 10))
        plain_s = timed(plain_step, k)
        b = Batch(r, items, targets)
        cfg = b.autotune(reps=1)
        b.set_graph(False)
        eager_s = timed(b.step, k)
        b.set_graph(True)
        graph_s = timed(b.step, k)
        log(f"bench: batch of {n} headline scenes: plain loop {plain_s * 1e3:.2f} ms, batched "
            f"{eager_s * 1e3:.2f} ms eager, {graph_s * 1e3:.2f} ms graph")
        del b, singles
        return {"scenes": n, "workload": "headline generator, seeds 0..%d" % (n - 1),
                "samples_per_step": n_samples, "boundary_pass": "unchecked f32 (batch and plain loop)",
                "plain_loop_ms_per_step": plain_s * 1e3, "batched_eager_ms_per_step": eager_s * 1e3,
                "batched_graph_ms_per_step": graph_s * 1e3, "batched_samples_per_s": n_samples / graph_s,
                "speedup_vs_plain_loop": plain_s / graph_s, "autotuned": cfg}
    finally:
        r.set_exact_boundary(exact and not args.unchecked_boundary)


def bench_batch(args, r, stream, dev):
    """Config 4: the batch stepped together against a plain loop that steps
    the same scenes one at a time (one set of launches per scene)."""
    import torch
    from pydvg import scenes
    from pydvg.batch import Batch
    from pydvg.renderer import DeviceScene, PreparedScene

    items = scenes.config4(seed=0)
    # the batch runs the unchecked boundary pass: the plain loop runs it too
    r.set_exact_boundary(False)
    n_samples = sum(sc.width * sc.height * o.samples_x * o.samples_y for sc, o in items)
    log(f"bench: c4 {len(items)} scenes, {n_samples} samples per step")
    targets, singles = [], []
    for sc, opts in items:
        tds = DeviceScene(r, PreparedScene(scenes.jittered(sc, 2.0, seed=1), opts))
        tgt = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device=dev)
        tds.render(tgt)
        targets.append(tgt)
        ds = DeviceScene(r, PreparedScene(sc, opts))
        bufs, sizes = ds.grad_buffers(torch)
        singles.append((ds, tgt, torch.empty_like(tgt), torch.empty_like(tgt), bufs, sizes))
    torch.cuda.synchronize()

    def plain_step():
        for ds, tgt, img, dr, bufs, sizes in singles:
            ds.render(img)
            torch.sub(img, tgt, out=dr)
            dr.mul_(2.0)
            ds.render_backward(dr, bufs, sizes)

    def timed(fn, k):
        for _ in range(max(args.warmup, 1)):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k

    plain_s = timed(plain_step, args.steps)
    log(f"bench: plain loop {plain_s * 1e3:.3f} ms/step")
    batch = Batch(r, items, targets)
    t_tune = time.perf_counter()
    cfg = batch.autotune(reps=1)
    tune_s = time.perf_counter() - t_tune
    log(f"bench: autotuned {cfg} in {tune_s:.2f} s")
    batch.set_graph(False)
    eager_s = timed(batch.step, args.steps)
    batch.set_graph(True)
    graph_s = timed(batch.step, args.steps)
    log(f"bench: batched {eager_s * 1e3:.3f} ms/step eager, {graph_s * 1e3:.3f} ms/step graph")
    out = {
        "metric": "fwd+bwd pixel-samples/s (c4, batched)", "value": n_samples / graph_s, "unit": "pixel-samples/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": graph_s * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": WORKLOADS["c4"][1], "global_batch": len(items), "samples_per_step": n_samples,
                   "parallelism": "single-gpu", "boundary_pass": "unchecked f32 (batch and plain loop)"},
        "plain_loop_ms_per_step": plain_s * 1e3, "batched_eager_ms_per_step": eager_s * 1e3,
        "speedup_vs_plain_loop": plain_s / graph_s, "autotuned": cfg, "autotune_seconds": tune_s,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
