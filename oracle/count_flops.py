"""Algorithmic FLOPs per pixel-sample of the benchmark workloads -- TEST
INFRASTRUCTURE (it runs the oracle; nothing in the product path loads it).

Runs liboracle_count.so -- the oracle's sources compiled with the op-counting
scalars of counted.h (Counted<float>, Counted<double>) -- over evenly spaced
slices of each workload's pixel-sample index range, and writes
profiles/flops_<workload>.json.  Rules (counted.h): +, -, *, /, comparisons,
min / max and sqrt are 1 flop each; acos / cos / sin / atan2 / pow / cbrt /
exp / log are 1 SFU op each, reported apart and added to the total.  The
work counted is the reference algorithm's with the reference GPU's culling:
a scene query counts only the groups of its 16x16 tile's list (the oracle
itself loops over every group).  Phases: the forward (scene evaluation and
splat per sample), the interior colour adjoint (scene evaluation with
fragments, reverse compositing, paint and filter adjoints, d_render gather)
and boundary sampling (edge point sampling, the two colour queries,
contribution, gradient scatter), each per pixel-sample of the workload.

Usage: python oracle/count_flops.py [workload ...]   (CPU; minutes)
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "diffvg-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

WORKLOADS = {"headline": "headline", "c1": "config1", "c2": "config2", "c3": "config3", "c5": "config5"}


def count(name, n_slices=16, per=2048):
    from pydvg import scenes
    from pydvg.scene import SceneDesc
    import oracle_py
    sc, opts = getattr(scenes, WORKLOADS[name])(seed=0)
    desc = SceneDesc(sc)
    total = sc.width * sc.height * opts.samples_x * opts.samples_y
    d_render = np.random.default_rng(0).normal(size=sc.width * sc.height * 4).astype(np.float32)
    acc = {"fwd": [0.0, 0.0], "interior": [0.0, 0.0], "boundary": [0.0, 0.0]}
    n = 0
    for k in range(n_slices):
        i0 = (total // n_slices) * k + 12345 % max(total // n_slices, 1)
        i1 = min(i0 + per, total)
        c = oracle_py.count_slice(desc, opts, d_render, i0, i1)
        for key in acc:
            acc[key][0] += c[key][0]
            acc[key][1] += c[key][1]
        n += i1 - i0
    per_sample = {k: (v[0] + v[1]) / n for k, v in acc.items()}
    return {
        "workload": f"{name}: {sc.width}x{sc.height}, {len(sc.shapes)} shapes, "
                    f"{opts.samples_x}x{opts.samples_y} spp, prefiltered={bool(opts.use_prefiltering)}",
        "method": "oracle compiled with op-counting scalars (oracle/counted.h), reference GPU tile culling",
        "samples_counted": n, "slices": n_slices,
        "fwd_flops_per_sample": per_sample["fwd"],
        "bwd_interior_flops_per_sample": per_sample["interior"],
        "bwd_boundary_flops_per_sample": per_sample["boundary"],
        "bwd_flops_per_sample": per_sample["interior"] + per_sample["boundary"],
        "sfu_per_sample": {k: v[1] / n for k, v in acc.items()},
    }


def main():
    names = sys.argv[1:] or list(WORKLOADS)
    for name in names:
        out = count(name)
        path = os.path.join(ROOT, "profiles", f"flops_{name}.json")
        with open(path, "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
