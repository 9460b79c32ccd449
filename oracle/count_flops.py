"""Algorithmic FLOPs per pixel-sample of the headline workload.

Runs the oracle in count mode over evenly spaced slices of the sample index
range and writes profiles/flops_headline.json.  Counting rules: +, -, *, /,
compare, min/max and sqrt are 1 flop each, acos/cos/sin/atan2/pow are 1 SFU op
each (reported apart and added to the flop total).  Scene evaluations count
the reference's work with its own culling: groups of the sample's 16x16 tile
list, every segment of an in-bounds path (the path BVH has one leaf for <= 8
segments), no distance queries in non-prefiltered mode (they do not reach the
output).  Fixed per-sample costs (gather of d_render over the 3x3 filter
support, filter-radius adjoint, reverse compositing, boundary point sampling
and gradient scatter) are counted from their formulas:
  gather_d_color   9 taps x 12      = 108
  filter adjoint   9 taps x 25      = 225
  reverse adjoint  25 per fragment
  boundary point   100; boundary gradient scatter 150; contribution 12
Usage: python oracle/count_flops.py  (CPU; about a minute)
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "diffvg-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def main():
    from pydvg import scenes
    from pydvg.scene import SceneDesc
    import oracle_py
    sc, opts = scenes.headline(seed=0)
    desc = SceneDesc(sc)
    total = sc.width * sc.height * opts.samples_x * opts.samples_y
    n_slices, per = 16, 4096
    acc = {}
    for k in range(n_slices):
        i0 = (total // n_slices) * k + 12345
        c = oracle_py.count_ops(desc, opts, i0, i0 + per)
        for key, v in c.items():
            acc[key] = acc.get(key, 0) + v
    n = acc["samples"]
    fwd = (acc["fwd_flops"] + acc["fwd_sfu"]) / n + 4 * 3  # + splat (4 channels)
    frags = acc["fragments"] / n
    interior = (acc["fwd_flops"] + acc["fwd_sfu"]) / n + 25 * frags + 108 + 225
    bnd_frac = acc["bnd_samples"] / n
    boundary = (acc["bnd_flops"] + acc["bnd_sfu"]) / n + bnd_frac * (100 + 150 + 12 + 108)
    out = {
        "workload": "headline 1024x1024, 1000 closed cubic paths, 2x2 spp",
        "samples_counted": n,
        "fwd_flops_per_sample": fwd,
        "fragments_per_sample": frags,
        "bwd_interior_flops_per_sample": interior,
        "bwd_boundary_flops_per_sample": boundary,
        "bwd_flops_per_sample": interior + boundary,
        "sfu_share_fwd": acc["fwd_sfu"] / max(1.0, acc["fwd_flops"] + acc["fwd_sfu"]),
        "boundary_samples_contributing": bnd_frac,
        "bwd_hbm_bytes": None,
    }
    path = os.path.join(ROOT, "profiles", "flops_headline.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
