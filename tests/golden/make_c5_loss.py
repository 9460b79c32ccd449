"""Writes tests/golden/c5_reduced_loss.json: the loss curve of the CPU
oracle's Adam loop on the optimisation workload scaled to 256x256.

Scene: the 2048x2048 / 10000-path generator at 256x256 with the path density
kept (156 closed cubic paths, 2x2 jittered samples, box r = 0.5); target: the
procedural sum of 256 anisotropic Gaussians at 256x256.  Each step renders
with the oracle (the reference's f32 forward), takes L = sum((img - tgt)^2),
d_render = 2 (img - tgt), runs the oracle's render_backward (the reference's
CPU backward, f64 solvers) and applies Adam (beta1 0.9, beta2 0.999,
eps 1e-8; lr 0.5 px on path points, 1e-2 on solid fill colours, colours
clipped to [0, 1]).  The losses L_0..L_4 are those before each update.

Run from the repository root after building the oracle (make -C oracle):
    python tests/golden/make_c5_loss.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "diffvg-rs_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

STEPS = 5


def main():
    import oracle_py
    from pydvg import scenes
    from pydvg.optim import Adam
    from pydvg.scene import Path, SceneDesc, Solid

    sc, opts = scenes.config5_reduced(seed=0)
    tgt = scenes.gaussian_target(sc.width, sc.height, 256, seed=0)
    paths = [s.geometry for s in sc.shapes if isinstance(s.geometry, Path)]
    solid = [gi for gi, g in enumerate(sc.groups) if isinstance(g.fill, Solid)]
    pts = np.concatenate([p.points.reshape(-1) for p in paths]).astype(np.float32)
    cols = np.array([sc.groups[gi].fill.color for gi in solid], np.float32).reshape(-1)
    opt_p, opt_c = Adam(pts.shape, 0.5), Adam(cols.shape, 1e-2)
    losses = []
    for step in range(STEPS):
        desc = SceneDesc(sc)
        img = oracle_py.render(desc, opts)
        diff = img.astype(np.float64) - tgt
        losses.append(float((diff * diff).sum()))
        d_render = (2.0 * (img - tgt)).astype(np.float32)
        g = oracle_py.render_backward(desc, opts, d_render)
        d_grp = g["d_group_data"].reshape(-1, 16)
        g_cols = np.concatenate([d_grp[gi, 8:12] for gi in solid])
        pts, cols = opt_p.update(pts, g["d_shape_points"]), opt_c.update(cols, g_cols)
        cols = np.clip(cols, 0.0, 1.0)
        o = 0
        for p in paths:
            n = p.points.size
            p.points = pts[o:o + n].reshape(-1, 2)
            o += n
        for k, gi in enumerate(solid):
            sc.groups[gi].fill = Solid(tuple(float(c) for c in cols[k * 4:k * 4 + 4]))
        print(f"step {step}: loss {losses[-1]:.6f}", flush=True)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_reduced_loss.json")
    with open(out, "w") as f:
        json.dump({"workload": "256x256, 156 closed cubic paths, 2x2 jittered spp, box r=0.5, "
                               "target: 256 anisotropic Gaussians, Adam lr 0.5 / 1e-2",
                   "losses": losses}, f, indent=1)
    print("wrote", out)


if __name__ == "__main__":
    main()
