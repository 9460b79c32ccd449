"""GPU: device-side re-packing (dvg_device_scene_update) and the device
Adam loop.

After moving path points and changing colours on the device, render and
backward must equal those of a scene packed on the host from the same
parameters (images within 1e-6, gradients within 1e-5 relative L2: only the
order of float atomics may differ), including the boundary tables and tile
lists the update rebuilds.  The device Adam loop must follow the host loop's
losses over its first steps.
"""
import copy

import numpy as np
import pytest

from parity_util import assert_grads
from pydvg import scenes
from pydvg.scene import Path, Solid

pytestmark = pytest.mark.gpu


def _render(renderer, ds, tgt, torch):
    img = torch.empty_like(tgt)
    ds.render(img)
    renderer.sync()
    dr = 2.0 * (img - tgt)
    torch.cuda.synchronize()
    bufs, sizes = ds.grad_buffers(torch)
    torch.cuda.synchronize()
    ds.render_backward(dr, bufs, sizes)
    renderer.sync()
    return img.cpu().numpy(), {k: bufs[k].cpu().numpy()[:sizes[k]] for k in bufs}


@pytest.mark.parametrize("name", ["cubic", "mixed"])
def test_update_matches_host_packing(renderer, name):
    import torch
    from pydvg.renderer import DeviceScene, PreparedScene
    if name == "cubic":
        sc, opts = scenes.small_cubic_scene(64, 48, 10, seed=21, spp=2)
    else:
        sc, opts = scenes.mixed_scene(64, 64, 18, seed=22), scenes.config2()[1]
    print(f"case update {name}", flush=True)
    tgt_ds = DeviceScene(renderer, PreparedScene(scenes.jittered(sc, 3.0, seed=5), opts))
    tgt = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device="cuda")
    tgt_ds.render(tgt)
    renderer.sync()
    prep = PreparedScene(sc, opts)
    ds = DeviceScene(renderer, prep)
    rng = np.random.default_rng(7)
    # new parameters: points moved by up to ~4 px, new solid fill colours
    moved = copy.deepcopy(sc)
    for s in moved.shapes:
        if isinstance(s.geometry, Path):
            s.geometry.points = (s.geometry.points + rng.normal(0, 2.0, s.geometry.points.shape)).astype(np.float32)
    for g in moved.groups:
        if isinstance(g.fill, Solid):
            g.fill = Solid(tuple(float(c) for c in rng.uniform(0.2, 0.9, 4)))
    mprep = PreparedScene(moved, opts)
    pts = torch.tensor(np.asarray(mprep.desc.path_points, np.float32), device="cuda")
    G = mprep.desc.num_groups
    cols = np.concatenate([np.asarray(mprep.desc.fill_color, np.float32).reshape(G, 4),
                           np.asarray(mprep.desc.stroke_color, np.float32).reshape(G, 4)], axis=1)
    cols = torch.tensor(cols.reshape(-1), device="cuda")
    torch.cuda.synchronize()
    for rep in range(2):  # the second update starts from updated buffers
        ds.update(pts, cols)
        ds.check()
        img, g = _render(renderer, ds, tgt, torch)
        ref_img, ref_g = _render(renderer, DeviceScene(renderer, mprep), tgt, torch)
        err = float(np.abs(img - ref_img).max())
        print(f"update {name} rep {rep}: image max |diff| {err:.3g}")
        total, per = assert_grads(g, ref_g, 1e-3, label=f"update {name} rep {rep}")
        assert err <= 1e-6
        assert total <= 1e-5, per


def test_device_adam_follows_host_adam(renderer):
    import torch
    from pydvg.optim import AdamLoop, DeviceAdamLoop
    from pydvg.renderer import DeviceScene, PreparedScene
    sc, opts = scenes.small_cubic_scene(64, 64, 12, seed=23, spp=2)
    print("case device adam", flush=True)
    tgt_ds = DeviceScene(renderer, PreparedScene(scenes.jittered(sc, 3.0, seed=6), opts))
    tgt = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device="cuda")
    tgt_ds.render(tgt)
    renderer.sync()
    dev = DeviceAdamLoop(renderer, copy.deepcopy(sc), opts, tgt, lr_points=0.5, lr_colors=1e-2)
    dev_losses = [dev.step() for _ in range(4)]
    host = AdamLoop(renderer, copy.deepcopy(sc), opts, tgt, lr_points=0.5, lr_colors=1e-2)
    host_losses = [host.step() for _ in range(4)]
    print("device", dev_losses, "host", host_losses)
    for a, b in zip(dev_losses, host_losses):
        assert abs(a - b) <= 1e-4 * abs(b)
    assert dev_losses[-1] < dev_losses[0]


def test_device_adam_follows_oracle_loss_curve(renderer):
    """The optimisation workload at 256x256 (156 paths, Gaussian target):
    the device loop's first five losses against the CPU oracle's Adam loop
    (tests/golden/c5_reduced_loss.json, written by make_c5_loss.py)."""
    import json
    import os
    import torch
    from pydvg.optim import DeviceAdamLoop
    with open(os.path.join(os.path.dirname(__file__), "golden", "c5_reduced_loss.json")) as f:
        ref = json.load(f)["losses"]
    sc, opts = scenes.config5_reduced(seed=0)
    tgt = torch.tensor(scenes.gaussian_target(sc.width, sc.height, 256, seed=0), device="cuda")
    print("case device adam vs oracle loss curve", flush=True)
    loop = DeviceAdamLoop(renderer, sc, opts, tgt, lr_points=0.5, lr_colors=1e-2)
    losses = [loop.step() for _ in range(len(ref))]
    rel = [abs(a - b) / b for a, b in zip(losses, ref)]
    print("device", losses, "oracle", ref, "rel", rel)
    assert rel[0] <= 1e-5
    assert max(rel) <= 1e-3


def test_device_adam_graph_replays_match_eager(renderer):
    """The device Adam loop with its step captured in a HIP graph against an
    eager step taken from the same state, step after step, with the
    program's own eager work and synchronisations between the replays.  Two
    runs differ only in the order of float atomic additions, which Adam's
    normalised update would amplify over steps, so before every step the
    eager loop takes the graph loop's state (parameters, moments, b^t).
    Bounds: image 1e-6, gradients 1e-5 relative (every field 1e-3),
    parameters 1e-3 px and 1e-5, loss 1e-6 relative."""
    import torch
    from pydvg.optim import DeviceAdamLoop
    sc, opts = scenes.config5_reduced(seed=0)
    tgt = torch.tensor(scenes.gaussian_target(sc.width, sc.height, 256, seed=0), device="cuda")
    print("case device adam graph replay", flush=True)
    graph = DeviceAdamLoop(renderer, copy.deepcopy(sc), opts, tgt, lr_points=0.5, lr_colors=1e-2, graph=True)
    eager = DeviceAdamLoop(renderer, copy.deepcopy(sc), opts, tgt, lr_points=0.5, lr_colors=1e-2)
    state = ("points", "colors", "m_p", "v_p", "m_c", "v_c", "pw1", "pw2")
    for k in range(6):
        for name in state:
            getattr(eager, name).copy_(getattr(graph, name))
        eager.t = graph.t
        torch.cuda.synchronize()
        lg = graph.step()
        le = eager.step()
        # eager work of the program between the replays
        x = torch.ones(1 << 16, device="cuda")
        (x * 3.0).sum().item()
        torch.cuda.synchronize()
        di = float((eager.img - graph.img).abs().max())
        dp = float((eager.points - graph.points).abs().max())
        dc = float((eager.colors - graph.colors).abs().max())
        print(f"step {k} ({'replay' if graph.graph is not None else 'eager'}): loss {lg:.7g} / {le:.7g}; "
              f"max |image diff| {di:.3g}, |points diff| {dp:.3g}, |colours diff| {dc:.3g}", flush=True)
        ge = {n: eager.bufs[n].cpu().numpy()[:eager.sizes[n]] for n in eager.bufs}
        gg = {n: graph.bufs[n].cpu().numpy()[:graph.sizes[n]] for n in graph.bufs}
        total, _ = assert_grads(gg, ge, 1e-3, label=f"step {k} gradients, graph against eager")
        assert total <= 1e-5
        assert di <= 1e-6 and dp <= 1e-3 and dc <= 1e-5
        assert abs(le - lg) <= 1e-6 * abs(le)
    assert graph.graph is not None
