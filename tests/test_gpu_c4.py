"""GPU: the batched workload itself, 64 scenes of uneven size (scenes.config4:
16-4096 mixed shapes, 128^2-1024^2, 2x2 jittered samples, 100 M samples in
all), stepped together (forward, L2 loss gradient, backward) with one
persistent launch per phase.

  * The 8 smallest scenes (128^2, 17-911 shapes) against the oracle, whole
    images: images within 1e-5 except at most 0.1 % of the pixels, gradients
    per field within 1e-3 relative, for d_render = 2 (image - target) as the
    batch computed it (the target: the scene with its points jittered).
  * The other 56 against the one-scene device path on the same inputs, in
    the batch's boundary pass (the unchecked f32 one), up to the order of
    float atomics: images 1e-6, gradients 1e-5 relative L2 and every field
    within 1e-3.
  * Linearity of the whole batched backward: the step is repeated with the
    targets T_a, T_b and (T_a + T_b) / 2, whose gradients must average.
"""
import numpy as np
import pytest

from parity_util import assert_grads, image_report
from pydvg import scenes
from pydvg.scene import GRAD_FIELDS, SceneDesc, finalize_background_gradients

pytestmark = pytest.mark.gpu

N_ORACLE = 8


@pytest.fixture(scope="module")
def c4(renderer):
    import torch
    from pydvg.batch import Batch
    from pydvg.renderer import DeviceScene, PreparedScene

    items = scenes.config4()
    print(f"setup c4: {len(items)} scenes, "
          f"{sum(sc.width * sc.height * o.samples_x * o.samples_y for sc, o in items) / 1e6:.1f} M samples", flush=True)
    targets = []
    for sc, opts in items:
        tds = DeviceScene(renderer, PreparedScene(scenes.jittered(sc, 2.0, seed=1), opts))
        t = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device="cuda")
        tds.render(t)
        renderer.sync()
        targets.append(t)
    batch = Batch(renderer, items, targets)
    batch.set_graph(False)
    print("setup c4: batch step", flush=True)
    batch.step()
    reads = [batch.read(i) for i in range(len(items))]
    order = sorted(range(len(items)), key=lambda i: (items[i][0].width * items[i][0].height, len(items[i][0].shapes)))
    return items, targets, batch, reads, order


def test_c4_smallest_scenes_match_oracle(oracle, c4):
    items, targets, _, reads, order = c4
    for i in order[:N_ORACLE]:
        sc, opts = items[i]
        print(f"case c4 scene {i}: {sc.width}x{sc.height}, {len(sc.shapes)} shapes (oracle)", flush=True)
        img, g = reads[i]
        desc = SceneDesc(sc)
        frac, mx = image_report(img, oracle.render(desc, opts))
        print(f"scene {i}: pixels beyond 1e-5: {frac:.2e}, max |diff| {mx:.3g}", flush=True)
        assert frac <= 1e-3
        d_render = (2.0 * (img - targets[i].cpu().numpy())).astype(np.float32)
        # the batch returns raw accumulations, like the packed seam (finalized
        # here on a copy: the fixture's reads are shared by the tests)
        g = {k: np.array(g[k]) for k in GRAD_FIELDS}
        finalize_background_gradients(sc, g)
        assert_grads(g, oracle.render_backward(desc, opts, d_render), 1e-3, label=f"c4 scene {i}")


def test_c4_other_scenes_match_single_scene_path(renderer, c4):
    items, targets, _, reads, order = c4
    renderer.set_exact_boundary(False)  # the batch's pass
    try:
        _one_scene_path(renderer, items, targets, reads, order[N_ORACLE:])
    finally:
        renderer.set_exact_boundary(True)


def _one_scene_path(renderer, items, targets, reads, which):
    import torch
    from pydvg.renderer import DeviceScene, PreparedScene
    for i in which:
        sc, opts = items[i]
        print(f"case c4 scene {i}: {sc.width}x{sc.height}, {len(sc.shapes)} shapes (one-scene path)", flush=True)
        ds = DeviceScene(renderer, PreparedScene(sc, opts))
        img = torch.empty_like(targets[i])
        ds.render(img)
        renderer.sync()
        dr = 2.0 * (img - targets[i])
        bufs, sizes = ds.grad_buffers(torch)
        torch.cuda.synchronize()  # the context's stream does not wait for torch's
        ds.render_backward(dr, bufs, sizes)
        renderer.sync()
        g_ref = {k: bufs[k].cpu().numpy()[:sizes[k]] for k in bufs}
        img_b, g_b = reads[i]
        err = float(np.abs(img_b - img.cpu().numpy()).max())
        total, _ = assert_grads(g_b, g_ref, 1e-3, label=f"c4 scene {i}, batch against one-scene path")
        assert err <= 1e-6
        assert total <= 1e-5


def test_c4_batched_backward_is_linear(renderer, c4):
    import torch
    items, targets, batch, reads, order = c4
    print("case c4 linearity", flush=True)
    saved = [t.clone() for t in targets]
    gens = torch.Generator(device="cuda")
    gens.manual_seed(7)
    try:
        for t in targets:
            t.add_(0.05 * torch.randn(t.shape, device="cuda", generator=gens))
        torch.cuda.synchronize()
        batch.step()
        g_b = [batch.read(i)[1] for i in range(len(items))]
        for t, s in zip(targets, saved):
            t.copy_(0.5 * (t + s))
        torch.cuda.synchronize()
        batch.step()
        g_c = [batch.read(i)[1] for i in range(len(items))]
    finally:
        for t, s in zip(targets, saved):
            t.copy_(s)
        torch.cuda.synchronize()
    for i in order:
        mean = {k: 0.5 * (reads[i][1][k] + g_b[i][k]) for k in GRAD_FIELDS}
        total, _ = assert_grads(g_c[i], mean, 1e-3, label=f"c4 scene {i} linearity")
        assert total <= 1e-4
