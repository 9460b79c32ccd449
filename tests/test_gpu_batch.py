"""GPU: batched stepping (dvg_batch_*) against the one-scene path.

A batch of scenes of different sizes and modes (non-prefiltered cubic paths,
mixed shapes with gradients and strokes, prefiltered strokes and polygons) is
stepped with forward + L2 loss gradient + backward.  Every scene's image and
gradients must equal those of the one-scene device path (dvg_device_render /
dvg_device_render_backward, in the batch's unchecked boundary pass) on the
same inputs, up to the summation order of
float atomics: images within 1e-6 absolute, gradients within 1e-5 relative L2.
The same holds for every replay of the captured HIP graph, interleaved with
eager work, and after the autotuner has changed the launch shapes.
"""
import numpy as np
import pytest

from parity_util import assert_grads
from pydvg import scenes
from pydvg.scene import RenderOptions

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-6
GRAD_REL = 1e-5


def _items():
    pf = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=True)
    return [
        scenes.small_cubic_scene(48, 40, 6, seed=11, spp=2),
        (scenes.mixed_scene(64, 64, 20, seed=12), scenes.config2()[1]),
        scenes.prefiltered_scene(40, 48, 6, 3, seed=13, step=8.0, radius=(4.0, 12.0), seg_range=(2, 5)),
        (scenes.mixed_scene(32, 32, 10, seed=14), pf),
        scenes.small_cubic_scene(20, 20, 3, seed=15, spp=3),
        # lists longer than the LDS bit set: the batch's global bit-set slots
        (scenes.mixed_scene(32, 32, 300, seed=16), scenes.config2()[1]),
    ]


@pytest.fixture(scope="module")
def setup(renderer):
    import torch
    from pydvg.batch import Batch
    from pydvg.renderer import DeviceScene, PreparedScene

    items = _items()
    targets, refs = [], []
    # the one-scene path in the batch's boundary pass (the unchecked one)
    renderer.set_exact_boundary(False)
    try:
        for i, (sc, opts) in enumerate(items):
            print(f"setup scene {i}: {sc.width}x{sc.height}, {len(sc.shapes)} shapes", flush=True)
            tds = DeviceScene(renderer, PreparedScene(scenes.jittered(sc, 2.0, seed=9), opts))
            tgt = torch.empty(sc.width * sc.height * 4, dtype=torch.float32, device="cuda")
            tds.render(tgt)
            renderer.sync()
            targets.append(tgt)
            # one-scene path
            ds = DeviceScene(renderer, PreparedScene(sc, opts))
            img = torch.empty_like(tgt)
            ds.render(img)
            renderer.sync()
            dr = 2.0 * (img - tgt)
            bufs, sizes = ds.grad_buffers(torch)
            # the context's stream does not wait for torch's: finish torch's work
            torch.cuda.synchronize()
            ds.render_backward(dr, bufs, sizes)
            renderer.sync()
            refs.append((img.cpu().numpy(), {k: bufs[k].cpu().numpy()[:sizes[k]] for k in bufs}))
    finally:
        renderer.set_exact_boundary(True)
    print("setup: batch of the scenes", flush=True)
    batch = Batch(renderer, items, targets)
    return batch, refs


def _check(batch, refs, label):
    for i, (img_ref, g_ref) in enumerate(refs):
        img, g = batch.read(i)
        err = float(np.abs(img - img_ref).max())
        print(f"{label} scene {i}: image max |diff| {err:.3g}", flush=True)
        total, per = assert_grads(g, g_ref, 1e-3, label=f"{label} scene {i}")
        assert err <= IMG_TOL
        assert total <= GRAD_REL, per


def test_batch_step_matches_single_scene_path(renderer, setup):
    batch, refs = setup
    batch.set_graph(False)
    batch.set_config(2, 2, 4, 8)
    print("case batch eager", flush=True)
    batch.step()
    _check(batch, refs, "eager")


def test_batch_graph_replays_match(renderer, setup):
    import torch
    batch, refs = setup
    batch.set_config(2, 2, 4, 0)
    batch.set_graph(True)
    print("case batch graph", flush=True)
    for rep in range(3):
        batch.step()
        # eager work of the program between replays
        x = torch.ones(1 << 16, device="cuda")
        (x * 2.0).sum().item()
        _check(batch, refs, f"graph replay {rep}")
    batch.set_graph(False)


def test_batch_autotuned_shapes_match(renderer, setup):
    batch, refs = setup
    batch.set_graph(False)
    print("case batch autotune", flush=True)
    cfg = batch.autotune(reps=1)
    print("autotuned", cfg)
    assert 1 <= cfg["wg_per_cu"] <= 64 and 0 <= cfg["sblock_log2"] <= 4 and 1 <= cfg["bnd_log2"] <= 4
    batch.step()
    _check(batch, refs, "autotuned")
    # every candidate shape gives the same results
    for sb, bb, ul in ((1, 1, 0), (3, 3, 1), (2, 2, 8)):
        batch.set_config(sb, bb, 2, ul)
        batch.step()
        _check(batch, refs, f"shape {sb}/{bb}/{ul}")
