"""GPU: signed distance (dvg_render_sdf / dvg_eval_positions, Renderer::
render_sdf / eval_positions, renderer.rs:586-848 over sdf.rs:9-262) against
the oracle.

Whole images on small scenes of every shape kind and sampling mode; points
scattered over the 1024^2 canvas of 1000 cubic paths.  The closest-point
solvers are the forward's (f32, bit-compatible with the oracle); the ellipse
Newton iteration calls the f32 atan2 / sin / cos of each side's library, so
values may differ in the last bits.  Tolerance: |gpu - oracle| <= 1e-5 +
1e-5 |v| for all but 0.1 % of the values (a sample next to an edge can take
the other sign there: both values are then within rounding of zero), and
<= 1e-3 everywhere."""
import numpy as np
import pytest

from pydvg import scenes
from pydvg.scene import Circle, Ellipse, RenderOptions, Scene, Shape, ShapeGroup, Solid, SceneDesc

pytestmark = pytest.mark.gpu

TOL, FRACTION, MAX_ABS = 1e-5, 1e-3, 1e-3


def _check(name, got, ref):
    d = np.abs(got - ref)
    bad = d > TOL + TOL * np.abs(ref)
    print(f"{name}: {got.size} values, max |diff| {d.max():.3g}, beyond tolerance {int(bad.sum())}", flush=True)
    assert np.all(np.isfinite(got))
    assert bad.mean() <= FRACTION
    assert d.max() <= MAX_ABS


def _scaled_scene():
    shapes = [Shape(Circle((8.0, 9.0), 4.0)), Shape(Ellipse((20.0, 12.0), (6.0, 3.0))),
              Shape(Circle((30.0, 30.0), 5.0), stroke_width=1.5)]
    g0 = ShapeGroup(shape_indices=[0, 1], fill=Solid((0.2, 0.3, 0.4, 1.0)))
    g0.set_shape_to_canvas(np.array([[1.5, 0.2, 3.0], [-0.1, 1.2, -2.0], [0, 0, 1]], np.float32))
    g1 = ShapeGroup(shape_indices=[2], stroke=Solid((0, 0, 0, 1)))
    return Scene(48, 44, shapes=shapes, groups=[g0, g1], background=(1, 1, 1, 1))


def _shrunk_groups():
    """Four groups of 12 mixed shapes (BVHs of two levels), shape-to-canvas
    scales 0.6, 1.0, 1.5 and 0.8 with a slight rotation: the reference's
    group-local node test then skips nodes by local distances."""
    sc = scenes.mixed_scene(96, 96, 48, 3)
    base = sc.groups
    sc.groups = []
    for k, cut in enumerate(range(0, 48, 12)):
        g = ShapeGroup(list(range(cut, cut + 12)), fill=base[cut].fill or Solid((0.3, 0.3, 0.3, 1.0)),
                       stroke=base[cut].stroke)
        s = (0.6, 1.0, 1.5, 0.8)[k]
        g.set_shape_to_canvas(np.array([[s, 0.1, 4.0 * k], [-0.1, s, 3.0], [0, 0, 1]], np.float32))
        sc.groups.append(g)
    return sc


def _cases():
    pf = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=True)
    return {
        "cubic48": scenes.small_cubic_scene(48, 40, 6, seed=11, spp=2),
        "mixed64": (scenes.mixed_scene(64, 64, 20, seed=12), scenes.config2()[1]),
        "prefiltered48": scenes.prefiltered_scene(48, 48, 6, 3, seed=13, step=8.0, radius=(4.0, 12.0),
                                                  seg_range=(2, 5)),
        "mixed32_pf": (scenes.mixed_scene(32, 32, 10, seed=14), pf),
        "scaled_groups": (_scaled_scene(), RenderOptions(samples_x=2, samples_y=2, seed=3)),
        "shrunk_groups": (_shrunk_groups(), RenderOptions(samples_x=2, samples_y=2, seed=5)),
    }


@pytest.mark.parametrize("name", list(_cases().keys()))
def test_render_sdf_matches_oracle(renderer, oracle, name):
    sc, opts = _cases()[name]
    print(f"case sdf {name} {sc.width}x{sc.height}", flush=True)
    got = renderer.render_sdf(sc, opts)
    ref = oracle.render_sdf(SceneDesc(sc), opts)
    _check(name, got, ref)
    assert (ref < 0).any() and (ref > 0).any()


@pytest.mark.parametrize("name", ["mixed64", "scaled_groups", "shrunk_groups"])
def test_eval_positions_matches_oracle(renderer, oracle, name):
    sc, opts = _cases()[name]
    rng = np.random.default_rng(4)
    p = np.stack([rng.uniform(-8, sc.width + 8, 3000), rng.uniform(-8, sc.height + 8, 3000)], 1).astype(np.float32)
    print(f"case eval_positions {name}", flush=True)
    _check(name, renderer.eval_positions(sc, p, opts), oracle.eval_positions(SceneDesc(sc), opts, p))


def test_shrinking_group_keeps_the_reference_node_test(renderer):
    """The reference skips a group BVH node by its group-local distance
    (distance.rs:346-360): at p, group 0's circle (12 px away, 24 local units)
    is skipped while group 1's circle is 15 px away, and wins when that one is
    40 px away (the closed forms of the oracle's pin)."""
    import test_oracle_sdf as t
    p = np.array([[50.0, 40.0]], np.float32)
    for far, want in ((False, 15.0), (True, 12.0)):
        sc = t._shrunk_pair(far)
        print(f"case shrunk pair far={far}", flush=True)
        assert abs(float(renderer.eval_positions(sc, p, t.OPTS)[0]) - want) < 1e-4
        img = renderer.render_sdf(sc, t.OPTS)
        assert abs(float(img[40 * sc.width + 50]) - float(
            renderer.eval_positions(sc, np.array([[50.5, 40.5]], np.float32), t.OPTS)[0])) < 1e-5


def test_tiled_sdf_on_the_full_canvas(renderer, oracle):
    """The 1024^2 image of 1000 cubic paths (2x2 samples, unjittered), computed
    by the tiled kernel (candidate lists per tile): against the walk at every
    sample (eval_positions, pixel means), and against the oracle at the
    samples of three 32x32 crops."""
    sc, _ = scenes.headline()
    opts = RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=False, use_prefiltering=False)
    print("case tiled sdf headline", flush=True)
    img = renderer.render_sdf(sc, opts).reshape(sc.height, sc.width)
    offs = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]], np.float32)
    yy, xx = np.mgrid[0:sc.height, 0:sc.width]
    pix = np.stack([xx.reshape(-1), yy.reshape(-1)], 1).astype(np.float32)
    pts = (pix[:, None, :] + offs[None]).reshape(-1, 2)
    walk = renderer.eval_positions(sc, pts, opts).reshape(sc.height, sc.width, 4).mean(2)
    _check("tiled against walk", img.reshape(-1), walk.reshape(-1).astype(np.float32))
    for y0, x0 in ((0, 0), (500, 320), (992, 992)):
        crop = (pix.reshape(sc.height, sc.width, 2)[y0:y0 + 32, x0:x0 + 32].reshape(-1, 2)[:, None, :]
                + offs[None]).reshape(-1, 2)
        ref = oracle.eval_positions(SceneDesc(sc), opts, crop).reshape(32, 32, 4).mean(2)
        _check(f"tiled against oracle at ({x0}, {y0})", img[y0:y0 + 32, x0:x0 + 32].reshape(-1),
               ref.reshape(-1).astype(np.float32))


def test_eval_positions_full_canvas(renderer, oracle):
    """2000 points over the 1024^2 canvas of 1000 cubic paths."""
    sc, opts = scenes.headline()
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 1024, (2000, 2)).astype(np.float32)
    print("case eval_positions headline", flush=True)
    _check("headline points", renderer.eval_positions(sc, p, opts), oracle.eval_positions(SceneDesc(sc), opts, p))


def test_empty_inputs(renderer):
    sc = Scene(16, 8)
    assert np.array_equal(renderer.render_sdf(sc, RenderOptions()), np.zeros(128, np.float32))
    assert renderer.eval_positions(sc, np.zeros((0, 2), np.float32), RenderOptions()).size == 0


@pytest.mark.parametrize("name", ["cubic48", "mixed64", "scaled_groups"])
def test_eval_positions_backward_matches_oracle(renderer, oracle, name):
    """eval_positions_backward (the reference runs it on the CPU,
    backward/render.rs:213-280) against the oracle's restatement, with
    translation gradients; relative L2 <= 1e-3 for every gradient array, the
    gradient tolerance of the render backward."""
    from parity_util import assert_grads
    from pydvg.scene import BackwardOptions
    sc, opts = _cases()[name]
    rng = np.random.default_rng(7)
    p = np.stack([rng.uniform(-4, sc.width + 4, 2000), rng.uniform(-4, sc.height + 4, 2000)], 1).astype(np.float32)
    d = rng.normal(size=2000).astype(np.float32)
    d[::7] = 0.0
    print(f"case eval_positions_backward {name}", flush=True)
    g = renderer.eval_positions_backward(sc, p, opts, d, BackwardOptions(True))
    ref = oracle.eval_positions_backward(SceneDesc(sc), p, d, translation=True)
    assert_grads(g, ref, 1e-3, label=f"{name}: positions backward")


@pytest.mark.parametrize("name", ["cubic48", "mixed64", "scaled_groups"])
def test_render_backward_sdf_matches_oracle(renderer, oracle, name):
    """The d_sdf term of render_backward (the reference's CPU
    backward/render.rs:171-186) against the oracle, with translation;
    relative L2 <= 1e-3 for every gradient array."""
    from parity_util import assert_grads
    from pydvg.scene import BackwardOptions
    sc, opts = _cases()[name]
    rng = np.random.default_rng(8)
    d = rng.normal(size=sc.width * sc.height).astype(np.float32)
    d[::5] = 0.0
    print(f"case render_backward_sdf {name}", flush=True)
    g = renderer.render_backward_sdf(sc, opts, d, BackwardOptions(True))
    ref = oracle.render_backward_sdf(SceneDesc(sc), opts, d, translation=True)
    assert_grads(g, ref, 1e-3, label=f"{name}: d_sdf backward")
