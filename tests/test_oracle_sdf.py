"""The oracle's signed distance (Renderer::render_sdf / eval_positions,
renderer.rs:586-848 over sdf.rs:9-262) pinned by closed forms.

A filled circle's signed distance is |p - c| - r; a filled axis-aligned box's
is the Euclidean distance to the box outside and minus the distance to the
nearest edge inside; a square polygon path gives the box's values; a group
scaled by s maps a circle of radius r to one of radius s r; an unfilled
(stroke-only) group is never "inside"; several shapes give the minimum with
the sign of the nearest one's group; render_sdf without jitter at one sample
a pixel is eval_positions at the pixel centres.  A wrong transform order, a
dropped sign or a min taken over the wrong set fails one of these."""
import numpy as np
import pytest

from pydvg import abi
from pydvg.scene import Circle, Path, Rect, RenderOptions, Scene, Shape, ShapeGroup, Solid, SceneDesc

OPTS = RenderOptions(samples_x=1, samples_y=1, jitter=False, use_prefiltering=False)


def _scene(shapes, groups, w=40, h=32):
    return Scene(w, h, shapes=shapes, groups=groups, background=(1.0, 1.0, 1.0, 1.0))


def _fill(idx, **kw):
    return ShapeGroup(shape_indices=idx, fill=Solid((0.2, 0.4, 0.6, 1.0)), **kw)


def _pts(rng, n=400, w=40, h=32):
    return np.stack([rng.uniform(-5, w + 5, n), rng.uniform(-5, h + 5, n)], 1).astype(np.float32)


def _box_sdf(p, lo, hi):
    dx = np.maximum(np.maximum(lo[0] - p[:, 0], 0), p[:, 0] - hi[0])
    dy = np.maximum(np.maximum(lo[1] - p[:, 1], 0), p[:, 1] - hi[1])
    out = np.hypot(dx, dy)
    inside = (p[:, 0] >= lo[0]) & (p[:, 0] <= hi[0]) & (p[:, 1] >= lo[1]) & (p[:, 1] <= hi[1])
    edge = np.minimum.reduce([p[:, 0] - lo[0], hi[0] - p[:, 0], p[:, 1] - lo[1], hi[1] - p[:, 1]])
    return np.where(inside, -edge, out)


def test_circle_is_distance_to_centre_minus_radius(oracle):
    sc = _scene([Shape(Circle((20.0, 16.0), 7.0))], [_fill([0])])
    p = _pts(np.random.default_rng(0))
    got = oracle.eval_positions(SceneDesc(sc), OPTS, p)
    want = np.hypot(p[:, 0] - 20.0, p[:, 1] - 16.0) - 7.0
    assert np.allclose(got, want, atol=2e-5, rtol=1e-5)
    assert (got < 0).any() and (got > 0).any()


@pytest.mark.parametrize("as_path", [False, True])
def test_box_and_square_polygon(oracle, as_path):
    lo, hi = (5.0, 6.0), (25.0, 20.0)
    if as_path:
        geom = Path(num_control_points=[0, 0, 0, 0],
                    points=[[lo[0], lo[1]], [hi[0], lo[1]], [hi[0], hi[1]], [lo[0], hi[1]]], is_closed=True)
    else:
        geom = Rect(lo, hi)
    sc = _scene([Shape(geom)], [_fill([0])])
    p = _pts(np.random.default_rng(1))
    got = oracle.eval_positions(SceneDesc(sc), OPTS, p)
    assert np.allclose(got, _box_sdf(p, lo, hi), atol=2e-5, rtol=1e-5)


def test_group_scale_maps_the_distance(oracle):
    s = 2.0
    g = _fill([0])
    g.set_shape_to_canvas(np.array([[s, 0, 3.0], [0, s, -2.0], [0, 0, 1]], np.float32))
    sc = _scene([Shape(Circle((8.0, 9.0), 4.0))], [g])
    p = _pts(np.random.default_rng(2))
    got = oracle.eval_positions(SceneDesc(sc), OPTS, p)
    c = np.array([8.0 * s + 3.0, 9.0 * s - 2.0])
    want = np.hypot(p[:, 0] - c[0], p[:, 1] - c[1]) - 4.0 * s
    assert np.allclose(got, want, atol=5e-5, rtol=1e-5)


def test_nearest_shape_and_its_group_sign(oracle):
    shapes = [Shape(Circle((10.0, 10.0), 5.0)), Shape(Circle((30.0, 20.0), 6.0))]
    stroke_only = ShapeGroup(shape_indices=[1], stroke=Solid((0, 0, 0, 1)))
    sc = _scene(shapes, [_fill([0]), stroke_only])
    p = _pts(np.random.default_rng(3))
    got = oracle.eval_positions(SceneDesc(sc), OPTS, p)
    d0 = np.hypot(p[:, 0] - 10.0, p[:, 1] - 10.0) - 5.0
    d1 = np.abs(np.hypot(p[:, 0] - 30.0, p[:, 1] - 20.0) - 6.0)  # unfilled: never inside
    want = np.where(np.abs(d0) < d1, d0, d1)
    assert np.allclose(got, want, atol=2e-5, rtol=1e-5)


def test_render_sdf_is_eval_at_pixel_centres(oracle):
    sc = _scene([Shape(Circle((20.0, 16.0), 7.0)), Shape(Rect((2.0, 3.0), (9.0, 30.0)))], [_fill([0]), _fill([1])])
    desc = SceneDesc(sc)
    img = oracle.render_sdf(desc, OPTS)
    yy, xx = np.mgrid[0:sc.height, 0:sc.width]
    p = np.stack([xx.reshape(-1) + 0.5, yy.reshape(-1) + 0.5], 1).astype(np.float32)
    assert np.array_equal(img, oracle.eval_positions(desc, OPTS, p))


def test_empty_scene_is_zero(oracle):
    sc = _scene([], [])
    assert np.array_equal(oracle.render_sdf(SceneDesc(sc), OPTS), np.zeros(40 * 32, np.float32))
    assert np.array_equal(oracle.eval_positions(SceneDesc(sc), OPTS, np.zeros((3, 2), np.float32)),
                          np.zeros(3, np.float32))


def test_positions_backward_circle_closed_form(oracle):
    """sdf = |p - c| - r for a filled circle: d/dr = -1 and d/dc = -(p - c)/|p - c|
    at every point, inside or outside (render_backward_positions)."""
    sc = _scene([Shape(Circle((20.0, 16.0), 7.0))], [_fill([0])])
    rng = np.random.default_rng(6)
    p = _pts(rng, 300)
    d = rng.normal(size=300).astype(np.float32)
    g = oracle.eval_positions_backward(SceneDesc(sc), p, d)
    v = p - np.array([20.0, 16.0], np.float32)
    n = v / np.linalg.norm(v, axis=1, keepdims=True)
    want_c = -(d[:, None] * n).sum(0)
    dp = g["d_shape_params"]
    assert np.allclose(dp[0:2], want_c, rtol=1e-4, atol=1e-4)
    assert np.isclose(dp[2], -d.sum(), rtol=1e-4, atol=1e-4)


def test_sdf_image_backward_is_points_backward_at_pixel_centres(oracle):
    """One unjittered sample a pixel: the d_sdf image term of render_backward
    equals eval_positions_backward at the pixel centres (with translation)."""
    sc = _scene([Shape(Circle((20.0, 16.0), 7.0)), Shape(Rect((2.0, 3.0), (9.0, 30.0)))], [_fill([0]), _fill([1])])
    desc = SceneDesc(sc)
    rng = np.random.default_rng(9)
    d = rng.normal(size=sc.width * sc.height).astype(np.float32)
    yy, xx = np.mgrid[0:sc.height, 0:sc.width]
    p = np.stack([xx.reshape(-1) + 0.5, yy.reshape(-1) + 0.5], 1).astype(np.float32)
    a = oracle.render_backward_sdf(desc, OPTS, d, translation=True)
    b = oracle.eval_positions_backward(desc, p, d, translation=True)
    for k in ("d_shape_params", "d_shape_transform", "d_group_transform", "d_translation"):
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(a["d_shape_params"]).max() > 0


def _shrunk_pair(far):
    """Group 0 maps shape space to canvas at scale 0.5: its circle (centre
    (100, 110), radius 6) lies at canvas (50, 55), radius 3, 12 px from
    p = (50, 40) but 24 group-local units away.  Group 1's circle (radius 5)
    is 15 px from p, or 40 px when `far`."""
    g0 = _fill([0])
    g0.set_shape_to_canvas(np.array([[0.5, 0, 0], [0, 0.5, 0], [0, 0, 1]], np.float32))
    c1 = (50.0, -5.0) if far else (50.0, 20.0)
    return _scene([Shape(Circle((100.0, 110.0), 6.0)), Shape(Circle(c1, 5.0))], [g0, _fill([1])], w=100, h=80)


def test_shrinking_group_keeps_the_reference_node_test(oracle):
    """compute_distance_group (distance.rs:346-360) skips a BVH node whose
    group-local box distance exceeds the best canvas distance so far.  Group 1
    is walked first (groups from the last): with its circle 15 px away, group
    0's node (24 local units) is skipped although its circle is 12 px away;
    with group 1's circle 40 px away the node is visited and 12 wins."""
    p = np.array([[50.0, 40.0]], np.float32)
    near = oracle.eval_positions(SceneDesc(_shrunk_pair(False)), OPTS, p)
    far = oracle.eval_positions(SceneDesc(_shrunk_pair(True)), OPTS, p)
    assert abs(float(near[0]) - 15.0) < 1e-5
    assert abs(float(far[0]) - 12.0) < 1e-5


def test_group_bvh_walk_finds_the_nearest_of_many(oracle):
    """A group of 40 circles (a BVH of several levels) scaled by 1.25: the
    walk's value is the closed-form nearest boundary distance, negative inside
    any circle (non-zero fill of circles of one orientation)."""
    rng = np.random.default_rng(8)
    c = rng.uniform(5, 75, size=(40, 2))
    r = rng.uniform(1.0, 4.0, size=40)
    g = _fill(list(range(40)))
    s = 1.25
    g.set_shape_to_canvas(np.array([[s, 0, -3.0], [0, s, 2.0], [0, 0, 1]], np.float32))
    sc = _scene([Shape(Circle(tuple(ci), float(ri))) for ci, ri in zip(c, r)], [g], w=100, h=100)
    p = _pts(rng, 600, 100, 100)
    got = oracle.eval_positions(SceneDesc(sc), OPTS, p)
    cc = c * s + np.array([-3.0, 2.0])
    d = np.hypot(p[:, None, 0] - cc[None, :, 0], p[:, None, 1] - cc[None, :, 1]) - s * r[None, :]
    want = np.where((d < 0).any(1), -np.abs(d).min(1), np.abs(d).min(1))
    assert np.allclose(got, want, atol=5e-5, rtol=1e-5)
