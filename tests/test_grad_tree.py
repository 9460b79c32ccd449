"""CPU: the SceneGrad tree (src/grad.rs) and apply_gpu_grads, which unpacks the
14 flat gradient arrays into it (src/renderer/backward_gpu.rs:599-864).

The flat arrays are filled with distinct values, so every leaf of the tree
must equal the one flat element the packed layout assigns it; the tree of a
sum of two flat gradients must equal the accumulate_from of the two trees;
to_diffvg_layout must carry the scene's path metadata.
"""
import numpy as np

from parity_util import tree_leaves
from pydvg import abi
from pydvg.grad import DCircle, DEllipse, DLinearGradient, DPath, DRadialGradient, DRect, DSolid, SceneGrad
from pydvg.scene import (Circle, Ellipse, FlatGrads, GradientStop, LinearGradient, Path, RadialGradient, Rect,
                         Scene, SceneDesc, Shape, ShapeGroup, Solid)


def _scene():
    sc = Scene(16, 12, background=(1.0, 1.0, 1.0, 1.0))
    sc.background_image = np.full((12, 16, 4), 0.5, np.float32)
    pts = np.arange(14, dtype=np.float32).reshape(7, 2)
    sc.shapes = [
        Shape(Circle((3.0, 4.0), 2.0)),
        Shape(Path(np.array([2, 2], np.uint8), pts, thickness=np.linspace(1, 2, 7).astype(np.float32),
                   is_closed=False), stroke_width=1.5),
        Shape(Ellipse((5.0, 5.0), (2.0, 3.0))),
        Shape(Rect((1.0, 1.0), (4.0, 6.0))),
        Shape(Path(np.array([1, 0], np.uint8), pts[:4].copy(), is_closed=True)),
    ]
    lin = LinearGradient((0.0, 0.0), (8.0, 8.0), [GradientStop(0.0, (1, 0, 0, 1)), GradientStop(1.0, (0, 0, 1, 1))])
    rad = RadialGradient((4.0, 4.0), (3.0, 2.0), [GradientStop(0.0, (1, 1, 0, 1)), GradientStop(0.5, (0, 1, 0, 1)),
                                                  GradientStop(1.0, (0, 1, 1, 1))])
    sc.groups = [
        ShapeGroup([0], fill=Solid((1, 0, 0, 1)), stroke=lin),
        ShapeGroup([1, 4], fill=rad, stroke=Solid((0, 0, 1, 0.5))),
        ShapeGroup([2], fill=lin),
        ShapeGroup([3], stroke=Solid((0.2, 0.3, 0.4, 1))),
    ]
    return sc


def _filled(desc, seed):
    g = FlatGrads(desc, translation=True)
    rng = np.random.default_rng(seed)
    for k, a in g.arrays.items():
        a[:] = rng.uniform(-1, 1, size=a.size).astype(np.float32)
    return g


def test_apply_gpu_grads_places_every_field():
    sc = _scene()
    desc = SceneDesc(sc)
    f = _filled(desc, 0)
    t = f.scene_grad()
    P = f["d_shape_params"]
    c = t.shapes[0].geometry
    assert isinstance(c, DCircle)
    assert np.array_equal(c.center, P[0:2]) and c.radius == P[2]
    e = t.shapes[2].geometry
    assert isinstance(e, DEllipse)
    assert np.array_equal(e.center, P[16:18]) and np.array_equal(e.radius, P[18:20])
    r = t.shapes[3].geometry
    assert isinstance(r, DRect)
    assert np.array_equal(r.min, P[24:26]) and np.array_equal(r.max, P[26:28])
    for i in (1, 4):
        p = t.shapes[i].geometry
        assert isinstance(p, DPath)
        o, n = int(desc.p_off[i]), int(desc.p_cnt[i])
        assert np.array_equal(p.points, f["d_shape_points"][2 * o:2 * (o + n)].reshape(n, 2))
    th = t.shapes[1].geometry.thickness
    assert np.array_equal(th, f["d_shape_thickness"][int(desc.t_off[1]):int(desc.t_off[1]) + 7])
    assert t.shapes[4].geometry.thickness is None
    for i in range(5):
        assert np.array_equal(t.shapes[i].transform, f["d_shape_transform"][9 * i:9 * i + 9].reshape(3, 3))
        assert t.shapes[i].stroke_width == f["d_shape_stroke_width"][i]
    D = f["d_group_data"]
    for gi in range(4):
        assert np.array_equal(t.shape_groups[gi].shape_to_canvas, f["d_group_transform"][9 * gi:9 * gi + 9].reshape(3, 3))
    assert isinstance(t.shape_groups[0].fill, DSolid) and t.shape_groups[0].fill.color == tuple(D[8:12])
    assert isinstance(t.shape_groups[1].stroke, DSolid) and t.shape_groups[1].stroke.color == tuple(D[16 + 12:16 + 16])
    assert t.shape_groups[2].stroke is None and t.shape_groups[3].fill is None
    # gradient paints: geometry at d_gradient_data[8 k + 1 ..], stops at the
    # gradient's stop offset, in the packer's order (fill before stroke)
    for gi, which, cls in ((0, "stroke", DLinearGradient), (1, "fill", DRadialGradient), (2, "fill", DLinearGradient)):
        paint = getattr(t.shape_groups[gi], which)
        k = int((desc.stroke_gradient if which == "stroke" else desc.fill_gradient)[gi])
        assert isinstance(paint, cls)
        gd = f["d_gradient_data"][8 * k + 1:8 * k + 5]
        a, b = (paint.start, paint.end) if cls is DLinearGradient else (paint.center, paint.radius)
        assert np.array_equal(a, gd[0:2]) and np.array_equal(b, gd[2:4])
        so = int(desc.gradient_stop_offset[k])
        assert len(paint.stops) == int(desc.gradient_stop_count[k])
        for j, st in enumerate(paint.stops):
            assert st.offset == f["d_stop_offsets"][so + j]
            assert st.color == tuple(f["d_stop_colors"][4 * (so + j):4 * (so + j) + 4])
    assert t.filter.radius == f["d_filter_radius"][0]
    assert t.background == tuple(f["d_background"])
    assert np.array_equal(t.background_image, f["d_background_image"])
    assert np.array_equal(t.translation, f["d_translation"])


def test_accumulate_from_adds_the_flat_gradients():
    sc = _scene()
    desc = SceneDesc(sc)
    a, b = _filled(desc, 1), _filled(desc, 2)
    s = FlatGrads(desc, translation=True)
    for k in s.arrays:
        s.arrays[k][:] = a.arrays[k] + b.arrays[k]
    ta, tb = a.scene_grad(), b.scene_grad()
    ta.accumulate_from(tb)
    np.testing.assert_allclose(tree_leaves(ta), tree_leaves(s.scene_grad()), rtol=0, atol=1e-6)
    # a zero tree plus a tree is that tree
    z = SceneGrad.zeros_from_scene(sc, True, True)
    z.accumulate_from(tb)
    assert np.array_equal(tree_leaves(z), tree_leaves(tb))


def test_accumulate_from_takes_missing_optional_parts():
    sc = _scene()
    desc = SceneDesc(sc)
    tb = _filled(desc, 3).scene_grad()
    z = SceneGrad.zeros_from_scene(sc, False, False)
    assert z.background_image is None and z.translation is None
    z.accumulate_from(tb)
    assert np.array_equal(z.background_image, tb.background_image)
    assert np.array_equal(z.translation, tb.translation)


def test_to_diffvg_layout_carries_path_metadata():
    sc = _scene()
    desc = SceneDesc(sc)
    t = _filled(desc, 4).scene_grad()
    d = t.to_diffvg_layout(sc)
    p1, p4 = d.shapes[1].geometry, d.shapes[4].geometry
    assert np.array_equal(p1.num_control_points, sc.shapes[1].geometry.num_control_points)
    assert p1.is_closed is False and p4.is_closed is True
    assert np.array_equal(p1.points, t.shapes[1].geometry.points)
    assert np.array_equal(p1.thickness, t.shapes[1].geometry.thickness)
    assert p4.thickness is None
    c0, c1 = d.shapes[0].geometry, t.shapes[0].geometry
    assert np.array_equal(c0.center, c1.center) and c0.radius == c1.radius
    assert d.background == t.background and d.filter.radius == t.filter.radius


def test_zeros_from_scene_mirrors_the_scene():
    sc = _scene()
    z = SceneGrad.zeros_from_scene(sc, True, True)
    assert [type(s.geometry) for s in z.shapes] == [DCircle, DPath, DEllipse, DRect, DPath]
    assert z.shapes[1].geometry.points.shape == (7, 2) and z.shapes[1].geometry.thickness.shape == (7,)
    assert z.background_image.size == 16 * 12 * 4 and z.translation.size == 16 * 12 * 2
    assert isinstance(z.shape_groups[1].fill, DRadialGradient) and len(z.shape_groups[1].fill.stops) == 3
    assert z.shape_groups[3].fill is None
    assert abi.PAINT_NONE == 0
