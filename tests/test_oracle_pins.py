"""Pins of the oracle (test infrastructure) against golden vectors, the
reference's own unit tests and closed forms.  CPU only."""
import math

import numpy as np
import pytest

from pydvg import abi, scenes
from pydvg.scene import (Circle, Filter, Path, Rect, RenderOptions, Scene, SceneDesc, Shape, ShapeGroup, Solid,
                         Ellipse)

M64 = (1 << 64) - 1


def test_pcg32_matches_published_demo(oracle):
    # O'Neill's pcg32-demo: pcg32_srandom_r(&rng, 42, 54) then pcg32_random_r
    # gives 0xa15c02b7 0x7b47f409 0xba1d3330 0x83d2f293 0xbfa4784b 0xcbed606e.
    # Pcg32::new(idx, seed) (renderer/rng.rs:16-26) is that seeding with
    # initseq = idx + 1 and initstate = 0x853c49e6748fea9b + seed.
    seed = (42 - 0x853c49e6748fea9b) & M64
    u, f = oracle.pcg32(53, seed, 6)
    assert [int(x) for x in u] == [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    # next_f32 keeps the top 23 bits (rng.rs:37-42)
    assert np.all(f == ((u >> 9).astype(np.float64) / 2 ** 23).astype(np.float32))


def test_pcg32_golden_fixture(oracle):
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "pcg32_streams.npz")
    g = np.load(path)
    for idx, seed, words in zip(g["idx"], g["seed"], g["words"]):
        u, _ = oracle.pcg32(int(idx), int(seed), words.size)
        assert np.array_equal(u, words)


@pytest.mark.parametrize("roots", [(0.1, 0.5, 0.9), (-1.0, 0.25, 2.0), (0.3, 0.3, 0.7)])
def test_solve_cubic_known_roots(oracle, roots):
    a, b, c = roots
    # (t-a)(t-b)(t-c) = t^3 - (a+b+c) t^2 + (ab+bc+ca) t - abc
    co = (1.0, -(a + b + c), a * b + b * c + c * a, -a * b * c)
    r64 = oracle.solve_cubic_f64(*co)
    assert len(r64) == 3
    assert np.allclose(r64, sorted(roots), atol=1e-6)
    r32 = oracle.solve_cubic_f32(*co)
    assert np.allclose(r32, sorted(roots), atol=2e-3)


def test_solve_cubic_single_real_root(oracle):
    # (t - 0.5)(t^2 + 1): one real root
    co = (1.0, -0.5, 1.0, -0.5)
    assert np.allclose(oracle.solve_cubic_f64(*co), [0.5], atol=1e-9)
    assert np.allclose(oracle.solve_cubic_f32(*co), [0.5], atol=1e-5)


def _bezier(pts, t):
    pts = np.asarray(pts, np.float64).reshape(-1, 2)
    n = len(pts) - 1
    return sum(math.comb(n, i) * (1 - t) ** (n - i) * t ** i * pts[i] for i in range(n + 1))


@pytest.mark.parametrize("kind,npts", [(1, 3), (2, 4)])
def test_closest_point_matches_brute_force(oracle, kind, npts):
    rng = np.random.default_rng(3)
    for _ in range(40):
        pts = rng.uniform(0, 100, size=(npts, 2)).astype(np.float32)
        p = rng.uniform(-20, 120, size=2).astype(np.float32)
        ts = np.linspace(0, 1, 20001)
        curve = _bezier(pts, ts[:, None])
        d = np.sqrt(((curve - p) ** 2).sum(1))
        best = d.min()
        cp, t, dist = oracle.closest(kind, pts, p)
        assert abs(dist - best) < 1e-2 * max(1.0, best) + 1e-3
        d32, t32 = oracle.closest(kind, pts, p, f32=True)
        assert abs(d32 - best) < 2e-2 * max(1.0, best) + 2e-3


def _circle_cubics(cx, cy, r):
    # four cubic arcs approximating a circle (kappa = 0.5523)
    k = 0.5522847498 * r
    P = [(cx + r, cy), (cx + r, cy + k), (cx + k, cy + r), (cx, cy + r), (cx - k, cy + r), (cx - r, cy + k),
         (cx - r, cy), (cx - r, cy - k), (cx - k, cy - r), (cx, cy - r), (cx + k, cy - r), (cx + r, cy - k)]
    return [np.array([P[3 * i], P[3 * i + 1], P[3 * i + 2], P[(3 * i + 3) % 12]], np.float32) for i in range(4)]


def test_winding_of_circle_arcs(oracle):
    segs = _circle_cubics(10, 10, 4)
    for p, want in [((10.1, 10.2), 1), ((20, 10.3), 0), ((0.5, 10.1), 0), ((12.5, 11.0), 1), ((13.9, 13.9), 0)]:
        w64 = sum(oracle.winding(2, s, p) for s in segs)
        w32 = sum(oracle.winding(2, s, p, f32=True) for s in segs)
        assert abs(w64) == want and abs(w32) == want, (p, w64, w32)


def test_cpu_winding_counts_line_like_cubic_twice(oracle):
    # A cubic whose y(t) is linear reaches solve_cubic's |a| < 1e-6 branch;
    # the f64 solve_quadratic then returns the single root twice
    # (distance/utils.rs:44-56, 86-91), so winding.rs counts the crossing
    # twice.  The f32 GPU solver divides by a = 0 and keeps one finite root.
    seg = np.array([(14, 6), (14, 6 + 8 / 3), (14, 6 + 16 / 3), (14, 14)], np.float32)
    assert oracle.winding(2, seg, (10.0, 10.1)) == 2
    assert oracle.winding(2, seg, (10.0, 10.1), f32=True) == 1


def _line_path_scene(stroke_width):
    # distance/tests.rs:28-51: line (10,10)->(90,10), stroke width 4
    sc = Scene(100, 100)
    sc.shapes.append(Shape(Path([0], [(10, 10), (90, 10)]), stroke_width=stroke_width))
    sc.groups.append(ShapeGroup([0], fill=None, stroke=Solid((0, 0, 0, 1))))
    return SceneDesc(sc)


def test_within_distance_path_stroke_reference_kat(oracle):
    d = _line_path_scene(4.0)
    assert oracle.within_distance(d, 0, (50.0, 12.0))
    assert not oracle.within_distance(d, 0, (50.0, 20.0))


def test_within_distance_stroked_rect_reference_kat(oracle):
    # distance/tests.rs:84-104: stroked rect (w = 5), near (32, 20), far (50, 20)
    sc = Scene(100, 100)
    sc.shapes.append(Shape(Rect((10, 10), (30, 30)), stroke_width=5.0))
    sc.groups.append(ShapeGroup([0], fill=None, stroke=Solid((0, 0, 0, 1))))
    d = SceneDesc(sc)
    assert oracle.within_distance(d, 0, (32.0, 20.0))
    assert not oracle.within_distance(d, 0, (50.0, 20.0))


def test_inside_circle_and_ellipse(oracle):
    sc = Scene(100, 100)
    sc.shapes.append(Shape(Circle((50, 50), 10)))
    sc.shapes.append(Shape(Ellipse((20, 20), (8, 3))))
    sc.groups.append(ShapeGroup([0], fill=Solid((1, 0, 0, 1))))
    sc.groups.append(ShapeGroup([1], fill=Solid((1, 0, 0, 1))))
    d = SceneDesc(sc)
    assert oracle.is_inside(d, 0, (55, 55)) and not oracle.is_inside(d, 0, (58, 58))
    assert oracle.is_inside(d, 1, (27, 20)) and not oracle.is_inside(d, 1, (20, 24))


@pytest.mark.parametrize("ftype,r", [(abi.FILTER_BOX, 0.5), (abi.FILTER_TENT, 1.0),
                                     (abi.FILTER_RADIAL_PARABOLIC, 1.0), (abi.FILTER_HANN, 1.5)])
def test_filter_closed_forms(oracle, ftype, r):
    dx, dy = 0.3 * r, -0.2 * r
    w = oracle.filter_weight(ftype, r, dx, dy)
    if ftype == abi.FILTER_BOX:
        want = 1 / (2 * r) ** 2
    elif ftype == abi.FILTER_TENT:
        want = (r - abs(dx)) * (r - abs(dy)) / r ** 4
    elif ftype == abi.FILTER_RADIAL_PARABOLIC:
        want = (4 / 3) ** 2 * (1 - (dx / r) ** 2) * (1 - (dy / r) ** 2)
    else:
        want = 0.25 * (1 - math.cos(2 * math.pi * (dx / (2 * r) + 0.5))) * (1 - math.cos(2 * math.pi * (dy / (2 * r) + 0.5))) / r ** 2
    assert abs(w - want) < 1e-5 * max(1, want)
    assert oracle.filter_weight(ftype, r, 1.01 * r, 0.0) == 0.0


def _one_shape(geom, color=(0.2, 0.4, 0.6, 1.0), W=24, H=24, spp=1, stroke=None, sw=0.0):
    sc = Scene(W, H, background=(1, 1, 1, 1))
    sc.shapes.append(Shape(geom, stroke_width=sw))
    sc.groups.append(ShapeGroup([0], fill=Solid(color) if stroke is None else None,
                                stroke=Solid(stroke) if stroke is not None else None))
    return sc, RenderOptions(samples_x=spp, samples_y=spp, jitter=False, use_prefiltering=False)


def test_forward_opaque_shape_over_white(oracle):
    sc, opts = _one_shape(Rect((4.0, 4.0), (12.0, 12.0)))
    img = oracle.render(SceneDesc(sc), opts).reshape(24, 24, 4)
    assert np.allclose(img[8, 8], (0.2, 0.4, 0.6, 1.0))
    assert np.allclose(img[20, 20], (1, 1, 1, 1))


def test_forward_half_plane_coverage(oracle):
    # a rect edge at x = 10.5 through the pixel column 10: 4x4 stratified
    # unjittered samples at x = 10 + (k + 0.5)/4 -> 2 of 4 columns inside
    sc, opts = _one_shape(Rect((-5.0, -5.0), (10.5, 30.0)), color=(0, 0, 0, 1), spp=4)
    img = oracle.render(SceneDesc(sc), opts).reshape(24, 24, 4)
    assert np.allclose(img[5, 10, :3], 0.5, atol=1e-6)


def test_forward_compositing_order_and_alpha(oracle):
    # two overlapping rects: group order decides; 0.5 alpha over white
    sc = Scene(8, 8, background=(1, 1, 1, 1))
    sc.shapes += [Shape(Rect((0, 0), (8, 8))), Shape(Rect((0, 0), (8, 8)))]
    sc.groups += [ShapeGroup([0], fill=Solid((1, 0, 0, 0.5))), ShapeGroup([1], fill=Solid((0, 0, 1, 0.5)))]
    opts = RenderOptions(samples_x=1, samples_y=1, jitter=False, use_prefiltering=False)
    img = oracle.render(SceneDesc(sc), opts).reshape(8, 8, 4)
    # red over white = (1, 0.5, 0.5); blue over that = (0.5, 0.25, 0.75)
    assert np.allclose(img[3, 3], (0.5, 0.25, 0.75, 1.0), atol=1e-6)


def test_forward_box_weight_sum_is_exact(oracle):
    # Box r = 0.5 with stratified samples: every interior pixel's weight sum is
    # spp * 1/(2r)^2 = spp, so a constant scene renders to its colour exactly.
    sc, opts = _one_shape(Rect((-100.0, -100.0), (100.0, 100.0)), color=(0.3, 0.6, 0.9, 1.0), spp=2)
    opts.jitter = True
    img = oracle.render(SceneDesc(sc), opts).reshape(24, 24, 4)
    assert np.allclose(img[..., :3], (0.3, 0.6, 0.9), atol=1e-6)


def test_backward_solid_colour_gradient_is_linear(oracle):
    # d/d(fill colour) of L = <d_render, image>: for an opaque solid shape over
    # a constant background the colour gradient equals the coverage-weighted
    # sum of d_render (sampling.rs:200-216), so it is linear in d_render.
    sc, opts = scenes.small_cubic_scene(16, 16, 2, seed=5, spp=2)
    d = SceneDesc(sc)
    rng = np.random.default_rng(0)
    a = rng.normal(size=16 * 16 * 4).astype(np.float32)
    b = rng.normal(size=16 * 16 * 4).astype(np.float32)
    ga = oracle.render_backward(d, opts, a)["d_group_data"]
    gb = oracle.render_backward(d, opts, b)["d_group_data"]
    gab = oracle.render_backward(d, opts, a + 2 * b)["d_group_data"]
    assert np.allclose(gab, ga + 2 * gb, rtol=1e-4, atol=1e-4)


def test_backward_colour_gradient_matches_finite_differences(oracle):
    # L(c) = sum(w * render(c)); interior colours are smooth parameters.
    sc, opts = scenes.small_cubic_scene(16, 16, 3, seed=2, spp=2)
    rng = np.random.default_rng(1)
    w = rng.normal(size=16 * 16 * 4).astype(np.float32)
    g = oracle.render_backward(SceneDesc(sc), opts, w)["d_group_data"].reshape(-1, 16)
    h = 1e-2
    for gi in range(3):
        for ch in range(4):
            col = list(sc.groups[gi].fill.color)
            vals = []
            for s in (+1, -1):
                c2 = list(col)
                c2[ch] += s * h
                sc.groups[gi].fill = Solid(tuple(c2))
                vals.append(float(np.dot(w, oracle.render(SceneDesc(sc), opts))))
            sc.groups[gi].fill = Solid(tuple(col))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - g[gi, 8 + ch]) <= 1e-2 * max(1.0, abs(fd)), (gi, ch, fd, g[gi, 8 + ch])


def _one_group_scene(geom, colour, size=64):
    sc = Scene(size, size, background=(1.0, 1.0, 1.0, 1.0), filter=Filter(abi.FILTER_BOX, 0.5))
    sc.shapes.append(Shape(geom))
    sc.groups.append(ShapeGroup([0], fill=Solid(colour)))
    return sc


def test_boundary_gradient_of_circle_radius_is_the_area_derivative(oracle):
    # One opaque circle of colour c over white, d_render = 1 on every channel:
    # L = sum(image) = W*H*4 - (4 - sum c) * pi r^2 in expectation, so
    # dL/dr = -(4 - sum c) * 2 pi r and dL/dc_xy = 0 (the circle stays inside
    # the canvas).  Only boundary sampling (backward/boundary.rs:113-1148)
    # produces the geometric terms; a wrong pdf, normal orientation or 1/(W H)
    # normalisation fails this.
    r, c = 17.3, (0.2, 0.5, 0.8, 1.0)
    sc = _one_group_scene(Circle((31.7, 32.4), r), c)
    opts = RenderOptions(samples_x=4, samples_y=4, seed=3, jitter=True, use_prefiltering=False)
    g = oracle.render_backward(SceneDesc(sc), opts, np.ones(64 * 64 * 4, np.float32))
    dr = float(g["d_shape_params"][2])
    want = -(4.0 - sum(c)) * 2.0 * math.pi * r
    assert abs(dr - want) <= 0.01 * abs(want), (dr, want)
    # the centre derivative vanishes up to Monte Carlo noise
    assert abs(g["d_shape_params"][0]) <= 0.01 * abs(want)
    assert abs(g["d_shape_params"][1]) <= 0.01 * abs(want)


def test_boundary_gradient_of_polygon_vertices_is_the_area_derivative(oracle):
    # A closed 4-segment line path (counter-clockwise quadrilateral) of colour
    # c over white with d_render = 1: dL/dv_i = -(4 - sum c) dA/dv_i with the
    # shoelace derivative dA/dv_i = 1/2 (y_{i+1} - y_{i-1}, x_{i-1} - x_{i+1}).
    pts = np.array([[14.0, 12.0], [50.0, 17.0], [46.0, 52.0], [11.0, 45.0]], np.float32)
    c = (0.9, 0.1, 0.3, 1.0)
    sc = _one_group_scene(Path(np.zeros(4, np.uint8), pts.copy(), is_closed=True), c)
    opts = RenderOptions(samples_x=4, samples_y=4, seed=5, jitter=True, use_prefiltering=False)
    g = oracle.render_backward(SceneDesc(sc), opts, np.ones(64 * 64 * 4, np.float32))
    d = g["d_shape_points"].reshape(-1, 2)
    x, y = pts[:, 0].astype(np.float64), pts[:, 1].astype(np.float64)
    area = 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)
    sgn = 1.0 if area > 0 else -1.0
    dA = sgn * 0.5 * np.stack([np.roll(y, -1) - np.roll(y, 1), np.roll(x, 1) - np.roll(x, -1)], axis=1)
    want = -(4.0 - sum(c)) * dA
    scale = np.abs(want).max()
    assert np.abs(d - want).max() <= 0.02 * scale, (d, want)
