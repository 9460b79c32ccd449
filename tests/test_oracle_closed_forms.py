"""CPU: closed forms and exact identities that pin the oracle branches not
covered by the reference's own tests.

Boundary gradients (backward/boundary.rs:113-1148).  One opaque shape of
colour c over a white background with d_render = 1 on every channel gives
L = sum(image) = W H 4 - (4 - sum c) A in expectation (box filter r = 0.5:
the pixel coverages sum to the covered area A), so every geometric gradient
is -(4 - sum c) dA/d(parameter), with A a closed form:
  * a closed path of line, quadratic and cubic segments: Green's theorem,
    A = 1/2 sum over segments of the integral of x y' - y x' (a polynomial in
    the control points), differentiated exactly;  pins the Bernstein-weighted
    scatter of boundary points to control points and the closed-path
    wrap-around;
  * a stroked circle (the stroke width is a radius): the ring area
    pi (r + w)^2 - pi (r - w)^2 = 4 pi r w;
  * a stroked open line path: the capsule 2 w l + pi w^2 (stroke width or
    per-point thickness);
  * a rect (x1 - x0)(y1 - y0); an ellipse pi a b;
  * a shape or group transform T: A = |det T| A_local, dA/dT = A cof(T),
    translation derivative 0.
These are Monte Carlo estimates: the tolerances are a few times the
sampling noise measured over seeds.

Exact identities, per sample: with identity transforms, translating a shape
moves every control point (or the centre) by the same vector, so the sum of
a path's point gradients equals its transform's translation gradient, a
circle's or ellipse's centre gradient equals it, a rect's min + max gradient
equals it, and the group's translation gradient is the sum over its shapes.

Paints (sampling.rs:11-139 forward; backward/paint.rs:8-140 adjoint): pixel
colours of a canvas-filling rect with linear and radial gradients against
the closed-form piecewise-linear interpolation, and central finite
differences of the stop colours, stop offsets and gradient geometry.

Prefiltered coverage (forward.rs:339-403, 1267-1284): fill coverage
smoothstep((d + 1) / 2) of the signed distance, stroke coverage
S(d + r) - S(d - r), at pixel centres of known distance from a rect edge, a
polygon edge and a circle outline.
"""
import math

import numpy as np
import pytest

from pydvg import abi
from pydvg.scene import (Circle, Ellipse, Filter, GradientStop, LinearGradient, Path, RadialGradient, Rect,
                         RenderOptions, Scene, SceneDesc, Shape, ShapeGroup, Solid)

GL_X, GL_W = np.polynomial.legendre.leggauss(6)


def _one(geom, fill=None, stroke=None, sw=0.0, size=64, transform=None, group_xform=None):
    sc = Scene(size, size, background=(1.0, 1.0, 1.0, 1.0), filter=Filter(abi.FILTER_BOX, 0.5))
    sh = Shape(geom, stroke_width=sw)
    if transform is not None:
        sh.transform = np.asarray(transform, np.float32)
    sc.shapes.append(sh)
    g = ShapeGroup([0], fill=Solid(fill) if fill else None, stroke=Solid(stroke) if stroke else None)
    if group_xform is not None:
        g.set_shape_to_canvas(group_xform)
    sc.groups.append(g)
    return sc


def _ones_backward(oracle, sc, spp, seed):
    opts = RenderOptions(samples_x=spp, samples_y=spp, seed=seed, jitter=True, use_prefiltering=False)
    return oracle.render_backward(SceneDesc(sc), opts, np.ones(sc.width * sc.height * 4, np.float32))


# ---- Green's theorem area of a closed Bezier path --------------------------------
def _segments(ncp, pts):
    """Control-point index lists of each segment of a closed path
    (geometry.rs:78-89: the last segment wraps to point 0)."""
    segs, i, n = [], 0, len(pts)
    for k in ncp:
        idx = [(i + j) % n for j in range(int(k) + 2)]
        segs.append(idx)
        i += int(k) + 1
    return segs


def _bezier(P, t):
    n = len(P) - 1
    B = np.stack([math.comb(n, k) * t ** k * (1 - t) ** (n - k) for k in range(n + 1)], 0)
    dB = np.stack([math.comb(n, k) * (k * t ** max(k - 1, 0) * (1 - t) ** (n - k)
                                      - (n - k) * t ** k * (1 - t) ** max(n - k - 1, 0)) for k in range(n + 1)], 0)
    return B.T @ P, dB.T @ P


def _green_area(ncp, pts):
    """Signed area 1/2 sum of the integral of x y' - y x' (Gauss-Legendre of
    order 6 is exact for these degree-5 integrands)."""
    t = 0.5 * (GL_X + 1.0)
    a = 0.0
    for idx in _segments(ncp, pts):
        p, dp = _bezier(pts[idx], t)
        a += 0.5 * np.sum(0.5 * GL_W * (p[:, 0] * dp[:, 1] - p[:, 1] * dp[:, 0]))
    return a


def _green_area_grad(ncp, pts):
    # the area is a quadratic form in the coordinates: central differences
    # are exact (up to rounding)
    g = np.zeros_like(pts, dtype=np.float64)
    for i in range(len(pts)):
        for c in range(2):
            p1, p2 = pts.astype(np.float64).copy(), pts.astype(np.float64).copy()
            p1[i, c] += 0.5
            p2[i, c] -= 0.5
            g[i, c] = _green_area(ncp, p1) - _green_area(ncp, p2)
    return g


def test_green_area_helper_matches_polygon_and_circle():
    # the helper itself: shoelace on a polygon, and a 4-arc cubic circle
    sq = np.array([[0, 0], [3, 0], [3, 2], [0, 2]], np.float64)
    assert abs(_green_area([0, 0, 0, 0], sq) - 6.0) < 1e-12
    k = 0.5522847498
    r = 10.0
    pts = np.array([[r, 0], [r, k * r], [k * r, r], [0, r], [-k * r, r], [-r, k * r], [-r, 0], [-r, -k * r],
                    [-k * r, -r], [0, -r], [k * r, -r], [r, -k * r]], np.float64)
    assert abs(_green_area([2, 2, 2, 2], pts) / (math.pi * r * r) - 1.0) < 1e-3


def test_closed_curve_boundary_gradient_is_the_green_area_derivative(oracle):
    # cubic, quadratic, cubic, line segments of a closed, simple, clockwise loop
    ncp = np.array([2, 1, 2, 0], np.uint8)
    pts = np.array([[16, 14], [30, 6], [44, 12], [50, 22], [54, 34], [48, 48], [38, 52], [26, 54], [12, 40]],
                   np.float32)
    assert len(pts) == int(np.sum(ncp.astype(int) + 1))
    c = (0.15, 0.35, 0.55, 1.0)
    sc = _one(Path(ncp, pts.copy(), is_closed=True), fill=c)
    g = _ones_backward(oracle, sc, 6, seed=2)
    d = g["d_shape_points"].reshape(-1, 2)
    A = _green_area(ncp, pts)
    want = -(4.0 - sum(c)) * math.copysign(1.0, A) * _green_area_grad(ncp, pts)
    scale = np.abs(want).max()
    err = np.abs(d - want).max() / scale
    print(f"closed curve: area {abs(A):.1f}, max point-gradient error {err:.3g} of the largest")
    assert err <= 0.03, (d, want)
    # translation: dA/dt = 0, and the transform's linear part dA/dT = A I
    T = g["d_shape_transform"].reshape(3, 3)
    wantT = -(4.0 - sum(c)) * abs(A)
    assert abs(T[0, 2]) <= 0.03 * scale and abs(T[1, 2]) <= 0.03 * scale
    assert abs(T[0, 0] - wantT) <= 0.02 * abs(wantT) and abs(T[1, 1] - wantT) <= 0.02 * abs(wantT)
    assert abs(T[0, 1]) <= 0.02 * abs(wantT) and abs(T[1, 0]) <= 0.02 * abs(wantT)


def _rot(th, s=1.0, tx=0.0, ty=0.0):
    c, n = s * math.cos(th), s * math.sin(th)
    return np.array([[c, -n, tx], [n, c, ty], [0, 0, 1]], np.float32)


def test_transform_gradients_are_the_cofactor(oracle):
    # a circle under a rotating shape transform and a rotating group
    # transform (isometries): A = pi r^2, dA/dM_lin = A M^-T for either
    # matrix, zero translation derivative, dA/dr = 2 pi r
    r, c = 13.0, (0.7, 0.2, 0.1, 1.0)
    S, G = _rot(0.3, 1.0, 3.0, -2.0), _rot(-0.2, 1.0, 2.0, 1.0)
    sc = _one(Circle((24.0, 26.0), r), fill=c, transform=S, group_xform=G)
    g = _ones_backward(oracle, sc, 6, seed=4)
    k = -(4.0 - sum(c)) * math.pi * r * r
    for name, M in (("d_shape_transform", S), ("d_group_transform", G)):
        d = g[name].reshape(3, 3)
        want = k * np.linalg.inv(M[:2, :2].astype(np.float64)).T
        err = np.abs(d[:2, :2] - want).max() / np.abs(want).max()
        print(f"{name}: linear-part error {err:.3g}; translation {d[0, 2]:.3g}, {d[1, 2]:.3g} (of {abs(k):.3g})")
        assert err <= 0.02, (name, d, want)
        assert abs(d[0, 2]) <= 1e-3 * abs(k) and abs(d[1, 2]) <= 1e-3 * abs(k)
    want_r = -(4.0 - sum(c)) * 2 * math.pi * r
    assert abs(g["d_shape_params"][2] - want_r) <= 0.01 * abs(want_r)


def test_boundary_pdf_is_in_shape_space_arclength(oracle):
    # The reference samples boundary points with a pdf per unit of
    # shape-space arclength (backward/boundary.rs:527-540 divides by 2 pi r
    # of the untransformed circle) and applies no canvas Jacobian, so under a
    # group transform of uniform scale s its estimates are the canvas-area
    # derivatives divided by s: the transform's linear part
    # A s^2 M^-T / s, and the radius (whose boundary displacement is also
    # taken in shape space) 2 pi r s^2 / s^2.
    r, c, sc_ = 13.0, (0.7, 0.2, 0.1, 1.0), 0.8
    G = _rot(0.1, sc_, 5.0, 4.0)
    sc = _one(Circle((30.0, 31.0), r), fill=c, group_xform=G)
    g = _ones_backward(oracle, sc, 6, seed=4)
    k = -(4.0 - sum(c))
    d = g["d_group_transform"].reshape(3, 3)[:2, :2]
    want = k * math.pi * r * r * sc_ * sc_ * np.linalg.inv(G[:2, :2].astype(np.float64)).T / sc_
    err = np.abs(d - want).max() / np.abs(want).max()
    print(f"scaled group: linear-part error {err:.3g} against the 1/s form")
    assert err <= 0.02
    want_r = k * 2 * math.pi * r
    assert abs(g["d_shape_params"][2] - want_r) <= 0.01 * abs(want_r)


def test_stroked_circle_gradients(oracle):
    # The stroke is the set |dist to the circle| < w, area 4 pi r w: the
    # stroke-width derivative is 4 pi r.  The reference adds every stroke
    # sample's contribution to the radius unsigned (backward/boundary.rs:
    # 1065-1069), although the inner side's normal points inward
    # (:536-539), so its radius derivative is the sum of both sides' lengths,
    # 2 pi (r + w) + 2 pi (r - w) = 4 pi r, not the ring area's 4 pi w.
    r, w, c = 15.0, 3.0, (0.3, 0.6, 0.2, 1.0)
    sc = _one(Circle((32.3, 31.6), r), stroke=c, sw=w)
    g = _ones_backward(oracle, sc, 6, seed=6)
    k = -(4.0 - sum(c))
    dw, dr = float(g["d_shape_stroke_width"][0]), float(g["d_shape_params"][2])
    print(f"stroked circle: d/dw {dw:.4g} (want {k * 4 * math.pi * r:.4g}), d/dr {dr:.4g} "
          f"(reference form {k * 4 * math.pi * r:.4g}; the ring area's would be {k * 4 * math.pi * w:.4g})")
    assert abs(dw - k * 4 * math.pi * r) <= 0.02 * abs(k * 4 * math.pi * r)
    assert abs(dr - k * 4 * math.pi * r) <= 0.02 * abs(k * 4 * math.pi * r)
    # the centre derivative vanishes up to Monte Carlo noise (about 0.15 % of dw)
    assert abs(g["d_shape_params"][0]) <= 5e-3 * abs(dw) and abs(g["d_shape_params"][1]) <= 5e-3 * abs(dw)


@pytest.mark.parametrize("thickness", [False, True])
def test_stroked_line_width_is_the_capsule_area_derivative(oracle, thickness):
    # an open line path stroked with radius w is a capsule of area 2 w l + pi w^2
    p = np.array([[14.0, 20.0], [48.0, 41.0]], np.float32)
    w, c = 2.5, (0.1, 0.2, 0.9, 1.0)
    geom = Path(np.array([0], np.uint8), p.copy(), thickness=np.array([w, w], np.float32) if thickness else None)
    sc = _one(geom, stroke=c, sw=w)
    g = _ones_backward(oracle, sc, 8, seed=8)
    l = float(np.linalg.norm(p[1] - p[0]))
    want = -(4.0 - sum(c)) * (2 * l + 2 * math.pi * w)
    got = float(g["d_shape_thickness"].sum()) if thickness else float(g["d_shape_stroke_width"][0])
    print(f"capsule (thickness={thickness}): d/dw {got:.4g}, want {want:.4g}")
    assert abs(got - want) <= 0.02 * abs(want)
    if thickness:
        # a symmetric capsule: both end radii carry the same derivative
        t0, t1 = g["d_shape_thickness"][:2]
        assert abs(t0 - t1) <= 0.05 * abs(want)
        assert abs(g["d_shape_stroke_width"][0]) == 0.0


def test_rect_and_ellipse_parameter_gradients_are_the_area_derivatives(oracle):
    c = (0.4, 0.4, 0.1, 1.0)
    k = -(4.0 - sum(c))
    x0, y0, x1, y1 = 11.3, 14.6, 47.2, 39.9
    g = _ones_backward(oracle, _one(Rect((x0, y0), (x1, y1)), fill=c), 6, seed=10)
    d = g["d_shape_params"][:4]
    want = k * np.array([-(y1 - y0), -(x1 - x0), (y1 - y0), (x1 - x0)])
    print(f"rect: {d} want {want}")
    assert np.abs(d - want).max() <= 0.02 * np.abs(want).max()
    a, b = 19.0, 11.5
    g = _ones_backward(oracle, _one(Ellipse((31.0, 33.0), (a, b)), fill=c), 6, seed=12)
    d = g["d_shape_params"][:4]
    want = k * np.array([0.0, 0.0, math.pi * b, math.pi * a])
    print(f"ellipse: {d} want {want}")
    assert np.abs(d - want).max() <= 0.02 * np.abs(want).max()


# ---- exact per-sample identities -------------------------------------------------
def _identity_scene():
    sc = Scene(64, 56, background=(1.0, 1.0, 1.0, 1.0), filter=Filter(abi.FILTER_TENT, 1.0))
    k = 0.5522847498
    circ = np.array([[40, 20], [40, 20 + 9 * k], [31 + 9 * k, 29], [31, 29], [31 - 9 * k, 29], [22, 20 + 9 * k],
                     [22, 20], [22, 20 - 9 * k], [31 - 9 * k, 11], [31, 11], [31 + 9 * k, 11], [40, 20 - 9 * k]],
                    np.float32)
    sc.shapes += [
        Shape(Path(np.array([2, 2, 2, 2], np.uint8), circ, is_closed=True), stroke_width=1.5),
        Shape(Path(np.array([1, 0, 2], np.uint8), np.array([[4, 40], [14, 30], [20, 46], [28, 50], [34, 36],
                                                               [42, 54], [50, 44]], np.float32)), stroke_width=2.0),
        Shape(Circle((50.0, 14.0), 7.0), stroke_width=1.0),
        Shape(Rect((6.0, 4.0), (18.0, 22.0))),
        Shape(Ellipse((52.0, 38.0), (8.0, 5.0))),
    ]
    sc.groups += [
        ShapeGroup([0], fill=Solid((0.8, 0.2, 0.3, 0.7)), stroke=Solid((0.1, 0.1, 0.6, 0.9))),
        ShapeGroup([1], stroke=Solid((0.2, 0.7, 0.2, 1.0))),
        ShapeGroup([2, 3], fill=Solid((0.9, 0.6, 0.1, 0.6)), stroke=Solid((0.3, 0.3, 0.3, 1.0))),
        ShapeGroup([4], fill=Solid((0.2, 0.5, 0.9, 0.8))),
    ]
    return sc


@pytest.mark.parametrize("prefiltered", [False, True])
def test_translation_identities_hold_per_sample(oracle, prefiltered):
    sc = _identity_scene()
    opts = (RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=True) if prefiltered
            else RenderOptions(samples_x=3, samples_y=3, seed=1, jitter=True, use_prefiltering=False))
    rng = np.random.default_rng(3)
    dr = rng.normal(size=sc.width * sc.height * 4).astype(np.float32)
    g = oracle.render_backward(SceneDesc(sc), opts, dr)
    T = g["d_shape_transform"].reshape(-1, 3, 3)
    tr = T[:, :2, 2].astype(np.float64)
    pts = g["d_shape_points"].reshape(-1, 2).astype(np.float64)
    prm = g["d_shape_params"].reshape(-1, 8).astype(np.float64)
    scale = np.abs(tr).max()
    assert scale > 0
    # per sample exact up to f32 rounding; in the prefiltered adjoint the
    # closest point is a Newton root (|f| < 1e-5), and the implicit
    # derivative of its parameter leaves a residual of that order
    tol = (3e-3 if prefiltered else 2e-4) * scale
    np.testing.assert_allclose(pts[:12].sum(0), tr[0], atol=tol)
    np.testing.assert_allclose(pts[12:19].sum(0), tr[1], atol=tol)
    np.testing.assert_allclose(prm[2, :2], tr[2], atol=tol)
    np.testing.assert_allclose(prm[3, 0:2] + prm[3, 2:4], tr[3], atol=tol)
    if not prefiltered:
        # the prefiltered ellipse adjoint keeps the reference's omission of
        # the direct centre term (backward/distance.rs:489-561)
        np.testing.assert_allclose(prm[4, :2], tr[4], atol=tol)
    G = g["d_group_transform"].reshape(-1, 3, 3)[:, :2, 2].astype(np.float64)
    np.testing.assert_allclose(G[0], tr[0], atol=tol)
    np.testing.assert_allclose(G[2], tr[2] + tr[3], atol=tol)
    np.testing.assert_allclose(G[3], tr[4], atol=tol)


# ---- paints ------------------------------------------------------------------------
def _stops(offs, cols):
    return [GradientStop(float(o), tuple(float(x) for x in col)) for o, col in zip(offs, cols)]


def _interp(t, offs, cols):
    """Piecewise-linear stop interpolation (sampling.rs:47-139 away from the
    clamps): the first colour below the first stop, the last above the last."""
    cols = np.asarray(cols, np.float64)
    out = np.empty(t.shape + (4,))
    for i, tv in enumerate(t.ravel()):
        if tv <= offs[0]:
            out.flat[i * 4:i * 4 + 4] = cols[0]
        elif tv >= offs[-1]:
            out.flat[i * 4:i * 4 + 4] = cols[-1]
        else:
            j = int(np.searchsorted(offs, tv, side="right")) - 1
            u = (tv - offs[j]) / (offs[j + 1] - offs[j])
            out.flat[i * 4:i * 4 + 4] = cols[j] * (1 - u) + cols[j + 1] * u
    return out


OFFS = [0.1, 0.45, 0.9]
COLS = [(0.9, 0.1, 0.2, 1.0), (0.2, 0.8, 0.3, 1.0), (0.1, 0.3, 0.95, 1.0)]


def _paint_scene(paint, W=48, H=40):
    sc = Scene(W, H, background=(1.0, 1.0, 1.0, 1.0), filter=Filter(abi.FILTER_BOX, 0.5))
    sc.shapes.append(Shape(Rect((-10.0, -10.0), (W + 10.0, H + 10.0))))
    sc.groups.append(ShapeGroup([0], fill=paint))
    return sc


def test_gradient_paint_colours_are_the_closed_form(oracle):
    W, H = 48, 40
    opts = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=False)
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    s, e = np.array([6.0, 9.0]), np.array([41.0, 30.0])
    lin = LinearGradient(tuple(s), tuple(e), _stops(OFFS, COLS))
    img = oracle.render(SceneDesc(_paint_scene(lin)), opts).reshape(H, W, 4)
    t = ((x - s[0]) * (e[0] - s[0]) + (y - s[1]) * (e[1] - s[1])) / np.dot(e - s, e - s)
    want = _interp(t, OFFS, COLS)
    print(f"linear: max |diff| {np.abs(img - want).max():.3g}")
    assert np.abs(img - want).max() <= 1e-5
    cc, rr = np.array([20.0, 18.0]), np.array([26.0, 15.0])
    rad = RadialGradient(tuple(cc), tuple(rr), _stops(OFFS, COLS))
    img = oracle.render(SceneDesc(_paint_scene(rad)), opts).reshape(H, W, 4)
    t = np.hypot((x - cc[0]) / rr[0], (y - cc[1]) / rr[1])
    want = _interp(t, OFFS, COLS)
    print(f"radial: max |diff| {np.abs(img - want).max():.3g}")
    assert np.abs(img - want).max() <= 1e-5


def _paint_t(kind, geo, W, H):
    """t of every pixel centre: the linear projection or the radial norm
    (used only to choose difference steps that no sample's t crosses a stop
    within)."""
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    if kind == "linear":
        ex, ey = geo[2] - geo[0], geo[3] - geo[1]
        return ((x - geo[0]) * ex + (y - geo[1]) * ey) / (ex * ex + ey * ey)
    return np.hypot((x - geo[0]) / geo[2], (y - geo[1]) / geo[3])


@pytest.mark.parametrize("kind", ["linear", "radial"])
def test_gradient_paint_adjoint_matches_finite_differences(oracle, kind):
    # One unjittered sample per pixel centre, so the image is a smooth
    # function of every paint parameter as long as no sample's t crosses a
    # stop offset: each step below is chosen so that none does, which makes
    # the central differences accurate to O(h^2) and f32 rounding.
    W, H = 48, 40
    opts = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=False)
    cols = [(0.9, 0.1, 0.2, 0.8), (0.2, 0.8, 0.3, 0.6), (0.1, 0.3, 0.95, 0.9)]
    geo = [6.0, 9.0, 41.0, 30.0] if kind == "linear" else [20.0, 18.0, 26.0, 15.0]

    def make(geo, offs, cols):
        st = _stops(offs, cols)
        p = (LinearGradient((geo[0], geo[1]), (geo[2], geo[3]), st) if kind == "linear"
             else RadialGradient((geo[0], geo[1]), (geo[2], geo[3]), st))
        return _paint_scene(p, W, H)

    rng = np.random.default_rng(11)
    w = rng.normal(size=W * H * 4).astype(np.float32)
    g = oracle.render_backward(SceneDesc(make(geo, OFFS, cols)), opts, w)

    def L(geo_, offs_, cols_):
        return float(np.dot(w.astype(np.float64), oracle.render(SceneDesc(make(geo_, offs_, cols_)), opts)))

    t0 = _paint_t(kind, geo, W, H)
    gap = np.abs(t0[..., None] - np.asarray(OFFS)).min()
    assert gap > 1e-5
    checks = []
    for j in range(4):  # start/end or centre/radius
        e = np.zeros(4)
        e[j] = 1e-3
        rate = np.abs(_paint_t(kind, np.add(geo, e), W, H) - t0).max() / 1e-3
        hg = 0.4 * gap / max(rate, 1e-9)
        gp, gm = list(np.add(geo, hg * (e > 0))), list(np.subtract(geo, hg * (e > 0)))
        checks.append((f"geometry[{j}]", (L(gp, OFFS, cols) - L(gm, OFFS, cols)) / (2 * hg),
                       float(g["d_gradient_data"][1 + j])))
    ho = 0.4 * gap
    for i in range(3):
        op, om = list(OFFS), list(OFFS)
        op[i] += ho
        om[i] -= ho
        checks.append((f"offset[{i}]", (L(geo, op, cols) - L(geo, om, cols)) / (2 * ho), float(g["d_stop_offsets"][i])))
        for ch in range(4):  # the image is linear in the stop colours
            cp = [list(c) for c in cols]
            cm = [list(c) for c in cols]
            cp[i][ch] += 1e-2
            cm[i][ch] -= 1e-2
            checks.append((f"colour[{i}][{ch}]", (L(geo, OFFS, cp) - L(geo, OFFS, cm)) / 2e-2,
                           float(g["d_stop_colors"][4 * i + ch])))
    big = max(abs(fd) for _, fd, _ in checks)
    for name, fd, an in checks:
        print(f"{kind} {name}: fd {fd:.6g} adjoint {an:.6g}")
        assert abs(fd - an) <= 1e-2 * max(abs(fd), 0.02 * big), (name, fd, an)


# ---- prefiltered coverage ----------------------------------------------------------
def _smooth(d):
    t = np.clip((d + 1.0) * 0.5, 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def test_prefiltered_coverage_is_the_smoothstep_of_the_distance(oracle):
    W, H = 40, 32
    opts = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=True)
    xe = 20.3
    y, x = np.mgrid[0:H, 0:W].astype(np.float64) + 0.5
    c = np.array([0.2, 0.5, 0.7])

    def check(sc, cov, label):
        img = oracle.render(SceneDesc(sc), opts).reshape(H, W, 4)
        want = 1.0 - cov[..., None] * (1.0 - c)
        err = np.abs(img[..., :3] - want).max()
        print(f"{label}: max |diff| {err:.3g}, partially covered pixels {int(((cov > 0) & (cov < 1)).sum())}")
        assert err <= 1e-5
        assert ((cov > 0.05) & (cov < 0.95)).sum() >= H

    # a rect edge at x = xe (the other edges far outside the canvas)
    sc = Scene(W, H, background=(1.0, 1.0, 1.0, 1.0), shapes=[Shape(Rect((-50.0, -50.0), (xe, 100.0)))],
               groups=[ShapeGroup([0], fill=Solid((*c, 1.0)))])
    check(sc, _smooth(xe - x), "rect edge")
    # a polygon path with one edge crossing the canvas: the line x = xe + (y - 16) / 8
    pts = np.array([[-60.0, -40.0], [xe - 7.0, -40.0], [xe + 10.5, 100.0], [-60.0, 100.0]], np.float32)
    sc = Scene(W, H, background=(1.0, 1.0, 1.0, 1.0),
               shapes=[Shape(Path(np.zeros(4, np.uint8), pts, is_closed=True))],
               groups=[ShapeGroup([0], fill=Solid((*c, 1.0)))])
    a, b = pts[1].astype(np.float64), pts[2].astype(np.float64)
    n = np.array([b[1] - a[1], -(b[0] - a[0])]) / np.linalg.norm(b - a)
    d = -((x - a[0]) * n[0] + (y - a[1]) * n[1])  # positive inside (left of the edge)
    check(sc, _smooth(d), "polygon edge")
    # a stroked circle outline: coverage S(d + r) - S(d - r), d the distance to the outline
    R, r = 11.0, 1.75
    sc = Scene(W, H, background=(1.0, 1.0, 1.0, 1.0), shapes=[Shape(Circle((19.6, 16.2), R), stroke_width=r)],
               groups=[ShapeGroup([0], stroke=Solid((*c, 1.0)))])
    dd = np.abs(np.hypot(x - 19.6, y - 16.2) - R)
    check(sc, _smooth(dd + r) - _smooth(dd - r), "stroked circle")
