"""CPU checks of the packer's BVHs and flattened stroke segments, the parts of
PreparedScene that only the reference's kernels read (prepare.rs:114-320,
renderer/bvh.rs, geometry.rs:218-476).

  * Hand-built cases fix the BVH layout: the median split on the longer axis,
    leaves of at most 8, preorder nodes, skip pointers and the stable order of
    equal centres.
  * Structural invariants on every packed BVH: node bounds are the union of
    their items, a skip-pointer walk visits every node once in preorder, and
    the leaves hold each item once.
  * Flattening against the curves themselves, in float64 from the Bernstein
    form (not de Casteljau): the recursion's leaves are the dyadic intervals
    whose sub-curve first passes the flatness test, segment ends lie on the
    curve, radii follow the Bernstein interpolation of the point radii.
  * The packed buffers equal oracle/oracle_pack.py's float32 restatement
    element for element on mixed scenes, scaled groups and thick strokes.
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

from pydvg import scenes
from pydvg.renderer import PreparedScene
from pydvg.scene import Circle, Path, RenderOptions, Scene, Shape, ShapeGroup, Solid

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_pack as op  # noqa: E402

NONE = 0xFFFFFFFF


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype).copy()


def _pack(sc, opts):
    p = PreparedScene(sc, opts)
    pk = p.pack
    n, G = pk.num_prepared_shapes, pk.num_groups
    d = {
        "shape_data": _arr(pk.shape_data, 15 * n, np.float32).reshape(-1, 15),
        "segment_data": _arr(pk.segment_data, pk.segment_data_len, np.float32),
        "curve_data": _arr(pk.curve_data, pk.curve_data_len, np.float32).reshape(-1, 13),
        "shape_bounds": _arr(pk.shape_bounds, 4 * n, np.float32).reshape(-1, 4),
        "group_data": _arr(pk.group_data, 16 * G, np.float32).reshape(-1, 16),
        "group_shapes": _arr(pk.group_shapes, pk.group_shapes_len, np.float32),
        "gb": _arr(pk.group_bvh_bounds, pk.group_bvh_bounds_len, np.float32),
        "gn": _arr(pk.group_bvh_nodes, pk.group_bvh_nodes_len, np.uint32),
        "gi": _arr(pk.group_bvh_indices, pk.group_bvh_indices_len, np.uint32),
        "gm": _arr(pk.group_bvh_meta, 4 * G, np.uint32).reshape(-1, 4),
        "pb": _arr(pk.path_bvh_bounds, pk.path_bvh_bounds_len, np.float32),
        "pn": _arr(pk.path_bvh_nodes, pk.path_bvh_nodes_len, np.uint32),
        "pi": _arr(pk.path_bvh_indices, pk.path_bvh_indices_len, np.uint32),
        "pm": _arr(pk.path_bvh_meta, 4 * n, np.uint32).reshape(-1, 4),
    }
    del p
    return d


def _bvh(d, kind, k):
    meta = d[kind + "m"][k]
    nodes = d[kind + "n"].reshape(-1, 4)[meta[0]:meta[0] + meta[1]]
    bounds = d[kind + "b"].reshape(-1, 4)[meta[0]:meta[0] + meta[1]]
    idx = d[kind + "i"][meta[2]:meta[2] + meta[3]]
    return meta, nodes, bounds, idx


def _row_scene(xs, ys=None, r=2.0):
    sc = Scene(256, 64)
    ys = ys if ys is not None else [32.0] * len(xs)
    for x, y in zip(xs, ys):
        sc.shapes.append(Shape(Circle((x, y), r)))
    sc.groups.append(ShapeGroup(list(range(len(xs))), fill=Solid((1, 0, 0, 1))))
    return sc


def test_bvh_of_ten_boxes_in_a_row():
    xs = [95.0, 5.0, 55.0, 25.0, 85.0, 15.0, 65.0, 45.0, 35.0, 75.0]
    d = _pack(_row_scene(xs), RenderOptions(use_prefiltering=False))
    meta, nodes, bounds, idx = _bvh(d, "g", 0)
    # 10 > 8: root split at 10 // 2 into two leaves of 5 (x extent 94 > 4)
    assert list(meta) == [0, 3, 0, 10]
    assert [list(n) for n in nodes] == [[1, NONE, 0, 0], [NONE, 2, 0, 5], [NONE, NONE, 5, 5]]
    assert list(idx) == sorted(range(10), key=lambda i: xs[i])
    assert np.allclose(bounds[0], [3.0, 30.0, 97.0, 34.0])
    assert np.allclose(bounds[1], [3.0, 30.0, 47.0, 34.0])
    assert np.allclose(bounds[2], [53.0, 30.0, 97.0, 34.0])


def test_bvh_of_seventeen_boxes_on_the_longer_axis_with_stable_ties():
    # a column: y extent wins; centres 0..7 twice plus one, equal centres
    # keep their input order (sort_by is stable)
    ys = [float(4 + 6 * (i % 8)) for i in range(16)] + [52.0]
    xs = [8.0 + (i % 3) for i in range(17)]
    d = _pack(_row_scene(xs, ys, r=1.0), RenderOptions(use_prefiltering=False))
    meta, nodes, _, idx = _bvh(d, "g", 0)
    # root 17 -> 8 | 9, the 9 -> 4 | 5: nodes root, leaf(8), inner, leaf(4),
    # leaf(5); an inner node keeps its start with count 0
    assert list(meta) == [0, 5, 0, 17]
    assert [list(n) for n in nodes] == [[1, NONE, 0, 0], [NONE, 2, 0, 8], [3, NONE, 8, 0],
                                        [NONE, 4, 8, 4], [NONE, NONE, 12, 5]]
    expect = sorted(range(17), key=lambda i: ys[i])
    assert list(idx[:8]) == expect[:8]
    # the right half is sorted again on its own longer axis (y: extent 24 > 2)
    assert list(idx[8:]) == sorted(expect[8:], key=lambda i: ys[i])


def _check_bvh(meta, nodes, bounds, idx, item_boxes, items):
    """item_boxes: id -> (4,) box; items: the ids the BVH must hold."""
    if len(items) == 0:
        assert list(meta) == [0, 0, 0, 0]
        return
    assert sorted(int(i) for i in idx) == sorted(int(i) for i in items)
    # skip-pointer walk entering every node: preorder, every node once
    seen, nid = [], 0
    while nid != NONE:
        seen.append(nid)
        left, skip, start, count = (int(v) for v in nodes[nid])
        nid = left if count == 0 else skip
    assert seen == list(range(len(nodes)))
    for k, (left, skip, start, count) in enumerate(nodes):
        if count == 0:
            continue
        assert count <= 8
        boxes = np.array([item_boxes[int(i)] for i in idx[start:start + count]])
        assert np.array_equal(bounds[k], np.concatenate([boxes[:, :2].min(0), boxes[:, 2:].max(0)]))
    # interior bounds: the union of the leaves below (subtree = [k, skip))
    for k, (left, skip, start, count) in enumerate(nodes):
        if count:
            continue
        end = int(skip) if skip != NONE else len(nodes)
        sub = [j for j in range(k + 1, end) if nodes[j][3] > 0]
        lo = np.min([bounds[j][:2] for j in sub], 0)
        hi = np.max([bounds[j][2:] for j in sub], 0)
        assert np.array_equal(bounds[k], np.concatenate([lo, hi]))


def _scaled_multi_group_scene(seed=3):
    rng = np.random.default_rng(seed)
    sc = Scene(96, 96)
    base = scenes.mixed_scene(96, 96, 60, seed)
    sc.shapes = base.shapes
    cuts = [0, 13, 30, 31, 60]
    scales = [0.5, 1.0, 2.0, 0.75]
    for k in range(4):
        g = ShapeGroup(list(range(cuts[k], cuts[k + 1])), fill=Solid((0.2, 0.4, 0.6, 1.0)))
        c, s = np.cos(0.3 * k), np.sin(0.3 * k)
        g.set_shape_to_canvas(np.array([[scales[k] * c, -scales[k] * s, 10.0 * k],
                                        [scales[k] * s, scales[k] * c, 5.0], [0, 0, 1]], np.float32))
        sc.groups.append(g)
    for sh in sc.shapes[::7]:
        sh.transform = np.array([[1.3, 0.2, 1.0], [-0.1, 0.8, 2.0], [0, 0, 1]], np.float32)
    # a thick open path and an empty one
    th = Path(np.array([2, 1, 0], np.uint8), rng.uniform(10, 80, size=(7, 2)).astype(np.float32),
              thickness=rng.uniform(0.5, 4.0, size=7).astype(np.float32))
    sc.shapes.append(Shape(th, stroke_width=2.0))
    sc.shapes.append(Shape(Path(np.zeros(0, np.uint8), np.zeros((0, 2), np.float32)), stroke_width=1.0))
    sc.groups.append(ShapeGroup([len(sc.shapes) - 2, len(sc.shapes) - 1], stroke=Solid((0, 0, 0, 1))))
    return sc


CASES = {
    "mixed": lambda: (scenes.mixed_scene(128, 128, 80, 11), RenderOptions(use_prefiltering=False)),
    "mixed_prefiltered": lambda: (scenes.mixed_scene(96, 96, 40, 5), RenderOptions(use_prefiltering=True)),
    "scaled_groups": lambda: (_scaled_multi_group_scene(), RenderOptions(use_prefiltering=False, path_tolerance=0.2)),
    "cubic": lambda: scenes.small_cubic_scene(64, 64, 12, seed=2, spp=2),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_packed_bvhs_hold_their_invariants(case):
    sc, opts = CASES[case]()
    d = _pack(sc, opts)
    sb = d["shape_bounds"]
    for g in range(len(sc.groups)):
        off, cnt = int(d["group_data"][g, 0]), int(d["group_data"][g, 1])
        preps = d["group_shapes"][off:off + cnt].astype(np.int64)
        valid = [p for p in preps if sb[p, 0] <= sb[p, 2] and sb[p, 1] <= sb[p, 3]]
        _check_bvh(*_bvh(d, "g", g), {int(p): sb[p] for p in valid}, valid)
    cd = d["curve_data"]
    for p in range(d["shape_data"].shape[0]):
        co, cc = int(d["shape_data"][p, 12]), int(d["shape_data"][p, 13])
        boxes = {co + k: np.array(op.curve_box(cd[co + k]), np.float32) for k in range(cc)}
        _check_bvh(*_bvh(d, "p", p), boxes, list(range(co, co + cc)))


@pytest.mark.parametrize("case", sorted(CASES))
def test_packed_bvhs_and_segments_equal_the_restatement(case):
    sc, opts = CASES[case]()
    d = _pack(sc, opts)
    sb, cd, sd = d["shape_bounds"], d["curve_data"], d["shape_data"]
    gout = {"bounds": [], "nodes": [], "indices": []}
    pout = {"bounds": [], "nodes": [], "indices": []}
    gmeta, pmeta, segs = [], [], []
    prep = 0
    for gi, g in enumerate(sc.groups):
        boxes, ids = [], []
        for si in g.shape_indices:
            sh = sc.shapes[si]
            if isinstance(sh.geometry, Path):
                flat = op.flatten_with_thickness(sh.geometry, op.shape_tolerance(opts, g, sh), sh.stroke_width)
                assert int(sd[prep, 1]) == len(segs) // 12 and int(sd[prep, 2]) == len(flat), (gi, si)
                segs.extend(op.segments_to_f32(flat).tolist())
                co, cc = int(sd[prep, 12]), int(sd[prep, 13])
                pmeta.append(op.pack_bvh([op.curve_box(cd[co + k]) for k in range(cc)],
                                         list(range(co, co + cc)), pout))
            else:
                assert sd[prep, 1] == 0 and sd[prep, 2] == 0
                pmeta.append([0, 0, 0, 0])
            if sb[prep, 0] <= sb[prep, 2] and sb[prep, 1] <= sb[prep, 3]:
                boxes.append(tuple(sb[prep]))
                ids.append(prep)
            prep += 1
        gmeta.append(op.pack_bvh(boxes, ids, gout))
    assert np.array_equal(d["segment_data"], np.asarray(segs, np.float32))
    assert np.array_equal(d["gm"], np.asarray(gmeta, np.uint32).reshape(-1, 4))
    assert np.array_equal(d["gb"], np.asarray(gout["bounds"], np.float32))
    assert np.array_equal(d["gn"], np.asarray(gout["nodes"], np.uint32))
    assert np.array_equal(d["gi"], np.asarray(gout["indices"], np.uint32))
    assert np.array_equal(d["pm"], np.asarray(pmeta, np.uint32).reshape(-1, 4))
    assert np.array_equal(d["pb"], np.asarray(pout["bounds"], np.float32))
    assert np.array_equal(d["pn"], np.asarray(pout["nodes"], np.uint32))
    assert np.array_equal(d["pi"], np.asarray(pout["indices"], np.uint32))


# ---- flattening against the curves themselves ----------------------------------

def _bez(P, t):
    n = len(P) - 1
    from math import comb
    return sum(comb(n, k) * (1 - t) ** (n - k) * t ** k * P[k] for k in range(n + 1))


def _dbez(P, t):
    n = len(P) - 1
    D = [n * (P[k + 1] - P[k]) for k in range(n)]
    return _bez(D, t) if n > 1 else D[0]


def _sub(P, a, b):
    """Control points of the Bezier P restricted to [a, b] (Hermite form)."""
    h = b - a
    if len(P) == 3:
        return [_bez(P, a), _bez(P, a) + 0.5 * h * _dbez(P, a), _bez(P, b)]
    return [_bez(P, a), _bez(P, a) + h / 3.0 * _dbez(P, a), _bez(P, b) - h / 3.0 * _dbez(P, b), _bez(P, b)]


def _line_dist(p, a, b):
    ab, ap = b - a, p - a
    ln = np.hypot(*ab)
    return np.hypot(*ap) if ln == 0 else abs(ab[0] * ap[1] - ab[1] * ap[0]) / ln


def _flatness(Q):
    return max(_line_dist(q, Q[0], Q[-1]) for q in Q[1:-1])


def _leaves(P, tol, a=0.0, b=1.0, depth=0, out=None, margin=None):
    out = [] if out is None else out
    f = _flatness(_sub(P, a, b))
    if margin is not None:
        margin.append(abs(f - tol))
    if depth >= 16 or f <= tol:
        out.append((a, b))
        return out
    m = 0.5 * (a + b)
    _leaves(P, tol, a, m, depth + 1, out, margin)
    _leaves(P, tol, m, b, depth + 1, out, margin)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_flattening_leaves_are_the_first_flat_dyadic_intervals(seed):
    rng = np.random.default_rng(seed)
    checked = []
    for ctrl in (1, 2):
        pts = rng.uniform(0, 60, size=(ctrl + 2, 2))
        rad = rng.uniform(0.5, 3.0, size=ctrl + 2)
        path = Path(np.array([ctrl], np.uint8), pts.astype(np.float32), thickness=rad.astype(np.float32))
        tol = 0.05
        flat = op.flatten_with_thickness(path, tol, 1.0)
        P = [np.asarray(p, np.float64) for p in path.points]
        R = [np.array([float(r)]) for r in path.thickness]
        margin = []
        leaves = _leaves(P, tol, margin=margin)
        if min(margin) < 1e-4:
            continue  # a sub-curve within rounding of the tolerance
        checked.append(len(leaves))
        assert len(flat) == len(leaves)
        for (a, b), s in zip(leaves, flat):
            assert np.allclose(s[0], _bez(P, a), atol=1e-4) and np.allclose(s[1], _bez(P, b), atol=1e-4)
            assert abs(s[2] - _bez(R, a)[0]) < 1e-5 and abs(s[3] - _bez(R, b)[0]) < 1e-5
        # the packer emits the same, through a one-shape scene
        sc = Scene(64, 64, shapes=[Shape(path, stroke_width=1.0)])
        sc.groups.append(ShapeGroup([0], stroke=Solid((0, 0, 0, 1))))
        d = _pack(sc, RenderOptions(use_prefiltering=False, path_tolerance=tol))
        assert np.array_equal(d["segment_data"], op.segments_to_f32(flat))
    assert checked and max(checked) > 4


def test_open_polyline_segments_caps_and_directions():
    pts = np.array([[0, 0], [3, 4], [3, 10], [0, 10]], np.float32)
    sc = Scene(16, 16, shapes=[Shape(Path(np.zeros(3, np.uint8), pts), stroke_width=1.5)])
    sc.groups.append(ShapeGroup([0], stroke=Solid((0, 0, 0, 1))))
    d = _pack(sc, RenderOptions(use_prefiltering=False))
    s = d["segment_data"].reshape(-1, 12)
    assert s.shape[0] == 3
    assert np.array_equal(s[:, 0:2], pts[:3]) and np.array_equal(s[:, 2:4], pts[1:])
    assert np.all(s[:, 4:6] == 1.5)
    dirs = np.array([[0.6, 0.8], [0, 1], [-1, 0]], np.float32)
    assert np.allclose(s[:, 6:8], [[0, 0], -dirs[0], -dirs[1]])
    assert np.allclose(s[:, 8:10], [dirs[1], dirs[2], [0, 0]])
    assert list(s[:, 10]) == [1, 0, 0] and list(s[:, 11]) == [0, 0, 1]
    assert list(d["shape_data"][0, 1:3]) == [0, 3]


def test_closed_path_wraps_directions_without_caps():
    pts = np.array([[0, 0], [4, 0], [4, 4]], np.float32)
    sc = Scene(16, 16, shapes=[Shape(Path(np.zeros(3, np.uint8), pts, is_closed=True))])
    sc.groups.append(ShapeGroup([0], fill=Solid((0, 0, 0, 1))))
    s = _pack(sc, RenderOptions(use_prefiltering=False))["segment_data"].reshape(-1, 12)
    assert s.shape[0] == 3 and np.array_equal(s[2, 2:4], pts[0])
    d_last = (pts[0] - pts[2]) / np.linalg.norm(pts[0] - pts[2])
    assert np.allclose(s[0, 6:8], -d_last) and np.allclose(s[2, 8:10], [1, 0])
    assert not s[:, 10:12].any()
    # a collinear quadratic is flat at once: one segment
    q = Path(np.array([1], np.uint8), np.array([[0, 0], [2, 2], [4, 4]], np.float32))
    sc2 = Scene(16, 16, shapes=[Shape(q, stroke_width=1.0)])
    sc2.groups.append(ShapeGroup([0], stroke=Solid((0, 0, 0, 1))))
    assert _pack(sc2, RenderOptions())["segment_data"].reshape(-1, 12).shape[0] == 1
