"""Restatement of the reference's BVH packing and stroke flattening --
TEST INFRASTRUCTURE ONLY.

Only tests/ import this module.  It recomputes, in float32 arithmetic with
the reference's operation order, the parts of PreparedScene that the packer
(dvg_pack.cpp) emits for the reference's consumers only:

  build_bvh / pack_bvh     renderer/bvh.rs:62-226 (median split by centre on
                           the longer axis, stable sort, leaves of <= 8 items
                           (renderer/constants.rs:23), preorder nodes, skip
                           pointers of assign_bvh_skip)
  flatten_with_thickness   geometry.rs:218-340 with flatten_quad_thick /
                           flatten_cubic_thick (:346-424), the flatness tests
                           (:478-504) and annotate_segments (:430-476)
  shape_tolerance          prepare.rs:145-150, 181-183 (Mat3::average_scale,
                           math.rs:250-260)
  segments_to_f32          path.rs:176-193 (12 floats a segment)
"""
import numpy as np

f32 = np.float32
LEAF = 8
NONE = 0xFFFFFFFF
HALF = f32(0.5)


def _fmax(a, b):  # f32::max: the other operand when one is NaN
    return b if np.isnan(a) else (a if np.isnan(b) or a >= b else b)


def _fmin(a, b):
    return b if np.isnan(a) else (a if np.isnan(b) or a <= b else b)


# ---- renderer/bvh.rs ---------------------------------------------------------

def build_bvh(boxes):
    """boxes: list of (minx, miny, maxx, maxy) float32.  Returns (nodes, order):
    nodes = [[minx, miny, maxx, maxy, left, right, skip, start, count]] in
    preorder, order = item indices in leaf order (build_bvh_node, :149-190)."""
    boxes = [tuple(f32(v) for v in b) for b in boxes]
    order = list(range(len(boxes)))
    nodes = []

    def node(start, end):
        mnx = mny = f32(np.inf)
        mxx = mxy = f32(-np.inf)
        for i in order[start:end]:
            b = boxes[i]
            mnx, mny = _fmin(mnx, b[0]), _fmin(mny, b[1])
            mxx, mxy = _fmax(mxx, b[2]), _fmax(mxy, b[3])
        count = end - start
        ni = len(nodes)
        nodes.append([mnx, mny, mxx, mxy, NONE, NONE, NONE, start, count])
        if count <= LEAF:
            return ni
        axis = 0 if (mxx - mnx) >= (mxy - mny) else 1

        def centre(i):
            b = boxes[i]
            return (b[axis] + b[axis + 2]) * HALF

        order[start:end] = sorted(order[start:end], key=centre)  # stable, as sort_by
        mid = start + count // 2
        left = node(start, mid)
        right = node(mid, end)
        nodes[ni][4], nodes[ni][5], nodes[ni][8] = left, right, 0
        return ni

    if boxes:
        root = node(0, len(boxes))
        stack = [(root, NONE)]  # assign_bvh_skip, :192-215
        while stack:
            nid, skip = stack.pop()
            if nid == NONE:
                continue
            n = nodes[nid]
            n[6] = skip
            if n[8] == 0:
                if n[5] != NONE:
                    stack.append((n[5], skip))
                if n[4] != NONE:
                    stack.append((n[4], n[5]))
    return nodes, order


def pack_bvh(boxes, ids, out):
    """append_group_bvh / append_path_bvh (:62-147): appends to
    out = {'bounds': [], 'nodes': [], 'indices': []}; returns the meta
    [node_off, node_count, index_off, index_count] ([0] * 4 when empty)."""
    if not boxes:
        return [0, 0, 0, 0]
    node_off, idx_off = len(out["nodes"]) // 4, len(out["indices"])
    nodes, order = build_bvh(boxes)
    out["indices"].extend(int(ids[i]) for i in order)
    for n in nodes:
        out["bounds"].extend(n[:4])
        out["nodes"].extend([n[4], n[6], n[7], n[8]])
    return [node_off, len(nodes), idx_off, len(order)]


def curve_box(cv):
    """curve_segment_bounds (:217-226) of one 13-float curve_data record
    [kind, p0..p3, r0..r3]: the control-point box padded by max |r|."""
    xs, ys = [f32(cv[1 + 2 * j]) for j in range(4)], [f32(cv[2 + 2 * j]) for j in range(4)]
    mnx, mny, mxx, mxy = xs[0], ys[0], xs[0], ys[0]
    for j in range(1, 4):  # bounds_from_points, path_utils.rs
        mnx, mny = _fmin(mnx, xs[j]), _fmin(mny, ys[j])
        mxx, mxy = _fmax(mxx, xs[j]), _fmax(mxy, ys[j])
    pad = abs(f32(cv[9]))
    for j in range(10, 13):
        pad = _fmax(pad, abs(f32(cv[j])))
    pad = _fmax(pad, f32(0.0))
    return (mnx - pad, mny - pad, mxx + pad, mxy + pad)


# ---- geometry.rs flattening ----------------------------------------------------

def _lerp(a, b):
    return (a[0] + (b[0] - a[0]) * HALF, a[1] + (b[1] - a[1]) * HALF)


def _dist_point_line(p, a, b):  # distance_point_line, :492-503
    abx, aby = b[0] - a[0], b[1] - a[1]
    apx, apy = p[0] - a[0], p[1] - a[1]
    area = abs(abx * apy - aby * apx)
    ln = np.sqrt(abx * abx + aby * aby)
    if ln == f32(0.0):
        return np.sqrt(apx * apx + apy * apy)
    return area / ln


def _quad(p0, p1, p2, r0, r1, r2, tol, depth, out):
    if depth >= 16 or _dist_point_line(p1, p0, p2) <= tol:
        out.append([p0, p2, r0, r2])
        return
    p01, p12 = _lerp(p0, p1), _lerp(p1, p2)
    p012 = _lerp(p01, p12)
    r01, r12 = HALF * (r0 + r1), HALF * (r1 + r2)
    r012 = HALF * (r01 + r12)
    _quad(p0, p01, p012, r0, r01, r012, tol, depth + 1, out)
    _quad(p012, p12, p2, r012, r12, r2, tol, depth + 1, out)


def _cubic(p0, p1, p2, p3, r0, r1, r2, r3, tol, depth, out):
    if depth >= 16 or _fmax(_dist_point_line(p1, p0, p3), _dist_point_line(p2, p0, p3)) <= tol:
        out.append([p0, p3, r0, r3])
        return
    p01, p12, p23 = _lerp(p0, p1), _lerp(p1, p2), _lerp(p2, p3)
    p012, p123 = _lerp(p01, p12), _lerp(p12, p23)
    p0123 = _lerp(p012, p123)
    r01, r12, r23 = HALF * (r0 + r1), HALF * (r1 + r2), HALF * (r2 + r3)
    r012, r123 = HALF * (r01 + r12), HALF * (r12 + r23)
    r0123 = HALF * (r012 + r123)
    _cubic(p0, p01, p012, p0123, r0, r01, r012, r0123, tol, depth + 1, out)
    _cubic(p0123, p123, p23, p3, r0123, r123, r23, r3, tol, depth + 1, out)


def flatten_with_thickness(path, tol, default_radius):
    """Segments [start, end, r0, r1, prev_dir, next_dir, start_cap, end_cap]
    of a pydvg Path (thickness_scale 1)."""
    tol, default_radius = f32(tol), f32(default_radius)
    pts = np.asarray(path.points, np.float32).reshape(-1, 2)
    n = len(pts)
    ctrl = list(np.asarray(path.num_control_points).reshape(-1))
    out = []
    if n == 0 or not ctrl:
        return out
    thick = None if path.thickness is None else np.asarray(path.thickness, np.float32).reshape(-1)

    def pr(i):  # point_and_radius, :321-339
        if path.is_closed:
            i %= n
        elif i >= n:
            return None
        r = thick[i] * f32(1.0) if thick is not None and i < len(thick) else default_radius
        return (pts[i, 0], pts[i, 1]), f32(r)

    pid = 0
    for c in ctrl:
        if c > 2:
            break
        got = [pr(pid + j) for j in range(c + 2)]
        if any(g is None for g in got):
            break
        P = [g[0] for g in got]
        R = [g[1] for g in got]
        if c == 0:
            out.append([P[0], P[1], R[0], R[1]])
        elif c == 1:
            _quad(P[0], P[1], P[2], R[0], R[1], R[2], tol, 0, out)
        else:
            _cubic(P[0], P[1], P[2], P[3], R[0], R[1], R[2], R[3], tol, 0, out)
        pid += c + 1
    # annotate_segments
    cnt = len(out)
    dirs = []
    for s in out:
        vx, vy = s[1][0] - s[0][0], s[1][1] - s[0][1]
        ln = np.sqrt(vx * vx + vy * vy)
        dirs.append((vx / ln, vy / ln) if ln > f32(0.0) else (f32(0.0), f32(0.0)))
    closed = bool(path.is_closed)
    res = []
    for i, s in enumerate(out):
        if i == 0:
            prev = (-dirs[cnt - 1][0], -dirs[cnt - 1][1]) if closed else (f32(0.0), f32(0.0))
        else:
            prev = (-dirs[i - 1][0], -dirs[i - 1][1])
        if i + 1 < cnt:
            nxt = dirs[i + 1]
        elif closed:
            nxt = dirs[0]
        else:
            nxt = (f32(0.0), f32(0.0))
        res.append([s[0], s[1], s[2], s[3], prev, nxt, (not closed) and i == 0, (not closed) and i + 1 == cnt])
    return res


def segments_to_f32(segs):
    out = []
    for a, b, r0, r1, pd, nd, sc, ec in segs:
        out.extend([a[0], a[1], b[0], b[1], r0, r1, pd[0], pd[1], nd[0], nd[1], 1.0 if sc else 0.0,
                    1.0 if ec else 0.0])
    return np.asarray(out, np.float32)


def _avg_scale(m):
    m = np.asarray(m, np.float32).reshape(3, 3)
    sx = np.sqrt(m[0, 0] * m[0, 0] + m[1, 0] * m[1, 0])
    sy = np.sqrt(m[0, 1] * m[0, 1] + m[1, 1] * m[1, 1])
    return HALF * (sx + sy)


def shape_tolerance(options, group, shape):
    group_scale = _fmax(_avg_scale(group.shape_to_canvas), f32(1.0e-6))
    base_tol = _fmax(f32(options.path_tolerance), f32(0.01))
    total = _fmax(group_scale * _avg_scale(shape.transform), f32(1.0e-6))
    return _fmax(base_tol / total, f32(1.0e-3))
