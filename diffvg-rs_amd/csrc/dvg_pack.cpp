// Host packer of libdvg: builds the packed scene buffers from a flat Scene.
//
//   prepare_scene               src/renderer/prepare.rs:114-357
//   prepare_shape               src/renderer/prepare.rs:459-561
//   pack_paint / push_*         src/renderer/prepare.rs:359-457
//   path_to_curve_segments      src/renderer/path.rs:100-173
//   prepare_scene_backward      src/renderer/prepare_backward.rs:45-115
//   build_boundary_sampling_data src/backward/boundary.rs:46-110, 200-451
//
//   append_group_bvh / append_path_bvh  src/renderer/bvh.rs:62-226
//   flatten_with_thickness / annotate_segments  src/geometry.rs:218-476
//
// Every buffer of PreparedScene is filled as the reference fills it,
// including the flattened stroke segments and the group / path BVHs.
#include "dvg_pack.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace dvgh {

static const size_t MAX_F32_INT = 16777216;  // renderer/constants.rs:17

struct V2 {
    float x, y;
};

// src/math.rs Mat3 helpers on row-major 3x3 arrays.
static void scale_factors(const float *m, float &sx, float &sy) {
    sx = std::sqrt(m[0] * m[0] + m[3] * m[3]);
    sy = std::sqrt(m[1] * m[1] + m[4] * m[4]);
}
static float average_scale(const float *m) {
    float sx, sy;
    scale_factors(m, sx, sy);
    return 0.5f * (sx + sy);
}
static float max_scale(const float *m) {
    float sx, sy;
    scale_factors(m, sx, sy);
    return std::fmax(sx, sy);
}
static void inverse_or_identity(const float *m, float out[6]) {  // math.rs:269-290
    float a = m[0], b = m[1], c = m[3], d = m[4], tx = m[2], ty = m[5];
    float det = a * d - b * c;
    if (std::fabs(det) < 1.0e-6f) {
        const float id[6] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f};
        std::memcpy(out, id, sizeof(id));
        return;
    }
    float inv_det = 1.0f / det;
    float m00 = d * inv_det, m01 = -b * inv_det, m10 = -c * inv_det, m11 = a * inv_det;
    out[0] = m00; out[1] = m01; out[2] = -(m00 * tx + m01 * ty);
    out[3] = m10; out[4] = m11; out[5] = -(m10 * tx + m11 * ty);
}
static V2 xform(const float *m, V2 p) { return {m[0] * p.x + m[1] * p.y + m[2], m[3] * p.x + m[4] * p.y + m[5]}; }

struct PathView {
    const float *pts;
    uint32_t np;
    const uint32_t *ctrl;
    uint32_t nc;
    const float *th;
    uint32_t nt;
    bool closed;
};

// path_utils.rs:7-39
static bool point_radius(const PathView &p, size_t idx, float def_r, V2 &pt, float &r) {
    if (p.np == 0) return false;
    size_t i;
    if (p.closed) i = idx % p.np;
    else {
        if (idx >= p.np) return false;
        i = idx;
    }
    pt = {p.pts[2 * i], p.pts[2 * i + 1]};
    r = (p.nt > 0 && i < p.nt) ? p.th[i] * 1.0f : def_r;
    return true;
}

struct Curve {
    uint32_t kind;
    V2 p[4];
    float r[4];
};

// path.rs:100-173
static void curve_segments(const PathView &p, float sw, std::vector<Curve> &out) {
    out.clear();
    if (p.nc == 0 || p.np == 0) return;
    size_t pid = 0;
    for (uint32_t k = 0; k < p.nc; ++k) {
        uint32_t c = p.ctrl[k];
        if (c > 2) break;
        Curve cv;
        cv.kind = c;
        bool ok = true;
        for (uint32_t j = 0; j <= c + 1 && ok; ++j) ok = point_radius(p, pid + j, sw, cv.p[j], cv.r[j]);
        if (!ok) break;
        for (uint32_t j = c + 2; j < 4; ++j) {
            cv.p[j] = cv.p[c + 1];
            cv.r[j] = cv.r[c + 1];
        }
        out.push_back(cv);
        pid += c + 1;
    }
}

// geometry.rs:388-401
static float dist_point_line(V2 p, V2 a, V2 b) {
    V2 ab = {b.x - a.x, b.y - a.y}, ap = {p.x - a.x, p.y - a.y};
    float area = std::fabs(ab.x * ap.y - ab.y * ap.x);
    float len = std::sqrt(ab.x * ab.x + ab.y * ab.y);
    if (len == 0.0f) return std::sqrt(ap.x * ap.x + ap.y * ap.y);
    return area / len;
}
static V2 lerp(V2 a, V2 b, float t) { return {a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t}; }

// Flattened stroke segments (geometry.rs:218-375, StrokeSegment of
// geometry.rs:30-56): recursive halving until flat within the tolerance,
// radii interpolated, then prev / next directions and cap flags.
struct StrokeSeg {
    V2 a, b;
    float r0, r1;
    V2 prev_dir, next_dir;
    bool start_cap, end_cap;
};
static void flat_quad(V2 p0, V2 p1, V2 p2, float r0, float r1, float r2, float tol, int depth,
                      std::vector<StrokeSeg> &out) {
    if (depth >= 16 || dist_point_line(p1, p0, p2) <= tol) {
        out.push_back({p0, p2, r0, r2, {0.f, 0.f}, {0.f, 0.f}, false, false});
        return;
    }
    V2 p01 = lerp(p0, p1, 0.5f), p12 = lerp(p1, p2, 0.5f), p012 = lerp(p01, p12, 0.5f);
    float r01 = 0.5f * (r0 + r1), r12 = 0.5f * (r1 + r2), r012 = 0.5f * (r01 + r12);
    flat_quad(p0, p01, p012, r0, r01, r012, tol, depth + 1, out);
    flat_quad(p012, p12, p2, r012, r12, r2, tol, depth + 1, out);
}
static void flat_cubic(V2 p0, V2 p1, V2 p2, V2 p3, float r0, float r1, float r2, float r3, float tol, int depth,
                       std::vector<StrokeSeg> &out) {
    if (depth >= 16 || std::fmax(dist_point_line(p1, p0, p3), dist_point_line(p2, p0, p3)) <= tol) {
        out.push_back({p0, p3, r0, r3, {0.f, 0.f}, {0.f, 0.f}, false, false});
        return;
    }
    V2 p01 = lerp(p0, p1, 0.5f), p12 = lerp(p1, p2, 0.5f), p23 = lerp(p2, p3, 0.5f);
    V2 p012 = lerp(p01, p12, 0.5f), p123 = lerp(p12, p23, 0.5f), p0123 = lerp(p012, p123, 0.5f);
    float r01 = 0.5f * (r0 + r1), r12 = 0.5f * (r1 + r2), r23 = 0.5f * (r2 + r3);
    float r012 = 0.5f * (r01 + r12), r123 = 0.5f * (r12 + r23), r0123 = 0.5f * (r012 + r123);
    flat_cubic(p0, p01, p012, p0123, r0, r01, r012, r0123, tol, depth + 1, out);
    flat_cubic(p0123, p123, p23, p3, r0123, r123, r23, r3, tol, depth + 1, out);
}
// annotate_segments (geometry.rs:432-476)
static void annotate(std::vector<StrokeSeg> &segs, size_t first, bool closed) {
    size_t count = segs.size() - first;
    if (count == 0) return;
    std::vector<V2> dirs(count);
    for (size_t i = 0; i < count; ++i) {
        const StrokeSeg &sg = segs[first + i];
        V2 v = {sg.b.x - sg.a.x, sg.b.y - sg.a.y};
        float len = std::sqrt(v.x * v.x + v.y * v.y);
        dirs[i] = len > 0.f ? V2{v.x / len, v.y / len} : V2{0.f, 0.f};
    }
    for (size_t i = 0; i < count; ++i) {
        StrokeSeg &sg = segs[first + i];
        sg.start_cap = !closed && i == 0;
        sg.end_cap = !closed && i + 1 == count;
        V2 pd = {0.f, 0.f};
        if (i == 0) {
            if (closed) pd = {-dirs[count - 1].x, -dirs[count - 1].y};
        } else {
            pd = {-dirs[i - 1].x, -dirs[i - 1].y};
        }
        V2 nd = {0.f, 0.f};
        if (i + 1 < count) nd = dirs[i + 1];
        else if (closed) nd = dirs[0];
        sg.prev_dir = pd;
        sg.next_dir = nd;
    }
}

// ---- BVHs (renderer/bvh.rs:62-226) ------------------------------------------
// Median split by centre on the longer axis (a stable sort, as the
// reference's sort_by), leaves of at most 8 items (constants.rs:23), nodes in
// preorder with skip pointers for stackless traversal.
static const uint32_t BVH_LEAF = 8, BVH_NONE = 0xffffffffu;
struct Box {
    V2 mn, mx;
};
struct BvhNode {
    Box b;
    uint32_t left, right, skip, start, count;
};
static uint32_t bvh_node(std::vector<BvhNode> &nodes, const std::vector<Box> &bx, std::vector<size_t> &idx,
                         size_t start, size_t end) {
    Box nb = {{INFINITY, INFINITY}, {-INFINITY, -INFINITY}};
    for (size_t i = start; i < end; ++i) {
        const Box &o = bx[idx[i]];
        nb.mn = {std::fmin(nb.mn.x, o.mn.x), std::fmin(nb.mn.y, o.mn.y)};
        nb.mx = {std::fmax(nb.mx.x, o.mx.x), std::fmax(nb.mx.y, o.mx.y)};
    }
    size_t count = end - start;
    uint32_t ni = (uint32_t)nodes.size();
    nodes.push_back({nb, BVH_NONE, BVH_NONE, BVH_NONE, (uint32_t)start, (uint32_t)count});
    if (count <= BVH_LEAF) return ni;
    int axis = (nb.mx.x - nb.mn.x) >= (nb.mx.y - nb.mn.y) ? 0 : 1;
    auto ctr = [&](size_t k) {
        const Box &o = bx[k];
        return axis == 0 ? (o.mn.x + o.mx.x) * 0.5f : (o.mn.y + o.mx.y) * 0.5f;
    };
    std::stable_sort(idx.begin() + start, idx.begin() + end, [&](size_t a, size_t b) { return ctr(a) < ctr(b); });
    size_t mid = start + count / 2;
    uint32_t l = bvh_node(nodes, bx, idx, start, mid);
    uint32_t r = bvh_node(nodes, bx, idx, mid, end);
    nodes[ni].left = l;
    nodes[ni].right = r;
    nodes[ni].count = 0;
    return ni;
}
static void bvh_skip(std::vector<BvhNode> &nodes, uint32_t root) {  // assign_bvh_skip
    if (nodes.empty()) return;
    std::vector<std::pair<uint32_t, uint32_t>> st = {{root, BVH_NONE}};
    while (!st.empty()) {
        auto [id, sk] = st.back();
        st.pop_back();
        if (id == BVH_NONE) continue;
        BvhNode &n = nodes[id];
        n.skip = sk;
        if (n.count == 0) {
            if (n.right != BVH_NONE) st.push_back({n.right, sk});
            if (n.left != BVH_NONE) st.push_back({n.left, n.right});
        }
    }
}
// Appends a BVH over `bx` (item ids `ids`) and its meta [node_off, node_cnt,
// idx_off, idx_cnt] (append_group_bvh / append_path_bvh).
static void append_bvh(const std::vector<Box> &bx, const std::vector<uint32_t> &ids, std::vector<float> &ob,
                       std::vector<uint32_t> &on, std::vector<uint32_t> &oi, std::vector<uint32_t> &meta) {
    if (bx.empty()) {
        meta.insert(meta.end(), {0u, 0u, 0u, 0u});
        return;
    }
    uint32_t node_off = (uint32_t)(on.size() / 4), idx_off = (uint32_t)oi.size();
    std::vector<size_t> idx(bx.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::vector<BvhNode> nodes;
    uint32_t root = bvh_node(nodes, bx, idx, 0, idx.size());
    bvh_skip(nodes, root);
    for (size_t i : idx) oi.push_back(ids[i]);
    for (const BvhNode &n : nodes) {
        ob.insert(ob.end(), {n.b.mn.x, n.b.mn.y, n.b.mx.x, n.b.mx.y});
        on.insert(on.end(), {n.left, n.skip, n.start, n.count});
    }
    meta.insert(meta.end(), {node_off, (uint32_t)nodes.size(), idx_off, (uint32_t)idx.size()});
}

static float polyline_length(const Curve &c) {  // backward/boundary.rs:230-309 per segment
    auto len = [](V2 a, V2 b) { return std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y)); };
    if (c.kind == 0) return len(c.p[0], c.p[1]);
    if (c.kind == 1) {
        float t = 0.5f, tt = 1.0f - t;
        V2 v1 = {((c.p[0].x * (tt * tt)) + (c.p[1].x * (2.0f * tt * t))) + (c.p[2].x * (t * t)),
                 ((c.p[0].y * (tt * tt)) + (c.p[1].y * (2.0f * tt * t))) + (c.p[2].y * (t * t))};
        return len(c.p[0], v1) + len(v1, c.p[2]);
    }
    auto ev = [&](float t) {
        float tt = 1.0f - t;
        return V2{(((c.p[0].x * (tt * tt * tt)) + (c.p[1].x * (3.0f * tt * tt * t))) + (c.p[2].x * (3.0f * tt * t * t))) +
                      (c.p[3].x * (t * t * t)),
                  (((c.p[0].y * (tt * tt * tt)) + (c.p[1].y * (3.0f * tt * tt * t))) + (c.p[2].y * (3.0f * tt * t * t))) +
                      (c.p[3].y * (t * t * t))};
    };
    V2 v1 = ev(1.0f / 3.0f), v2 = ev(2.0f / 3.0f);
    return (len(c.p[0], v1) + len(v1, v2)) + len(v2, c.p[3]);
}

int prepare(const dvg_scene_desc *d, const dvg_render_options *opt, dvg_prepared &P, const char **err) {
    P = dvg_prepared();
    const uint32_t S = d->num_shapes, G = d->num_groups;
    size_t npix = (size_t)d->width * d->height;
    if (d->background_image && d->background_image_len != npix * 4) {
        *err = "background image size mismatch";
        return DVG_INVALID_SCENE;
    }
    if (S > MAX_F32_INT) {
        *err = "too many shapes for f32 indexing";
        return DVG_INVALID_SCENE;
    }
    bool prefilter = opt->use_prefiltering != 0;
    // Gradient paints must be listed in packing order (fill, then stroke, per
    // group) so that desc gradient i is packed gradient i.
    {
        uint32_t next = 0;
        for (uint32_t g = 0; g < G; ++g) {
            uint32_t kinds[2] = {d->fill_kind[g], d->stroke_kind[g]};
            uint32_t idx[2] = {d->fill_gradient[g], d->stroke_gradient[g]};
            for (int k = 0; k < 2; ++k)
                if (kinds[k] == DVG_PAINT_LINEAR || kinds[k] == DVG_PAINT_RADIAL) {
                    if (idx[k] != next || next >= d->num_gradients) {
                        *err = "gradient table not in group order";
                        return DVG_INVALID_SCENE;
                    }
                    ++next;
                }
        }
    }
    auto path_view = [&](uint32_t i) {
        PathView v;
        v.pts = d->path_points + 2 * (size_t)d->path_point_offset[i];
        v.np = d->path_point_count[i];
        v.ctrl = d->path_num_controls + d->path_ctrl_offset[i];
        v.nc = d->path_ctrl_count[i];
        v.th = d->path_thickness + d->path_thickness_offset[i];
        v.nt = d->path_thickness_count[i];
        v.closed = d->path_is_closed[i] != 0;
        return v;
    };
    std::vector<Curve> curves;
    std::vector<StrokeSeg> segs;
    std::vector<Box> path_boxes, group_boxes;
    std::vector<uint32_t> path_ids, group_ids;
    for (uint32_t g = 0; g < G; ++g) {
        const float *s2c = d->shape_to_canvas + 9 * (size_t)g;
        const float *c2s = d->canvas_to_shape + 9 * (size_t)g;
        float group_scale = std::fmax(average_scale(s2c), 1.0e-6f);
        float inv_scale = 1.0f / group_scale;
        float base_tol = std::fmax(opt->path_tolerance, 0.01f);
        float aa_local = prefilter ? inv_scale : 0.0f;
        uint32_t shape_offset = (uint32_t)P.group_shapes.size();
        uint32_t shape_count = 0;
        bool gb_valid = false;
        float gb[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
        group_boxes.clear();
        group_ids.clear();
        P.group_xform.insert(P.group_xform.end(), {c2s[0], c2s[1], c2s[2], c2s[3], c2s[4], c2s[5]});
        P.group_shape_xform.insert(P.group_shape_xform.end(), {s2c[0], s2c[1], s2c[2], s2c[3], s2c[4], s2c[5]});
        P.group_inv_scale.push_back(inv_scale);
        for (uint32_t k = 0; k < d->group_shape_count[g]; ++k) {
            uint32_t si = d->group_shape_indices[d->group_shape_offset[g] + k];
            if (si >= S) {
                *err = "group references invalid shape index";
                return DVG_INVALID_SCENE;
            }
            const float *T = d->shape_transform + 9 * (size_t)si;
            float total_scale = std::fmax(group_scale * average_scale(T), 1.0e-6f);
            float tol = std::fmax(base_tol / total_scale, 1.0e-3f);
            uint32_t kind = d->shape_kind[si];
            const float *sp = d->shape_params + 4 * (size_t)si;
            float sw = d->stroke_width[si];
            float max_r = std::fabs(sw);
            float params[4] = {0.f, 0.f, 0.f, 0.f};
            bool has_bounds = true, has_thickness = false, approx = false;
            uint32_t seg_off = 0, seg_cnt = 0;
            path_boxes.clear();
            path_ids.clear();
            V2 bmin{0.f, 0.f}, bmax{0.f, 0.f};
            uint32_t curve_off = (uint32_t)(P.curve_data.size() / 13), curve_cnt = 0;
            if (kind == DVG_SHAPE_CIRCLE) {
                float r = std::fabs(sp[2]);
                params[0] = sp[0]; params[1] = sp[1]; params[2] = r;
                bmin = {sp[0] - r, sp[1] - r};
                bmax = {sp[0] + r, sp[1] + r};
            } else if (kind == DVG_SHAPE_ELLIPSE) {
                float rx = std::fabs(sp[2]), ry = std::fabs(sp[3]);
                params[0] = sp[0]; params[1] = sp[1]; params[2] = rx; params[3] = ry;
                bmin = {sp[0] - rx, sp[1] - ry};
                bmax = {sp[0] + rx, sp[1] + ry};
            } else if (kind == DVG_SHAPE_RECT) {
                std::memcpy(params, sp, sizeof(params));
                bmin = {sp[0], sp[1]};
                bmax = {sp[2], sp[3]};
            } else {
                PathView pv = path_view(si);
                has_thickness = pv.nt > 0;
                approx = d->path_use_distance_approx[si] != 0;
                curve_segments(pv, sw, curves);
                // the flattened stroke segments (prepare.rs:520-530), in
                // the order flatten_with_thickness visits the path
                size_t first = segs.size();
                for (const Curve &c : curves) {
                    if (c.kind == 0) segs.push_back({c.p[0], c.p[1], c.r[0], c.r[1], {0.f, 0.f}, {0.f, 0.f}, false, false});
                    else if (c.kind == 1) flat_quad(c.p[0], c.p[1], c.p[2], c.r[0], c.r[1], c.r[2], tol, 0, segs);
                    else flat_cubic(c.p[0], c.p[1], c.p[2], c.p[3], c.r[0], c.r[1], c.r[2], c.r[3], tol, 0, segs);
                }
                annotate(segs, first, pv.closed);
                for (size_t k = first; k < segs.size(); ++k)
                    max_r = std::fmax(std::fmax(max_r, std::fabs(segs[k].r0)), std::fabs(segs[k].r1));
                seg_off = (uint32_t)first;
                seg_cnt = (uint32_t)(segs.size() - first);
                // path BVH over the curve segments (prepare.rs:241-250)
                path_boxes.clear();
                path_ids.clear();
                for (size_t k = 0; k < curves.size(); ++k) {
                    const Curve &c = curves[k];
                    Box b = {{INFINITY, INFINITY}, {-INFINITY, -INFINITY}};
                    for (int j = 0; j < 4; ++j) {
                        b.mn = {std::fmin(b.mn.x, c.p[j].x), std::fmin(b.mn.y, c.p[j].y)};
                        b.mx = {std::fmax(b.mx.x, c.p[j].x), std::fmax(b.mx.y, c.p[j].y)};
                    }
                    float pad = std::fmax(std::fmax(std::fmax(std::fabs(c.r[0]), std::fabs(c.r[1])), std::fabs(c.r[2])),
                                          std::fabs(c.r[3]));
                    pad = std::fmax(pad, 0.0f);
                    path_boxes.push_back({{b.mn.x - pad, b.mn.y - pad}, {b.mx.x + pad, b.mx.y + pad}});
                    path_ids.push_back(curve_off + (uint32_t)k);
                }
                for (const Curve &c : curves) {
                    P.curve_data.push_back((float)c.kind);
                    for (int j = 0; j < 4; ++j) {
                        P.curve_data.push_back(c.p[j].x);
                        P.curve_data.push_back(c.p[j].y);
                    }
                    for (int j = 0; j < 4; ++j) P.curve_data.push_back(c.r[j]);
                }
                curve_cnt = (uint32_t)curves.size();
                if (pv.np == 0) {
                    has_bounds = false;
                } else {
                    bmin = {INFINITY, INFINITY};
                    bmax = {-INFINITY, -INFINITY};
                    for (uint32_t j = 0; j < pv.np; ++j) {
                        bmin = {std::fmin(bmin.x, pv.pts[2 * j]), std::fmin(bmin.y, pv.pts[2 * j + 1])};
                        bmax = {std::fmax(bmax.x, pv.pts[2 * j]), std::fmax(bmax.y, pv.pts[2 * j + 1])};
                    }
                    params[0] = bmin.x; params[1] = bmin.y; params[2] = bmax.x; params[3] = bmax.y;
                }
            }
            float inv[6];
            inverse_or_identity(T, inv);
            P.shape_xform.insert(P.shape_xform.end(), inv, inv + 6);
            uint32_t prep = (uint32_t)(P.shape_data.size() / 15);
            P.group_shapes.push_back((float)prep);
            P.prepared_to_scene.push_back(si);
            float sd[15] = {(float)kind, (float)seg_off, (float)seg_cnt, sw, params[0], params[1], params[2], params[3],
                            has_thickness ? 1.f : 0.f, (float)d->stroke_join[si], (float)d->stroke_cap[si],
                            d->stroke_miter_limit[si], (float)curve_off, (float)curve_cnt, approx ? 1.f : 0.f};
            P.shape_data.insert(P.shape_data.end(), sd, sd + 15);
            float stroke_pad = std::fabs(max_r) * std::fmax(max_scale(T), 0.0f);
            if (d->stroke_join[si] == DVG_JOIN_MITER) stroke_pad *= std::fmax(d->stroke_miter_limit[si], 1.0f);
            float pad = std::fmax(aa_local + stroke_pad, 0.0f);
            if (has_bounds) {
                V2 cs[4] = {{bmin.x, bmin.y}, {bmax.x, bmin.y}, {bmax.x, bmax.y}, {bmin.x, bmax.y}};
                V2 mn{INFINITY, INFINITY}, mx{-INFINITY, -INFINITY};
                for (V2 c : cs) {
                    V2 t = xform(T, c);
                    mn = {std::fmin(mn.x, t.x), std::fmin(mn.y, t.y)};
                    mx = {std::fmax(mx.x, t.x), std::fmax(mx.y, t.y)};
                }
                float b[4] = {mn.x - pad, mn.y - pad, mx.x + pad, mx.y + pad};
                P.shape_bounds.insert(P.shape_bounds.end(), b, b + 4);
                group_boxes.push_back({{b[0], b[1]}, {b[2], b[3]}});
                group_ids.push_back(prep);
                if (!gb_valid) {
                    std::memcpy(gb, b, sizeof(b));
                    gb_valid = true;
                } else {
                    gb[0] = std::fmin(gb[0], b[0]); gb[1] = std::fmin(gb[1], b[1]);
                    gb[2] = std::fmax(gb[2], b[2]); gb[3] = std::fmax(gb[3], b[3]);
                }
            } else {
                P.shape_bounds.insert(P.shape_bounds.end(), {1.f, 1.f, 0.f, 0.f});
            }
            append_bvh(path_boxes, path_ids, P.path_bvh_bounds, P.path_bvh_nodes, P.path_bvh_indices,
                       P.path_bvh_meta);
            ++shape_count;
        }
        if (P.group_shapes.size() > MAX_F32_INT) {
            *err = "too many group shape indices for f32 indexing";
            return DVG_INVALID_SCENE;
        }
        // pack_paint (prepare.rs:359-457)
        float gd[16] = {(float)shape_offset, (float)shape_count, 0.f, 0.f, 0.f, 0.f, 1.f, (float)d->fill_rule[g],
                        0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        uint32_t kinds[2] = {d->fill_kind[g], d->stroke_kind[g]};
        const float *cols[2] = {d->fill_color + 4 * (size_t)g, d->stroke_color + 4 * (size_t)g};
        uint32_t gidx[2] = {d->fill_gradient[g], d->stroke_gradient[g]};
        for (int k = 0; k < 2; ++k) {
            gd[2 + 2 * k] = (float)kinds[k];
            if (kinds[k] == DVG_PAINT_SOLID) {
                for (int c = 0; c < 4; ++c) gd[8 + 4 * k + c] = cols[k][c];
            } else if (kinds[k] == DVG_PAINT_LINEAR || kinds[k] == DVG_PAINT_RADIAL) {
                uint32_t gi = gidx[k];
                uint32_t so = d->gradient_stop_offset[gi], sc = d->gradient_stop_count[gi];
                uint32_t stop_off = (uint32_t)P.stop_offsets.size();
                for (uint32_t j = 0; j < sc; ++j) {
                    P.stop_offsets.push_back(d->stop_offsets[so + j]);
                    for (int c = 0; c < 4; ++c) P.stop_colors.push_back(d->stop_colors[4 * (size_t)(so + j) + c]);
                }
                uint32_t index = (uint32_t)(P.gradient_data.size() / 8);
                const float *gp = d->gradient_params + 4 * (size_t)gi;
                float hdr[8] = {kinds[k] == DVG_PAINT_LINEAR ? 0.f : 1.f, gp[0], gp[1], gp[2], gp[3],
                                (float)stop_off, (float)sc, 0.f};
                P.gradient_data.insert(P.gradient_data.end(), hdr, hdr + 8);
                gd[3 + 2 * k] = (float)index;
            }
        }
        if (gb_valid) P.group_bounds.insert(P.group_bounds.end(), gb, gb + 4);
        else P.group_bounds.insert(P.group_bounds.end(), {1.f, 1.f, 0.f, 0.f});
        // group BVH over the shapes with bounds (prepare.rs:290-297)
        append_bvh(group_boxes, group_ids, P.group_bvh_bounds, P.group_bvh_nodes, P.group_bvh_indices,
                   P.group_bvh_meta);
        P.group_data.insert(P.group_data.end(), gd, gd + 16);
    }
    // segments_to_f32 (path.rs:176-193)
    P.segment_data.reserve(segs.size() * 12);
    for (const StrokeSeg &sg : segs)
        P.segment_data.insert(P.segment_data.end(), {sg.a.x, sg.a.y, sg.b.x, sg.b.y, sg.r0, sg.r1, sg.prev_dir.x,
                                                     sg.prev_dir.y, sg.next_dir.x, sg.next_dir.y,
                                                     sg.start_cap ? 1.f : 0.f, sg.end_cap ? 1.f : 0.f});
    if (segs.size() > MAX_F32_INT) {
        *err = "too many segments for f32 indexing";
        return DVG_INVALID_SCENE;
    }
    if (P.curve_data.size() / 13 > MAX_F32_INT) {
        *err = "too many curve segments for f32 indexing";
        return DVG_INVALID_SCENE;
    }
    P.num_groups = G;

    // prepare_scene_backward (prepare_backward.rs:45-115)
    P.shape_path_offsets.assign(S, 0);
    P.shape_path_point_counts.assign(S, 0);
    P.shape_path_ctrl_offsets.assign(S, 0);
    P.shape_path_ctrl_counts.assign(S, 0);
    P.shape_path_thickness_offsets.assign(S, 0);
    P.shape_path_thickness_counts.assign(S, 0);
    P.shape_path_is_closed.assign(S, 0);
    for (uint32_t si = 0; si < S; ++si) {
        const float *T = d->shape_transform + 9 * (size_t)si;
        P.shape_transform.insert(P.shape_transform.end(), {T[0], T[1], T[2], T[3], T[4], T[5]});
        if (d->shape_kind[si] != DVG_SHAPE_PATH) continue;
        PathView pv = path_view(si);
        P.shape_path_offsets[si] = (uint32_t)(P.path_points.size() / 2);
        P.shape_path_point_counts[si] = pv.np;
        P.shape_path_ctrl_offsets[si] = (uint32_t)P.path_num_control_points.size();
        P.shape_path_ctrl_counts[si] = pv.nc;
        P.shape_path_thickness_offsets[si] = (uint32_t)P.path_thickness.size();
        P.shape_path_thickness_counts[si] = pv.nt;
        P.shape_path_is_closed[si] = pv.closed ? 1u : 0u;
        P.path_points.insert(P.path_points.end(), pv.pts, pv.pts + 2 * (size_t)pv.np);
        P.path_num_control_points.insert(P.path_num_control_points.end(), pv.ctrl, pv.ctrl + pv.nc);
        P.path_thickness.insert(P.path_thickness.end(), pv.th, pv.th + pv.nt);
    }

    // build_boundary_sampling_data (backward/boundary.rs:46-110), used by the
    // non-prefiltered backward only (backward_gpu.rs:217-221)
    if (!prefilter && G > 0) {
        std::vector<float> L(S, 0.f);
        const float PI = 3.14159274101257324219f;
        for (uint32_t si = 0; si < S; ++si) {  // compute_shape_lengths, boundary.rs:200-227
            const float *sp = d->shape_params + 4 * (size_t)si;
            uint32_t kind = d->shape_kind[si];
            if (kind == DVG_SHAPE_CIRCLE) {
                L[si] = 2.0f * PI * std::fabs(sp[2]);
            } else if (kind == DVG_SHAPE_ELLIPSE) {
                float a = std::fabs(sp[2]), b = std::fabs(sp[3]);
                L[si] = (a == 0.f || b == 0.f) ? 0.f : PI * (3.0f * (a + b) - std::sqrt((3.0f * a + b) * (a + 3.0f * b)));
            } else if (kind == DVG_SHAPE_RECT) {
                L[si] = 2.0f * (std::fabs(sp[2] - sp[0]) + std::fabs(sp[3] - sp[1]));
            } else {
                PathView pv = path_view(si);
                curve_segments(pv, 0.f, curves);
                float len = 0.f;
                for (const Curve &c : curves) len += polyline_length(c);
                L[si] = len;
            }
        }
        std::vector<float> cdf, pmf;
        std::vector<uint32_t> sids, gids;
        float accum = 0.f;
        for (uint32_t g = 0; g < G; ++g)  // build_shape_cdf, boundary.rs:312-346
            for (uint32_t k = 0; k < d->group_shape_count[g]; ++k) {
                uint32_t si = d->group_shape_indices[d->group_shape_offset[g] + k];
                float len = L[si];
                sids.push_back(si);
                gids.push_back(g);
                accum += len;
                cdf.push_back(accum);
                pmf.push_back(len);
            }
        if (accum > 0.f && std::isfinite(accum) && !cdf.empty()) {
            for (float &v : cdf) v /= accum;
            for (float &v : pmf) v /= accum;
            P.shape_lengths = L;
            P.shape_cdf = cdf;
            P.shape_pmf = pmf;
            P.shape_ids = sids;
            P.group_ids = gids;
            P.path_cdf_offsets.assign(S, 0);
            P.path_cdf_counts.assign(S, 0);
            P.path_point_offsets.assign(S, 0);
            for (uint32_t si = 0; si < S; ++si) {  // build_path_cdfs, boundary.rs:349-451
                if (d->shape_kind[si] != DVG_SHAPE_PATH) continue;
                float len = L[si];
                PathView pv = path_view(si);
                if (!(len > 0.f) || !std::isfinite(len) || pv.nc == 0) continue;
                float inv = 1.0f / len;
                P.path_cdf_offsets[si] = (uint32_t)P.path_cdf.size();
                P.path_cdf_counts[si] = pv.nc;
                P.path_point_offsets[si] = (uint32_t)P.path_point_ids.size();
                size_t pid = 0;
                float prev = 0.f;
                for (uint32_t k = 0; k < pv.nc; ++k) {
                    uint32_t c = pv.ctrl[k];
                    P.path_point_ids.push_back((uint32_t)pid);
                    float sl = 0.f;
                    if (c <= 2) {
                        Curve cv;
                        cv.kind = c;
                        bool ok = true;
                        for (uint32_t j = 0; j <= c + 1 && ok; ++j) ok = point_radius(pv, pid + j, 0.f, cv.p[j], cv.r[j]);
                        if (ok) {
                            for (uint32_t j = c + 2; j < 4; ++j) cv.p[j] = cv.p[c + 1];
                            sl = polyline_length(cv);
                        }
                        pid += c + 1;
                    }
                    float sn = sl * inv;
                    P.path_pmf.push_back(sn);
                    prev = sn + prev;
                    P.path_cdf.push_back(prev);
                }
            }
        }
    }

    dvg_render_params &rp = P.params;
    std::memset(&rp, 0, sizeof(rp));
    rp.width = d->width;
    rp.height = d->height;
    rp.samples_x = std::max(opt->samples_x, 1u);
    rp.samples_y = std::max(opt->samples_y, 1u);
    rp.seed = opt->seed;
    rp.jitter = prefilter ? 0u : (opt->jitter ? 1u : 0u);
    rp.use_prefiltering = prefilter ? 1u : 0u;
    rp.filter_type = d->filter_type;
    rp.filter_radius = d->filter_radius;
    for (int c = 0; c < 4; ++c) rp.background[c] = d->background[c];
    if (d->background_image) P.background_image.assign(d->background_image, d->background_image + npix * 4);
    rp.background_image = P.background_image.empty() ? nullptr : P.background_image.data();
    return DVG_OK;
}

template <typename T>
static const T *ptr(const std::vector<T> &v) {
    return v.empty() ? nullptr : v.data();
}

void pack_views(const dvg_prepared &P, dvg_scene_pack *pk, dvg_backward_pack *bp, dvg_render_params *params) {
    if (pk) {
        std::memset(pk, 0, sizeof(*pk));
        pk->num_groups = P.num_groups;
        pk->num_prepared_shapes = (uint32_t)(P.shape_data.size() / 15);
        pk->shape_data = ptr(P.shape_data);
        pk->segment_data = ptr(P.segment_data);
        pk->segment_data_len = P.segment_data.size();
        pk->curve_data = ptr(P.curve_data);
        pk->curve_data_len = P.curve_data.size();
        pk->shape_bounds = ptr(P.shape_bounds);
        pk->group_bounds = ptr(P.group_bounds);
        pk->group_data = ptr(P.group_data);
        pk->group_xform = ptr(P.group_xform);
        pk->group_shape_xform = ptr(P.group_shape_xform);
        pk->group_inv_scale = ptr(P.group_inv_scale);
        pk->group_shapes = ptr(P.group_shapes);
        pk->group_shapes_len = P.group_shapes.size();
        pk->shape_xform = ptr(P.shape_xform);
        pk->group_bvh_bounds = ptr(P.group_bvh_bounds);
        pk->group_bvh_bounds_len = P.group_bvh_bounds.size();
        pk->group_bvh_nodes = ptr(P.group_bvh_nodes);
        pk->group_bvh_nodes_len = P.group_bvh_nodes.size();
        pk->group_bvh_indices = ptr(P.group_bvh_indices);
        pk->group_bvh_indices_len = P.group_bvh_indices.size();
        pk->group_bvh_meta = ptr(P.group_bvh_meta);
        pk->path_bvh_bounds = ptr(P.path_bvh_bounds);
        pk->path_bvh_bounds_len = P.path_bvh_bounds.size();
        pk->path_bvh_nodes = ptr(P.path_bvh_nodes);
        pk->path_bvh_nodes_len = P.path_bvh_nodes.size();
        pk->path_bvh_indices = ptr(P.path_bvh_indices);
        pk->path_bvh_indices_len = P.path_bvh_indices.size();
        pk->path_bvh_meta = ptr(P.path_bvh_meta);
        pk->gradient_data = ptr(P.gradient_data);
        pk->gradient_data_len = P.gradient_data.size();
        pk->stop_offsets = ptr(P.stop_offsets);
        pk->stop_offsets_len = P.stop_offsets.size();
        pk->stop_colors = ptr(P.stop_colors);
    }
    if (bp) {
        std::memset(bp, 0, sizeof(*bp));
        bp->num_shapes = (uint32_t)P.shape_path_offsets.size();
        bp->prepared_to_scene = ptr(P.prepared_to_scene);
        bp->shape_transform = ptr(P.shape_transform);
        bp->path_points = ptr(P.path_points);
        bp->path_points_len = P.path_points.size();
        bp->path_num_control_points = ptr(P.path_num_control_points);
        bp->path_num_control_points_len = P.path_num_control_points.size();
        bp->path_thickness = ptr(P.path_thickness);
        bp->path_thickness_len = P.path_thickness.size();
        bp->shape_path_offsets = ptr(P.shape_path_offsets);
        bp->shape_path_point_counts = ptr(P.shape_path_point_counts);
        bp->shape_path_ctrl_offsets = ptr(P.shape_path_ctrl_offsets);
        bp->shape_path_ctrl_counts = ptr(P.shape_path_ctrl_counts);
        bp->shape_path_thickness_offsets = ptr(P.shape_path_thickness_offsets);
        bp->shape_path_thickness_counts = ptr(P.shape_path_thickness_counts);
        bp->shape_path_is_closed = ptr(P.shape_path_is_closed);
        bp->shape_lengths = ptr(P.shape_lengths);
        bp->shape_cdf = ptr(P.shape_cdf);
        bp->shape_cdf_len = P.shape_cdf.size();
        bp->shape_pmf = ptr(P.shape_pmf);
        bp->shape_ids = ptr(P.shape_ids);
        bp->group_ids = ptr(P.group_ids);
        bp->path_cdf = ptr(P.path_cdf);
        bp->path_cdf_len = P.path_cdf.size();
        bp->path_pmf = ptr(P.path_pmf);
        bp->path_point_ids = ptr(P.path_point_ids);
        bp->path_point_ids_len = P.path_point_ids.size();
        bp->path_cdf_offsets = ptr(P.path_cdf_offsets);
        bp->path_cdf_counts = ptr(P.path_cdf_counts);
        bp->path_point_offsets = ptr(P.path_point_offsets);
    }
    if (params) *params = P.params;
}

}  // namespace dvgh
