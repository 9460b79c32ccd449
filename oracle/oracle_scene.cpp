// Oracle (test infrastructure only): scene reader, path segmentation and
// flattening, restated from src/scene.rs, src/renderer/path.rs and
// src/geometry.rs.
#include "oracle_common.h"

namespace oracle {

static M3 read_m3(const float *p) {
    M3 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = p[i * 3 + j];
    return r;
}

static void read_paint(const dvg_scene_desc *d, uint32_t kind, const float *color, uint32_t gidx,
                       Paint &out) {
    out.kind = kind;
    out.color = V4(color[0], color[1], color[2], color[3]);
    if (kind == DVG_PAINT_LINEAR || kind == DVG_PAINT_RADIAL) {
        out.grad_index = gidx;
        out.radial = d->gradient_kind[gidx] == 1u;
        const float *gp = d->gradient_params + 4 * gidx;
        out.p0 = V2(gp[0], gp[1]);
        out.p1 = V2(gp[2], gp[3]);
        uint32_t so = d->gradient_stop_offset[gidx], sc = d->gradient_stop_count[gidx];
        out.stop_offset = so;
        for (uint32_t i = 0; i < sc; ++i) {
            Stop s;
            s.offset = d->stop_offsets[so + i];
            const float *c = d->stop_colors + 4 * (so + i);
            s.color = V4(c[0], c[1], c[2], c[3]);
            out.stops.push_back(s);
        }
    }
}

Scene read_scene(const dvg_scene_desc *d) {
    Scene s;
    s.width = d->width;
    s.height = d->height;
    s.background = V4(d->background[0], d->background[1], d->background[2], d->background[3]);
    if (d->background_image) {
        s.has_bg_image = true;
        s.bg_image.assign(d->background_image, d->background_image + d->background_image_len);
    }
    s.filter_type = d->filter_type;
    s.filter_radius = d->filter_radius;
    s.num_gradients = d->num_gradients;
    s.num_stops = d->num_stops;
    s.path_points_len = d->path_points_len;
    s.path_thickness_len = d->path_thickness_len;
    for (uint32_t i = 0; i < d->num_shapes; ++i) {
        Shape sh;
        sh.kind = d->shape_kind[i];
        const float *p = d->shape_params + 4 * i;
        sh.center = V2(p[0], p[1]);
        sh.radius_c = p[2];
        sh.radius_e = V2(p[2], p[3]);
        sh.rmin = V2(p[0], p[1]);
        sh.rmax = V2(p[2], p[3]);
        sh.transform = read_m3(d->shape_transform + 9 * i);
        sh.stroke_width = d->stroke_width[i];
        sh.join = d->stroke_join[i];
        sh.cap = d->stroke_cap[i];
        sh.miter_limit = d->stroke_miter_limit[i];
        if (sh.kind == DVG_SHAPE_PATH) {
            Path &pa = sh.path;
            uint32_t po = d->path_point_offset[i], pc = d->path_point_count[i];
            pa.point_offset = po;
            for (uint32_t k = 0; k < pc; ++k)
                pa.points.push_back(V2(d->path_points[2 * (po + k)], d->path_points[2 * (po + k) + 1]));
            uint32_t co = d->path_ctrl_offset[i], cc = d->path_ctrl_count[i];
            for (uint32_t k = 0; k < cc; ++k) pa.num_control_points.push_back((uint8_t)d->path_num_controls[co + k]);
            uint32_t to = d->path_thickness_offset[i], tc = d->path_thickness_count[i];
            pa.thickness_offset = to;
            if (tc > 0) {
                pa.has_thickness = true;
                for (uint32_t k = 0; k < tc; ++k) pa.thickness.push_back(d->path_thickness[to + k]);
            }
            pa.is_closed = d->path_is_closed[i] != 0;
            pa.use_distance_approx = d->path_use_distance_approx[i] != 0;
        }
        s.shapes.push_back(sh);
    }
    for (uint32_t g = 0; g < d->num_groups; ++g) {
        Group gr;
        uint32_t o = d->group_shape_offset[g], c = d->group_shape_count[g];
        for (uint32_t k = 0; k < c; ++k) gr.shape_indices.push_back(d->group_shape_indices[o + k]);
        read_paint(d, d->fill_kind[g], d->fill_color + 4 * g, d->fill_gradient[g], gr.fill);
        read_paint(d, d->stroke_kind[g], d->stroke_color + 4 * g, d->stroke_gradient[g], gr.stroke);
        gr.fill_rule = d->fill_rule[g];
        gr.shape_to_canvas = read_m3(d->shape_to_canvas + 9 * g);
        gr.canvas_to_shape = read_m3(d->canvas_to_shape + 9 * g);
        s.groups.push_back(gr);
    }
    return s;
}

// path.rs:100-173 (and the identical walk of distance/bvh.rs:251-380).
std::vector<Seg> path_segments(const Path &path, float stroke_width) {
    std::vector<Seg> out;
    if (path.num_control_points.empty() || path.points.empty()) return out;
    size_t point_id = 0;
    uint32_t seg_index = 0;
    for (uint8_t nc : path.num_control_points) {
        Seg s;
        s.kind = nc;
        s.index = seg_index;
        s.point_id = (uint32_t)point_id;
        if (nc == 0) {
            if (!path_point_radius(path, point_id, stroke_width, s.p0, s.r0)) break;
            if (!path_point_radius(path, point_id + 1, stroke_width, s.p1, s.r1)) break;
            s.p2 = s.p1; s.p3 = s.p1; s.r2 = s.r1; s.r3 = s.r1;
            point_id += 1;
        } else if (nc == 1) {
            if (!path_point_radius(path, point_id, stroke_width, s.p0, s.r0)) break;
            if (!path_point_radius(path, point_id + 1, stroke_width, s.p1, s.r1)) break;
            if (!path_point_radius(path, point_id + 2, stroke_width, s.p2, s.r2)) break;
            s.p3 = s.p2; s.r3 = s.r2;
            point_id += 2;
        } else if (nc == 2) {
            if (!path_point_radius(path, point_id, stroke_width, s.p0, s.r0)) break;
            if (!path_point_radius(path, point_id + 1, stroke_width, s.p1, s.r1)) break;
            if (!path_point_radius(path, point_id + 2, stroke_width, s.p2, s.r2)) break;
            if (!path_point_radius(path, point_id + 3, stroke_width, s.p3, s.r3)) break;
            point_id += 3;
        } else {
            break;
        }
        out.push_back(s);
        ++seg_index;
    }
    return out;
}

// geometry.rs:388-401
static float distance_point_line(V2 p, V2 a, V2 b) {
    V2 ab = b - a;
    V2 ap = p - a;
    float area = std::fabs(ab.cross(ap));
    float len = ab.length();
    if (len == 0.0f) return ap.length();
    return area / len;
}

static void flat_quad(V2 p0, V2 p1, V2 p2, float r0, float r1, float r2, float tol, uint32_t depth,
                      float &maxr) {
    // geometry.rs:312-338
    if (depth >= 16 || distance_point_line(p1, p0, p2) <= tol) {
        maxr = rmax(rmax(maxr, std::fabs(r0)), std::fabs(r2));
        return;
    }
    V2 p01 = p0.lerp(p1, 0.5f), p12 = p1.lerp(p2, 0.5f), p012 = p01.lerp(p12, 0.5f);
    float r01 = 0.5f * (r0 + r1), r12 = 0.5f * (r1 + r2), r012 = 0.5f * (r01 + r12);
    flat_quad(p0, p01, p012, r0, r01, r012, tol, depth + 1, maxr);
    flat_quad(p012, p12, p2, r012, r12, r2, tol, depth + 1, maxr);
}

static void flat_cubic(V2 p0, V2 p1, V2 p2, V2 p3, float r0, float r1, float r2, float r3, float tol,
                       uint32_t depth, float &maxr) {
    // geometry.rs:340-375
    float d1 = distance_point_line(p1, p0, p3), d2 = distance_point_line(p2, p0, p3);
    if (depth >= 16 || rmax(d1, d2) <= tol) {
        maxr = rmax(rmax(maxr, std::fabs(r0)), std::fabs(r3));
        return;
    }
    V2 p01 = p0.lerp(p1, 0.5f), p12 = p1.lerp(p2, 0.5f), p23 = p2.lerp(p3, 0.5f);
    V2 p012 = p01.lerp(p12, 0.5f), p123 = p12.lerp(p23, 0.5f), p0123 = p012.lerp(p123, 0.5f);
    float r01 = 0.5f * (r0 + r1), r12 = 0.5f * (r1 + r2), r23 = 0.5f * (r2 + r3);
    float r012 = 0.5f * (r01 + r12), r123 = 0.5f * (r12 + r23), r0123 = 0.5f * (r012 + r123);
    flat_cubic(p0, p01, p012, p0123, r0, r01, r012, r0123, tol, depth + 1, maxr);
    flat_cubic(p0123, p123, p23, p3, r0123, r123, r23, r3, tol, depth + 1, maxr);
}

// prepare.rs:518-524 over geometry.rs:218-305 (flatten_with_thickness with
// thickness_scale 1): the largest |r0|, |r1| of the flattened segments.
float flattened_max_radius(const Path &path, float tolerance, float default_radius) {
    float maxr = 0.0f;
    std::vector<Seg> segs = path_segments(path, default_radius);
    for (const Seg &s : segs) {
        if (s.kind == 0) {
            maxr = rmax(rmax(maxr, std::fabs(s.r0)), std::fabs(s.r1));
        } else if (s.kind == 1) {
            flat_quad(s.p0, s.p1, s.p2, s.r0, s.r1, s.r2, tolerance, 0, maxr);
        } else {
            flat_cubic(s.p0, s.p1, s.p2, s.p3, s.r0, s.r1, s.r2, s.r3, tolerance, 0, maxr);
        }
    }
    return maxr;
}

}  // namespace oracle
