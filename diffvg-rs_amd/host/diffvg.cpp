// C++ host API over libdvg.so's C ABI (see diffvg.hpp for the reference
// interfaces each part mirrors).
#include "diffvg.hpp"

#include <cmath>
#include <type_traits>

namespace diffvg {

Mat3 Mat3::identity() {
    Mat3 r;
    r.m[0][0] = r.m[1][1] = r.m[2][2] = 1.f;
    return r;
}

// Mat3::inverse (src/math.rs:269-290): the affine inverse in f32, in the
// reference's order of operations; nullopt below its singularity threshold.
std::optional<Mat3> Mat3::inverse() const {
    float a = m[0][0], b = m[0][1], c = m[1][0], d = m[1][1], tx = m[0][2], ty = m[1][2];
    float det = a * d - b * c;
    if (std::fabs(det) < 1.0e-6f) return std::nullopt;
    float inv_det = 1.0f / det;
    Mat3 r;
    r.m[0][0] = d * inv_det;
    r.m[0][1] = -b * inv_det;
    r.m[1][0] = -c * inv_det;
    r.m[1][1] = a * inv_det;
    r.m[0][2] = -(r.m[0][0] * tx + r.m[0][1] * ty);
    r.m[1][2] = -(r.m[1][0] * tx + r.m[1][1] * ty);
    r.m[2][2] = 1.f;
    return r;
}

void ShapeGroup::set_shape_to_canvas(const Mat3 &mat) {
    shape_to_canvas = mat;
    auto inv = mat.inverse();
    canvas_to_shape = inv ? *inv : Mat3::identity();
}

RenderError::RenderError(Kind kind, const std::string &message)
    : std::runtime_error(std::string(kind == Kind::InvalidScene ? "InvalidScene: " : "Launch: ") + message),
      kind_(kind) {}

namespace {

void put_mat(std::vector<float> &v, const Mat3 &m) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) v.push_back(m.m[r][c]);
}

void put_color(std::vector<float> &v, const Color &c) {
    v.push_back(c.r);
    v.push_back(c.g);
    v.push_back(c.b);
    v.push_back(c.a);
}

template <class T>
const T *ptr_or_null(const std::vector<T> &v) {
    return v.empty() ? nullptr : v.data();
}

Mat3 mat_at(const std::vector<float> &v, size_t base) {
    Mat3 m;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) m.m[r][c] = v[base + 3 * r + c];
    return m;
}

}  // namespace

FlatScene::FlatScene(const Scene &scene) {
    const size_t S = scene.shapes.size(), G = scene.groups.size();
    uint32_t npts = 0, nctrl = 0, nthick = 0;
    for (const Shape &sh : scene.shapes) {
        put_mat(shape_transform, sh.transform);
        stroke_width.push_back(sh.stroke_width);
        stroke_join.push_back((uint32_t)sh.stroke_join);
        stroke_cap.push_back((uint32_t)sh.stroke_cap);
        stroke_miter_limit.push_back(sh.stroke_miter_limit);
        float p[4] = {0.f, 0.f, 0.f, 0.f};
        uint32_t kind = DVG_SHAPE_PATH, po = 0, pc = 0, co = 0, cc = 0, to = 0, tc = 0, closed = 0, approx = 0;
        if (auto *c = std::get_if<Circle>(&sh.geometry)) {
            kind = DVG_SHAPE_CIRCLE;
            p[0] = c->center.x; p[1] = c->center.y; p[2] = c->radius;
        } else if (auto *e = std::get_if<Ellipse>(&sh.geometry)) {
            kind = DVG_SHAPE_ELLIPSE;
            p[0] = e->center.x; p[1] = e->center.y; p[2] = e->radius.x; p[3] = e->radius.y;
        } else if (auto *r = std::get_if<Rect>(&sh.geometry)) {
            kind = DVG_SHAPE_RECT;
            p[0] = r->min.x; p[1] = r->min.y; p[2] = r->max.x; p[3] = r->max.y;
        } else {
            const Path &path = std::get<Path>(sh.geometry);
            po = npts; pc = (uint32_t)path.points.size();
            co = nctrl; cc = (uint32_t)path.num_control_points.size();
            for (const Vec2 &q : path.points) {
                path_points.push_back(q.x);
                path_points.push_back(q.y);
            }
            for (uint8_t k : path.num_control_points) path_num_controls.push_back(k);
            npts += pc;
            nctrl += cc;
            to = nthick;
            if (path.thickness) {
                tc = (uint32_t)path.thickness->size();
                path_thickness.insert(path_thickness.end(), path.thickness->begin(), path.thickness->end());
                nthick += tc;
            }
            closed = path.is_closed ? 1u : 0u;
            approx = path.use_distance_approx ? 1u : 0u;
        }
        shape_kind.push_back(kind);
        shape_params.insert(shape_params.end(), p, p + 4);
        path_point_offset.push_back(po);
        path_point_count.push_back(pc);
        path_ctrl_offset.push_back(co);
        path_ctrl_count.push_back(cc);
        path_thickness_offset.push_back(to);
        path_thickness_count.push_back(tc);
        path_is_closed.push_back(closed);
        path_use_distance_approx.push_back(approx);
    }
    // groups; gradient paints numbered fill before stroke, in group order
    auto add_paint = [&](const std::optional<Paint> &paint, std::vector<uint32_t> &kinds, std::vector<float> &colors,
                         std::vector<uint32_t> &grads) {
        uint32_t kind = DVG_PAINT_NONE, gidx = 0;
        Color solid;
        if (paint) {
            if (auto *c = std::get_if<Color>(&*paint)) {
                kind = DVG_PAINT_SOLID;
                solid = *c;
            } else {
                const std::vector<GradientStop> *stops;
                gidx = (uint32_t)gradient_kind.size();
                if (auto *l = std::get_if<LinearGradient>(&*paint)) {
                    kind = DVG_PAINT_LINEAR;
                    gradient_kind.push_back(0);
                    float q[4] = {l->start.x, l->start.y, l->end.x, l->end.y};
                    gradient_params.insert(gradient_params.end(), q, q + 4);
                    stops = &l->stops;
                } else {
                    const RadialGradient &rg = std::get<RadialGradient>(*paint);
                    kind = DVG_PAINT_RADIAL;
                    gradient_kind.push_back(1);
                    float q[4] = {rg.center.x, rg.center.y, rg.radius.x, rg.radius.y};
                    gradient_params.insert(gradient_params.end(), q, q + 4);
                    stops = &rg.stops;
                }
                gradient_stop_offset.push_back((uint32_t)stop_offsets.size());
                gradient_stop_count.push_back((uint32_t)stops->size());
                for (const GradientStop &st : *stops) {
                    stop_offsets.push_back(st.offset);
                    put_color(stop_colors, st.color);
                }
            }
        }
        kinds.push_back(kind);
        put_color(colors, solid);
        grads.push_back(gidx);
    };
    for (const ShapeGroup &g : scene.groups) {
        group_shape_offset.push_back((uint32_t)group_shape_indices.size());
        group_shape_count.push_back((uint32_t)g.shape_indices.size());
        group_shape_indices.insert(group_shape_indices.end(), g.shape_indices.begin(), g.shape_indices.end());
        add_paint(g.fill, fill_kind, fill_color, fill_gradient);
        add_paint(g.stroke, stroke_kind, stroke_color, stroke_gradient);
        fill_rule.push_back((uint32_t)g.fill_rule);
        put_mat(shape_to_canvas, g.shape_to_canvas);
        put_mat(canvas_to_shape, g.canvas_to_shape);
        opacity.push_back(g.opacity);
    }
    if (scene.background_image) background_image = *scene.background_image;

    dvg_scene_desc &d = desc_;
    d.width = scene.width;
    d.height = scene.height;
    d.background[0] = scene.background.r;
    d.background[1] = scene.background.g;
    d.background[2] = scene.background.b;
    d.background[3] = scene.background.a;
    d.background_image = scene.background_image ? background_image.data() : nullptr;
    d.background_image_len = scene.background_image ? background_image.size() : 0;
    d.filter_type = (uint32_t)scene.filter.filter_type;
    d.filter_radius = scene.filter.radius;
    d.num_shapes = (uint32_t)S;
    d.shape_kind = ptr_or_null(shape_kind);
    d.shape_params = ptr_or_null(shape_params);
    d.shape_transform = ptr_or_null(shape_transform);
    d.stroke_width = ptr_or_null(stroke_width);
    d.stroke_join = ptr_or_null(stroke_join);
    d.stroke_cap = ptr_or_null(stroke_cap);
    d.stroke_miter_limit = ptr_or_null(stroke_miter_limit);
    d.path_point_offset = ptr_or_null(path_point_offset);
    d.path_point_count = ptr_or_null(path_point_count);
    d.path_ctrl_offset = ptr_or_null(path_ctrl_offset);
    d.path_ctrl_count = ptr_or_null(path_ctrl_count);
    d.path_thickness_offset = ptr_or_null(path_thickness_offset);
    d.path_thickness_count = ptr_or_null(path_thickness_count);
    d.path_is_closed = ptr_or_null(path_is_closed);
    d.path_use_distance_approx = ptr_or_null(path_use_distance_approx);
    d.path_points = ptr_or_null(path_points);
    d.path_points_len = path_points.size();
    d.path_num_controls = ptr_or_null(path_num_controls);
    d.path_num_controls_len = path_num_controls.size();
    d.path_thickness = ptr_or_null(path_thickness);
    d.path_thickness_len = path_thickness.size();
    d.num_groups = (uint32_t)G;
    d.group_shape_offset = ptr_or_null(group_shape_offset);
    d.group_shape_count = ptr_or_null(group_shape_count);
    d.group_shape_indices = ptr_or_null(group_shape_indices);
    d.group_shape_indices_len = group_shape_indices.size();
    d.fill_kind = ptr_or_null(fill_kind);
    d.stroke_kind = ptr_or_null(stroke_kind);
    d.fill_color = ptr_or_null(fill_color);
    d.stroke_color = ptr_or_null(stroke_color);
    d.fill_gradient = ptr_or_null(fill_gradient);
    d.stroke_gradient = ptr_or_null(stroke_gradient);
    d.fill_rule = ptr_or_null(fill_rule);
    d.shape_to_canvas = ptr_or_null(shape_to_canvas);
    d.canvas_to_shape = ptr_or_null(canvas_to_shape);
    d.opacity = ptr_or_null(opacity);
    d.num_gradients = (uint32_t)gradient_kind.size();
    d.gradient_kind = ptr_or_null(gradient_kind);
    d.gradient_params = ptr_or_null(gradient_params);
    d.gradient_stop_offset = ptr_or_null(gradient_stop_offset);
    d.gradient_stop_count = ptr_or_null(gradient_stop_count);
    d.num_stops = (uint32_t)stop_offsets.size();
    d.stop_offsets = ptr_or_null(stop_offsets);
    d.stop_colors = ptr_or_null(stop_colors);
}

FlatGrads::FlatGrads(const FlatScene &flat, bool compute_translation) {
    const dvg_scene_desc &d = flat.desc();
    d_shape_params.assign(8ull * d.num_shapes, 0.f);
    d_shape_points.assign(d.path_points_len, 0.f);
    d_shape_thickness.assign(d.path_thickness_len, 0.f);
    d_shape_stroke_width.assign(d.num_shapes, 0.f);
    d_shape_transform.assign(9ull * d.num_shapes, 0.f);
    d_group_transform.assign(9ull * d.num_groups, 0.f);
    d_group_data.assign(16ull * d.num_groups, 0.f);
    d_gradient_data.assign(8ull * d.num_gradients, 0.f);
    d_stop_offsets.assign(d.num_stops, 0.f);
    d_stop_colors.assign(4ull * d.num_stops, 0.f);
    d_filter_radius.assign(1, 0.f);
    d_background.assign(4, 0.f);
    d_background_image.assign(d.background_image ? (size_t)d.width * d.height * 4 : 0, 0.f);
    d_translation.assign(compute_translation ? (size_t)d.width * d.height * 2 : 0, 0.f);
}

dvg_grads FlatGrads::view() {
    dvg_grads g;
    g.d_shape_params = d_shape_params.data();
    g.d_shape_points = d_shape_points.data();
    g.d_shape_thickness = d_shape_thickness.data();
    g.d_shape_stroke_width = d_shape_stroke_width.data();
    g.d_shape_transform = d_shape_transform.data();
    g.d_group_transform = d_group_transform.data();
    g.d_group_data = d_group_data.data();
    g.d_gradient_data = d_gradient_data.data();
    g.d_stop_offsets = d_stop_offsets.data();
    g.d_stop_colors = d_stop_colors.data();
    g.d_filter_radius = d_filter_radius.data();
    g.d_background = d_background.data();
    g.d_background_image = d_background_image.empty() ? nullptr : d_background_image.data();
    g.d_translation = d_translation.empty() ? nullptr : d_translation.data();
    return g;
}

// ---- gradient tree --------------------------------------------------------------

namespace {

DPaint zero_paint(const Paint &p) {
    if (std::holds_alternative<Color>(p)) return Color{};
    if (auto *l = std::get_if<LinearGradient>(&p)) return DLinearGradient{{}, {}, std::vector<DGradientStop>(l->stops.size())};
    return DRadialGradient{{}, {}, std::vector<DGradientStop>(std::get<RadialGradient>(p).stops.size())};
}

void add(Vec2 &a, const Vec2 &b) { a.x += b.x; a.y += b.y; }
void add(Color &a, const Color &b) { a.r += b.r; a.g += b.g; a.b += b.b; a.a += b.a; }
void add(Mat3 &a, const Mat3 &b) {
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) a.m[r][c] += b.m[r][c];
}

void add_floats(std::vector<float> &a, const std::vector<float> &b) {
    size_t n = std::min(a.size(), b.size());
    for (size_t i = 0; i < n; ++i) a[i] += b[i];
}

void add_optional(std::optional<std::vector<float>> &a, const std::optional<std::vector<float>> &b) {
    if (a) {
        if (b) add_floats(*a, *b);
    } else if (b) {
        a = b;
    }
}

void add_stops(std::vector<DGradientStop> &a, const std::vector<DGradientStop> &b) {
    size_t n = std::min(a.size(), b.size());
    for (size_t i = 0; i < n; ++i) {
        a[i].offset += b[i].offset;
        add(a[i].color, b[i].color);
    }
}

void add_paint(std::optional<DPaint> &a, const std::optional<DPaint> &b) {
    if (!b) return;
    if (!a) {
        a = b;
        return;
    }
    if (auto *x = std::get_if<Color>(&*a)) {
        if (auto *y = std::get_if<Color>(&*b)) add(*x, *y);
    } else if (auto *x = std::get_if<DLinearGradient>(&*a)) {
        if (auto *y = std::get_if<DLinearGradient>(&*b)) {
            add(x->start, y->start);
            add(x->end, y->end);
            add_stops(x->stops, y->stops);
        }
    } else if (auto *x = std::get_if<DRadialGradient>(&*a)) {
        if (auto *y = std::get_if<DRadialGradient>(&*b)) {
            add(x->center, y->center);
            add(x->radius, y->radius);
            add_stops(x->stops, y->stops);
        }
    }
}

void add_geometry(DShapeGeometry &a, const DShapeGeometry &b) {
    if (auto *x = std::get_if<DCircle>(&a)) {
        if (auto *y = std::get_if<DCircle>(&b)) { add(x->center, y->center); x->radius += y->radius; }
    } else if (auto *x = std::get_if<DEllipse>(&a)) {
        if (auto *y = std::get_if<DEllipse>(&b)) { add(x->center, y->center); add(x->radius, y->radius); }
    } else if (auto *x = std::get_if<DRect>(&a)) {
        if (auto *y = std::get_if<DRect>(&b)) { add(x->min, y->min); add(x->max, y->max); }
    } else if (auto *x = std::get_if<DPath>(&a)) {
        if (auto *y = std::get_if<DPath>(&b)) {
            size_t n = std::min(x->points.size(), y->points.size());
            for (size_t i = 0; i < n; ++i) add(x->points[i], y->points[i]);
            add_optional(x->thickness, y->thickness);
        }
    }
}

}  // namespace

SceneGrad SceneGrad::zeros_from_scene(const Scene &scene, bool include_background_image, bool include_translation) {
    SceneGrad g;
    for (const Shape &s : scene.shapes) {
        DShape d;
        if (std::holds_alternative<Circle>(s.geometry)) d.geometry = DCircle{};
        else if (std::holds_alternative<Ellipse>(s.geometry)) d.geometry = DEllipse{};
        else if (std::holds_alternative<Rect>(s.geometry)) d.geometry = DRect{};
        else {
            const Path &p = std::get<Path>(s.geometry);
            DPath dp;
            dp.points.assign(p.points.size(), Vec2{});
            if (p.thickness) dp.thickness = std::vector<float>(p.thickness->size(), 0.f);
            d.geometry = dp;
        }
        g.shapes.push_back(d);
    }
    for (const ShapeGroup &grp : scene.groups) {
        DShapeGroup d;
        if (grp.fill) d.fill = zero_paint(*grp.fill);
        if (grp.stroke) d.stroke = zero_paint(*grp.stroke);
        g.shape_groups.push_back(d);
    }
    if (include_background_image && scene.background_image)
        g.background_image = std::vector<float>(scene.background_image->size(), 0.f);
    if (include_translation) g.translation = std::vector<float>((size_t)scene.width * scene.height * 2, 0.f);
    return g;
}

void SceneGrad::accumulate_from(const SceneGrad &other) {
    size_t n = std::min(shapes.size(), other.shapes.size());
    for (size_t i = 0; i < n; ++i) {
        shapes[i].stroke_width += other.shapes[i].stroke_width;
        add(shapes[i].transform, other.shapes[i].transform);
        add_geometry(shapes[i].geometry, other.shapes[i].geometry);
    }
    size_t m = std::min(shape_groups.size(), other.shape_groups.size());
    for (size_t i = 0; i < m; ++i) {
        add(shape_groups[i].shape_to_canvas, other.shape_groups[i].shape_to_canvas);
        add_paint(shape_groups[i].fill, other.shape_groups[i].fill);
        add_paint(shape_groups[i].stroke, other.shape_groups[i].stroke);
    }
    filter.radius += other.filter.radius;
    add(background, other.background);
    add_optional(background_image, other.background_image);
    add_optional(translation, other.translation);
}

namespace {

float at(const std::vector<float> &v, size_t i) { return i < v.size() ? v[i] : 0.f; }

// build_paint_grad (src/renderer/backward_gpu.rs:803-864)
DPaint paint_grad(const Paint &paint, uint32_t kind, uint32_t grad_index, size_t color_base, const FlatScene &flat,
                  const FlatGrads &g) {
    if (std::holds_alternative<Color>(paint))
        return Color{at(g.d_group_data, color_base), at(g.d_group_data, color_base + 1),
                     at(g.d_group_data, color_base + 2), at(g.d_group_data, color_base + 3)};
    if (kind == DVG_PAINT_NONE) return Color{};
    size_t base = (size_t)grad_index * 8;
    Vec2 p0{at(g.d_gradient_data, base + 1), at(g.d_gradient_data, base + 2)};
    Vec2 p1{at(g.d_gradient_data, base + 3), at(g.d_gradient_data, base + 4)};
    uint32_t so = flat.gradient_stop_offset[grad_index], sc = flat.gradient_stop_count[grad_index];
    std::vector<DGradientStop> stops(sc);
    for (uint32_t i = 0; i < sc; ++i) {
        size_t o = so + i;
        stops[i].offset = at(g.d_stop_offsets, o);
        stops[i].color = Color{at(g.d_stop_colors, 4 * o), at(g.d_stop_colors, 4 * o + 1),
                               at(g.d_stop_colors, 4 * o + 2), at(g.d_stop_colors, 4 * o + 3)};
    }
    if (flat.gradient_kind[grad_index] == 0) return DLinearGradient{p0, p1, stops};
    return DRadialGradient{p0, p1, stops};
}

}  // namespace

void apply_gpu_grads(const Scene &scene, const FlatScene &flat, const FlatGrads &g, SceneGrad &grads) {
    for (size_t i = 0; i < scene.shapes.size() && i < grads.shapes.size(); ++i) {
        DShape &d = grads.shapes[i];
        if (i < g.d_shape_stroke_width.size()) d.stroke_width = g.d_shape_stroke_width[i];
        if (9 * i + 8 < g.d_shape_transform.size()) d.transform = mat_at(g.d_shape_transform, 9 * i);
        size_t base = 8 * i;
        const ShapeGeometry &geom = scene.shapes[i].geometry;
        if (std::holds_alternative<Circle>(geom)) {
            if (auto *c = std::get_if<DCircle>(&d.geometry); c && base + 2 < g.d_shape_params.size()) {
                c->center = {g.d_shape_params[base], g.d_shape_params[base + 1]};
                c->radius = g.d_shape_params[base + 2];
            }
        } else if (std::holds_alternative<Ellipse>(geom)) {
            if (auto *e = std::get_if<DEllipse>(&d.geometry); e && base + 3 < g.d_shape_params.size()) {
                e->center = {g.d_shape_params[base], g.d_shape_params[base + 1]};
                e->radius = {g.d_shape_params[base + 2], g.d_shape_params[base + 3]};
            }
        } else if (std::holds_alternative<Rect>(geom)) {
            if (auto *r = std::get_if<DRect>(&d.geometry); r && base + 3 < g.d_shape_params.size()) {
                r->min = {g.d_shape_params[base], g.d_shape_params[base + 1]};
                r->max = {g.d_shape_params[base + 2], g.d_shape_params[base + 3]};
            }
        } else if (auto *p = std::get_if<DPath>(&d.geometry)) {
            uint32_t po = flat.path_point_offset[i], pc = flat.path_point_count[i];
            for (uint32_t k = 0; k < pc && k < p->points.size(); ++k) {
                size_t b = (size_t)(po + k) * 2;
                if (b + 1 < g.d_shape_points.size()) p->points[k] = {g.d_shape_points[b], g.d_shape_points[b + 1]};
            }
            if (p->thickness) {
                uint32_t to = flat.path_thickness_offset[i], tc = flat.path_thickness_count[i];
                for (uint32_t k = 0; k < tc && k < p->thickness->size(); ++k)
                    if (to + k < g.d_shape_thickness.size()) (*p->thickness)[k] = g.d_shape_thickness[to + k];
            }
        }
    }
    for (size_t gi = 0; gi < scene.groups.size() && gi < grads.shape_groups.size(); ++gi) {
        DShapeGroup &d = grads.shape_groups[gi];
        if (9 * gi + 8 < g.d_group_transform.size()) d.shape_to_canvas = mat_at(g.d_group_transform, 9 * gi);
        const ShapeGroup &grp = scene.groups[gi];
        if (grp.fill)
            d.fill = paint_grad(*grp.fill, flat.fill_kind[gi], flat.fill_gradient[gi], 16 * gi + 8, flat, g);
        if (grp.stroke)
            d.stroke = paint_grad(*grp.stroke, flat.stroke_kind[gi], flat.stroke_gradient[gi], 16 * gi + 12, flat, g);
    }
    if (!g.d_filter_radius.empty()) grads.filter.radius = g.d_filter_radius[0];
    if (g.d_background.size() >= 4)
        grads.background = {g.d_background[0], g.d_background[1], g.d_background[2], g.d_background[3]};
    if (grads.background_image && g.d_background_image.size() >= grads.background_image->size())
        std::copy(g.d_background_image.begin(), g.d_background_image.begin() + grads.background_image->size(),
                  grads.background_image->begin());
    if (grads.translation && g.d_translation.size() >= grads.translation->size())
        std::copy(g.d_translation.begin(), g.d_translation.begin() + grads.translation->size(),
                  grads.translation->begin());
}

// ---- Renderer ---------------------------------------------------------------------

namespace {

dvg_render_options c_options(const RenderOptions &o) {
    dvg_render_options r;
    r.aa = o.aa;
    r.path_tolerance = o.path_tolerance;
    r.samples_x = o.samples_x;
    r.samples_y = o.samples_y;
    r.seed = o.seed;
    r.jitter = o.jitter ? 1u : 0u;
    r.use_prefiltering = o.use_prefiltering ? 1u : 0u;
    return r;
}

std::vector<float> flat_positions(const std::vector<Vec2> &positions) {
    std::vector<float> v;
    v.reserve(positions.size() * 2);
    for (const Vec2 &p : positions) {
        v.push_back(p.x);
        v.push_back(p.y);
    }
    return v;
}

SceneGrad tree(const Scene &scene, const FlatScene &flat, const FlatGrads &g) {
    SceneGrad t = SceneGrad::zeros_from_scene(scene, !g.d_background_image.empty(), !g.d_translation.empty());
    apply_gpu_grads(scene, flat, g, t);
    return t;
}

}  // namespace

Renderer::Renderer(int device_ordinal) {
    ctx_ = dvg_ctx_create(device_ordinal);
    if (!ctx_) throw RenderError(RenderError::Kind::Launch, "no usable HIP device");
}

Renderer::~Renderer() {
    if (ctx_) dvg_ctx_destroy(ctx_);
}

void Renderer::check(int32_t status) const {
    if (status == DVG_OK) return;
    const char *msg = dvg_last_error(ctx_);
    throw RenderError(status == DVG_INVALID_SCENE ? RenderError::Kind::InvalidScene : RenderError::Kind::Launch,
                      msg ? msg : "");
}

void Renderer::set_exact_boundary(bool on) { check(dvg_ctx_set_exact_boundary(ctx_, on ? 1 : 0)); }

Image Renderer::render(const Scene &scene, const RenderOptions &options) {
    FlatScene flat(scene);
    dvg_render_options o = c_options(options);
    Image img{scene.width, scene.height, std::vector<float>((size_t)scene.width * scene.height * 4)};
    check(dvg_render(ctx_, &flat.desc(), &o, img.pixels.data()));
    return img;
}

std::vector<float> Renderer::render_sdf(const Scene &scene, const RenderOptions &options) {
    FlatScene flat(scene);
    dvg_render_options o = c_options(options);
    std::vector<float> out((size_t)scene.width * scene.height);
    check(dvg_render_sdf(ctx_, &flat.desc(), &o, out.data()));
    return out;
}

std::vector<float> Renderer::eval_positions(const Scene &scene, const RenderOptions &options,
                                            const std::vector<Vec2> &positions) {
    FlatScene flat(scene);
    dvg_render_options o = c_options(options);
    std::vector<float> pos = flat_positions(positions), out(positions.size());
    check(dvg_eval_positions(ctx_, &flat.desc(), &o, pos.data(), positions.size(), out.data()));
    return out;
}

void Renderer::render_backward_flat(const Scene &scene, const RenderOptions &options,
                                    const std::vector<float> &d_render, const BackwardOptions &backward_options,
                                    const FlatScene &flat, FlatGrads &out) {
    if (d_render.size() != (size_t)scene.width * scene.height * 4)
        throw RenderError(RenderError::Kind::InvalidScene, "d_render_image size mismatch");
    dvg_render_options o = c_options(options);
    dvg_backward_options b{backward_options.compute_translation ? 1u : 0u};
    dvg_grads gv = out.view();
    check(dvg_render_backward(ctx_, &flat.desc(), &o, &b, d_render.data(), &gv));
}

SceneGrad Renderer::render_backward(const Scene &scene, const RenderOptions &options,
                                    const std::vector<float> &d_render, const BackwardOptions &backward_options,
                                    const std::vector<float> *d_sdf) {
    FlatScene flat(scene);
    FlatGrads g(flat, backward_options.compute_translation);
    render_backward_flat(scene, options, d_render, backward_options, flat, g);
    SceneGrad t = tree(scene, flat, g);
    if (d_sdf) {
        // the d_sdf term (backward_gpu.rs:512-521), added as accumulate_from
        if (d_sdf->size() != (size_t)scene.width * scene.height)
            throw RenderError(RenderError::Kind::InvalidScene, "d_sdf_image size mismatch");
        FlatGrads gs(flat, backward_options.compute_translation);
        dvg_render_options o = c_options(options);
        dvg_backward_options b{backward_options.compute_translation ? 1u : 0u};
        dvg_grads gv = gs.view();
        check(dvg_render_backward_sdf(ctx_, &flat.desc(), &o, &b, d_sdf->data(), &gv));
        t.accumulate_from(tree(scene, flat, gs));
    }
    return t;
}

SceneGrad Renderer::eval_positions_backward(const Scene &scene, const RenderOptions &options,
                                            const std::vector<Vec2> &positions, const std::vector<float> &d_sdf,
                                            const BackwardOptions &backward_options) {
    if (d_sdf.size() != positions.size())
        throw RenderError(RenderError::Kind::InvalidScene, "d_sdf_image size mismatch for eval_positions");
    FlatScene flat(scene);
    FlatGrads g(flat, backward_options.compute_translation);
    dvg_render_options o = c_options(options);
    dvg_backward_options b{backward_options.compute_translation ? 1u : 0u};
    std::vector<float> pos = flat_positions(positions);
    dvg_grads gv = g.view();
    check(dvg_eval_positions_backward(ctx_, &flat.desc(), &o, &b, pos.data(), positions.size(), d_sdf.data(), &gv));
    return tree(scene, flat, g);
}

}  // namespace diffvg
