// Test library for the C++ host API (diffvg-rs_amd/host): builds one fixed
// scene with the C++ types and exposes, through a small C interface for
// ctypes, its flat arrays, the gradient tree apply_gpu_grads makes from given
// flat gradients, and the Renderer's results.  tests/test_host_cpp.py builds
// the same scene with pydvg and compares.
#include <cstring>
#include <string>

#include "../diffvg-rs_amd/host/diffvg.hpp"

using namespace diffvg;

namespace {

std::string g_error;

Scene make_scene() {
    Scene sc;
    sc.width = 24;
    sc.height = 20;
    sc.background = {0.9f, 0.95f, 1.f, 1.f};
    std::vector<float> bg((size_t)sc.width * sc.height * 4);
    for (size_t k = 0; k < bg.size(); ++k) bg[k] = (float)(k % 13) / 13.0f;
    sc.background_image = bg;

    Shape s0;
    s0.geometry = Circle{{6.f, 6.f}, 4.f};
    s0.stroke_width = 1.f;
    Shape s1;
    Path p1;
    p1.num_control_points = {2, 2};
    p1.points = {{2.f, 14.f}, {6.f, 8.f}, {10.f, 18.f}, {14.f, 12.f}, {16.f, 8.f}, {20.f, 16.f}, {22.f, 10.f}};
    p1.thickness = std::vector<float>{1.f, 1.25f, 1.5f, 1.75f, 2.f, 1.5f, 1.f};
    s1.geometry = p1;
    s1.stroke_width = 1.5f;
    Shape s2;
    s2.geometry = Ellipse{{15.f, 6.f}, {5.f, 3.f}};
    Shape s3;
    s3.geometry = Rect{{3.f, 10.f}, {9.f, 17.f}};
    s3.stroke_width = 0.75f;
    s3.stroke_join = StrokeJoin::Miter;
    s3.stroke_cap = StrokeCap::Butt;
    Shape s4;
    Path p4;
    p4.num_control_points = {1, 0, 2};
    p4.points = {{12.f, 11.f}, {18.f, 9.f}, {21.f, 14.f}, {19.f, 18.f}, {16.f, 19.f}, {13.f, 17.f}};
    p4.is_closed = true;
    s4.geometry = p4;
    s4.stroke_width = 0.5f;
    s4.stroke_join = StrokeJoin::Bevel;
    s4.stroke_miter_limit = 2.f;
    sc.shapes = {s0, s1, s2, s3, s4};

    LinearGradient lin{{0.f, 0.f}, {8.f, 8.f}, {{0.f, {1.f, 0.f, 0.f, 1.f}}, {1.f, {0.f, 0.f, 1.f, 1.f}}}};
    RadialGradient rad{{15.f, 6.f}, {4.f, 3.f},
                       {{0.f, {1.f, 1.f, 0.f, 1.f}}, {0.5f, {0.f, 1.f, 0.f, 0.8f}}, {1.f, {0.f, 1.f, 1.f, 1.f}}}};
    LinearGradient lin2{{3.f, 10.f}, {21.f, 18.f}, {{0.2f, {0.5f, 0.f, 0.5f, 1.f}}, {0.9f, {0.f, 0.5f, 0.f, 0.6f}}}};
    ShapeGroup g0;
    g0.shape_indices = {0};
    g0.fill = Paint{Color{1.f, 0.f, 0.f, 1.f}};
    g0.stroke = Paint{lin};
    ShapeGroup g1;
    g1.shape_indices = {1};
    g1.stroke = Paint{Color{0.f, 0.f, 1.f, 0.5f}};
    ShapeGroup g2;
    g2.shape_indices = {2};
    g2.fill = Paint{rad};
    Mat3 rot = Mat3::identity();
    rot.m[0][0] = 0.8f; rot.m[0][1] = -0.6f; rot.m[0][2] = 6.6f;
    rot.m[1][0] = 0.6f; rot.m[1][1] = 0.8f; rot.m[1][2] = -7.8f;
    g2.set_shape_to_canvas(rot);
    g2.opacity = 0.7f;
    ShapeGroup g3;
    g3.shape_indices = {3, 4};
    g3.fill = Paint{lin2};
    g3.stroke = Paint{Color{0.2f, 0.3f, 0.4f, 1.f}};
    g3.fill_rule = FillRule::EvenOdd;
    sc.groups = {g0, g1, g2, g3};
    sc.filter = Filter{FilterType::Tent, 1.f};
    return sc;
}

RenderOptions make_options() {
    RenderOptions o;
    o.samples_x = o.samples_y = 2;
    o.seed = 7;
    o.use_prefiltering = false;
    return o;
}

const Scene &scene() {
    static const Scene sc = make_scene();
    return sc;
}

template <class T>
int copy_out(const std::vector<T> &v, double *out, int cap) {
    int n = (int)v.size();
    for (int i = 0; i < n && i < cap; ++i) out[i] = (double)v[i];
    return n;
}

// leaves of a SceneGrad, in declaration order: per shape the geometry's
// fields, the transform, the stroke width; per group the shape_to_canvas,
// then fill and stroke (colour, or the two points then offset + colour per
// stop); then filter radius, background, background image, translation
struct Leaves {
    std::vector<double> v;
    void add(float x) { v.push_back(x); }
    void add(const Vec2 &p) { add(p.x); add(p.y); }
    void add(const Color &c) { add(c.r); add(c.g); add(c.b); add(c.a); }
    void add(const Mat3 &m) {
        for (auto &r : m.m)
            for (float x : r) add(x);
    }
    void add(const std::vector<float> &xs) {
        for (float x : xs) add(x);
    }
    void add_stops(const std::vector<DGradientStop> &st) {
        for (auto &s : st) { add(s.offset); add(s.color); }
    }
    void add(const DPaint &p) {
        if (auto *c = std::get_if<Color>(&p)) add(*c);
        else if (auto *l = std::get_if<DLinearGradient>(&p)) { add(l->start); add(l->end); add_stops(l->stops); }
        else { auto &r = std::get<DRadialGradient>(p); add(r.center); add(r.radius); add_stops(r.stops); }
    }
    explicit Leaves(const SceneGrad &t) {
        for (auto &s : t.shapes) {
            if (auto *c = std::get_if<DCircle>(&s.geometry)) { add(c->center); add(c->radius); }
            else if (auto *e = std::get_if<DEllipse>(&s.geometry)) { add(e->center); add(e->radius); }
            else if (auto *r = std::get_if<DRect>(&s.geometry)) { add(r->min); add(r->max); }
            else {
                auto &p = std::get<DPath>(s.geometry);
                for (auto &q : p.points) add(q);
                if (p.thickness) add(*p.thickness);
            }
            add(s.transform);
            add(s.stroke_width);
        }
        for (auto &g : t.shape_groups) {
            add(g.shape_to_canvas);
            if (g.fill) add(*g.fill);
            if (g.stroke) add(*g.stroke);
        }
        add(t.filter.radius);
        add(t.background);
        if (t.background_image) add(*t.background_image);
        if (t.translation) add(*t.translation);
    }
};

std::vector<std::vector<float> *> grad_arrays(FlatGrads &g) {
    return {&g.d_shape_params, &g.d_shape_points, &g.d_shape_thickness, &g.d_shape_stroke_width,
            &g.d_shape_transform, &g.d_group_transform, &g.d_group_data, &g.d_gradient_data,
            &g.d_stop_offsets, &g.d_stop_colors, &g.d_filter_radius, &g.d_background,
            &g.d_background_image, &g.d_translation};
}

SceneGrad tree_of(const FlatScene &flat, const float *const *arrays) {
    FlatGrads g(flat, true);
    auto arr = grad_arrays(g);
    for (size_t k = 0; k < arr.size(); ++k)
        if (!arr[k]->empty()) std::memcpy(arr[k]->data(), arrays[k], arr[k]->size() * sizeof(float));
    SceneGrad t = SceneGrad::zeros_from_scene(scene(), true, true);
    apply_gpu_grads(scene(), flat, g, t);
    return t;
}

int leaves_out(const SceneGrad &t, double *out, int cap) { return copy_out(Leaves(t).v, out, cap); }

}  // namespace

extern "C" {

const char *hc_last_error() { return g_error.c_str(); }

// one flat array of the scene's FlatScene by its dvg_scene_desc name; -1 for
// an unknown name
int hc_flat_field(const char *name, double *out, int cap) {
    FlatScene f(scene());
    const std::string n(name);
#define U(x) if (n == #x) return copy_out(f.x, out, cap);
    U(shape_kind) U(stroke_join) U(stroke_cap) U(path_point_offset) U(path_point_count) U(path_ctrl_offset)
    U(path_ctrl_count) U(path_thickness_offset) U(path_thickness_count) U(path_is_closed)
    U(path_use_distance_approx) U(path_num_controls) U(group_shape_offset) U(group_shape_count)
    U(group_shape_indices) U(fill_kind) U(stroke_kind) U(fill_gradient) U(stroke_gradient) U(fill_rule)
    U(gradient_kind) U(gradient_stop_offset) U(gradient_stop_count) U(shape_params) U(shape_transform)
    U(stroke_width) U(stroke_miter_limit) U(path_points) U(path_thickness) U(fill_color) U(stroke_color)
    U(shape_to_canvas) U(canvas_to_shape) U(opacity) U(gradient_params) U(stop_offsets) U(stop_colors)
    U(background_image)
#undef U
    const dvg_scene_desc &d = f.desc();
    if (n == "header") {
        std::vector<double> h = {(double)d.width, (double)d.height, d.background[0], d.background[1],
                                 d.background[2], d.background[3], (double)d.filter_type, d.filter_radius,
                                 (double)d.num_shapes, (double)d.num_groups, (double)d.num_gradients,
                                 (double)d.num_stops, (double)d.path_points_len, (double)d.path_num_controls_len,
                                 (double)d.path_thickness_len, (double)d.group_shape_indices_len,
                                 (double)d.background_image_len};
        return copy_out(h, out, cap);
    }
    return -1;
}

// sizes of the 14 flat gradient arrays (compute_translation on)
int hc_grad_sizes(uint64_t *out) {
    FlatScene f(scene());
    FlatGrads g(f, true);
    auto arr = grad_arrays(g);
    for (size_t k = 0; k < arr.size(); ++k) out[k] = arr[k]->size();
    return (int)arr.size();
}

// leaves of apply_gpu_grads(arrays) (14 arrays of hc_grad_sizes)
int hc_tree_leaves(const float *const *arrays, double *out, int cap) {
    FlatScene f(scene());
    return leaves_out(tree_of(f, arrays), out, cap);
}

// leaves of tree(a).accumulate_from(tree(b))
int hc_accumulate_leaves(const float *const *a, const float *const *b, double *out, int cap) {
    FlatScene f(scene());
    SceneGrad ta = tree_of(f, a);
    ta.accumulate_from(tree_of(f, b));
    return leaves_out(ta, out, cap);
}

// leaves of zeros_from_scene(scene, false, false).accumulate_from(tree(b))
int hc_accumulate_into_zeros(const float *const *b, double *out, int cap) {
    FlatScene f(scene());
    SceneGrad z = SceneGrad::zeros_from_scene(scene(), false, false);
    z.accumulate_from(tree_of(f, b));
    return leaves_out(z, out, cap);
}

int hc_inverse(const float *m9, float *out9) {
    Mat3 m;
    for (int i = 0; i < 9; ++i) m.m[i / 3][i % 3] = m9[i];
    auto r = m.inverse();
    if (!r) return 0;
    for (int i = 0; i < 9; ++i) out9[i] = r->m[i / 3][i % 3];
    return 1;
}

// Renderer calls (GPU).  Return 0, or -1 with hc_last_error() set to the
// RenderError's text.
int hc_render(float *rgba) {
    try {
        Renderer r(0);
        Image img = r.render(scene(), make_options());
        std::memcpy(rgba, img.pixels.data(), img.pixels.size() * sizeof(float));
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

int hc_render_sdf(float *sdf) {
    try {
        Renderer r(0);
        auto v = r.render_sdf(scene(), make_options());
        std::memcpy(sdf, v.data(), v.size() * sizeof(float));
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

// render_backward with translation, plus the d_sdf term when d_sdf is non-null
int hc_render_backward(const float *d_render, uint64_t n_render, const float *d_sdf, uint64_t n_sdf, double *out,
                       int cap) {
    try {
        Renderer r(0);
        std::vector<float> dr(d_render, d_render + n_render);
        std::vector<float> ds;
        if (d_sdf) ds.assign(d_sdf, d_sdf + n_sdf);
        BackwardOptions b;
        b.compute_translation = true;
        SceneGrad t = r.render_backward(scene(), make_options(), dr, b, d_sdf ? &ds : nullptr);
        return leaves_out(t, out, cap);
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

int hc_eval_positions(const float *xy, uint64_t n, float *out) {
    try {
        Renderer r(0);
        std::vector<Vec2> pos(n);
        for (uint64_t i = 0; i < n; ++i) pos[i] = {xy[2 * i], xy[2 * i + 1]};
        auto v = r.eval_positions(scene(), make_options(), pos);
        std::memcpy(out, v.data(), v.size() * sizeof(float));
        return 0;
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

int hc_eval_positions_backward(const float *xy, uint64_t n, const float *d_sdf, uint64_t n_sdf, double *out,
                               int cap) {
    try {
        Renderer r(0);
        std::vector<Vec2> pos(n);
        for (uint64_t i = 0; i < n; ++i) pos[i] = {xy[2 * i], xy[2 * i + 1]};
        std::vector<float> ds(d_sdf, d_sdf + n_sdf);
        SceneGrad t = r.eval_positions_backward(scene(), make_options(), pos, ds);
        return leaves_out(t, out, cap);
    } catch (const std::exception &e) {
        g_error = e.what();
        return -1;
    }
}

}  // extern "C"
