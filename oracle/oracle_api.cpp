// Oracle (test infrastructure only): extern "C" entry points of liboracle.so,
// loaded by tests/ through ctypes.  Never linked into libdvg.
#include "oracle_backward.h"
#include "oracle_forward.h"

#include <chrono>

using namespace oracle;

static void fill_out(const std::vector<float> &v, float *dst) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(float));
}

namespace oracle {
void grads_zero(const Scene &sc, bool with_translation, Grads &gr) {
    size_t S = sc.shapes.size(), G = sc.groups.size();
    gr.d_shape_params.assign(S * 8, 0.f);
    gr.d_shape_points.assign(sc.path_points_len, 0.f);
    gr.d_shape_thickness.assign(sc.path_thickness_len, 0.f);
    gr.d_shape_stroke_width.assign(S, 0.f);
    gr.d_shape_transform.assign(S * 9, 0.f);
    gr.d_group_transform.assign(G * 9, 0.f);
    gr.d_group_data.assign(G * 16, 0.f);
    gr.d_gradient_data.assign((size_t)sc.num_gradients * 8, 0.f);
    gr.d_stop_offsets.assign(sc.num_stops, 0.f);
    gr.d_stop_colors.assign((size_t)sc.num_stops * 4, 0.f);
    gr.d_background_image.assign(sc.has_bg_image ? sc.bg_image.size() : 0, 0.f);
    gr.d_translation.assign(with_translation ? (size_t)sc.width * sc.height * 2 : 0, 0.f);
    gr.d_filter_radius[0] = 0.0;
    for (int i = 0; i < 4; ++i) gr.d_background[i] = 0.0;
}
}  // namespace oracle

extern "C" {

// Forward render with the reference GPU semantics (f32).  sample_colors may be
// NULL; otherwise it receives W*H*spp premultiplied RGBA sample colours.
int oracle_render(const dvg_scene_desc *desc, const dvg_render_options *opt, float *out, float *sample_colors) {
    Scene sc = read_scene(desc);
    const char *err = nullptr;
    return render_forward(sc, *opt, out, sample_colors, &err);
}

// Renderer::render_sdf: W*H signed distances (negative inside a fill).
int oracle_render_sdf(const dvg_scene_desc *desc, const dvg_render_options *opt, float *out) {
    Scene sc = read_scene(desc);
    return render_sdf(sc, *opt, out);
}

// Renderer::eval_positions: the signed distance at n canvas points.
int oracle_eval_positions(const dvg_scene_desc *desc, const dvg_render_options *opt, const float *pos, uint64_t n,
                          float *out) {
    Scene sc = read_scene(desc);
    return eval_positions(sc, *opt, pos, (size_t)n, out);
}

// Pixel rows [y0, y1) of the forward render into out ((y1 - y0) * W * 4),
// bit-identical to the same rows of oracle_render.
int oracle_render_rows(const dvg_scene_desc *desc, const dvg_render_options *opt, uint32_t y0, uint32_t y1,
                       float *out) {
    Scene sc = read_scene(desc);
    const char *err = nullptr;
    return render_forward(sc, *opt, out, nullptr, &err, y0, y1);
}

// Reference CPU render_backward into the flat gradient layout, for a d_render
// that is zero outside pixel rows [y0, y1) (the whole image: 0, UINT32_MAX).
int oracle_render_backward_rows(const dvg_scene_desc *desc, const dvg_render_options *opt,
                                const dvg_backward_options *bopt, const float *d_render, const dvg_grads *out,
                                uint32_t y0, uint32_t y1) {
    Scene sc = read_scene(desc);
    Grads gr;
    bool tr = bopt && bopt->compute_translation;
    grads_zero(sc, tr, gr);
    const char *err = nullptr;
    int st = render_backward(sc, *opt, tr, d_render, gr, &err, y0, y1);
    if (st != DVG_OK) return st;
    fill_out(gr.d_shape_params, out->d_shape_params);
    fill_out(gr.d_shape_points, out->d_shape_points);
    fill_out(gr.d_shape_thickness, out->d_shape_thickness);
    fill_out(gr.d_shape_stroke_width, out->d_shape_stroke_width);
    fill_out(gr.d_shape_transform, out->d_shape_transform);
    fill_out(gr.d_group_transform, out->d_group_transform);
    fill_out(gr.d_group_data, out->d_group_data);
    fill_out(gr.d_gradient_data, out->d_gradient_data);
    fill_out(gr.d_stop_offsets, out->d_stop_offsets);
    fill_out(gr.d_stop_colors, out->d_stop_colors);
    if (out->d_filter_radius) out->d_filter_radius[0] = (float)gr.d_filter_radius[0];
    if (out->d_background)
        for (int i = 0; i < 4; ++i) out->d_background[i] = (float)gr.d_background[i];
    fill_out(gr.d_background_image, out->d_background_image);
    fill_out(gr.d_translation, out->d_translation);
    return DVG_OK;
}

// Renderer::eval_positions_backward (render_backward_positions) into the
// flat gradient layout.
int oracle_eval_positions_backward(const dvg_scene_desc *desc, const dvg_backward_options *bopt, const float *pos,
                                   uint64_t n, const float *d_sdf, const dvg_grads *out) {
    Scene sc = read_scene(desc);
    Grads gr;
    bool tr = bopt && bopt->compute_translation;
    grads_zero(sc, tr, gr);
    cpu::backward_positions(sc, tr, pos, (size_t)n, d_sdf, gr);
    fill_out(gr.d_shape_params, out->d_shape_params);
    fill_out(gr.d_shape_points, out->d_shape_points);
    fill_out(gr.d_shape_thickness, out->d_shape_thickness);
    fill_out(gr.d_shape_stroke_width, out->d_shape_stroke_width);
    fill_out(gr.d_shape_transform, out->d_shape_transform);
    fill_out(gr.d_group_transform, out->d_group_transform);
    fill_out(gr.d_group_data, out->d_group_data);
    fill_out(gr.d_gradient_data, out->d_gradient_data);
    fill_out(gr.d_stop_offsets, out->d_stop_offsets);
    fill_out(gr.d_stop_colors, out->d_stop_colors);
    if (out->d_filter_radius) out->d_filter_radius[0] = (float)gr.d_filter_radius[0];
    if (out->d_background)
        for (int i = 0; i < 4; ++i) out->d_background[i] = (float)gr.d_background[i];
    fill_out(gr.d_background_image, out->d_background_image);
    fill_out(gr.d_translation, out->d_translation);
    return DVG_OK;
}

// The d_sdf image term of render_backward alone (d_render absent): what
// render_backward with d_sdf adds to the d_render gradients.
int oracle_render_backward_sdf(const dvg_scene_desc *desc, const dvg_render_options *opt,
                               const dvg_backward_options *bopt, const float *d_sdf, const dvg_grads *out) {
    Scene sc = read_scene(desc);
    Grads gr;
    bool tr = bopt && bopt->compute_translation;
    grads_zero(sc, tr, gr);
    cpu::backward_sdf_image(sc, *opt, tr, d_sdf, gr);
    fill_out(gr.d_shape_params, out->d_shape_params);
    fill_out(gr.d_shape_points, out->d_shape_points);
    fill_out(gr.d_shape_thickness, out->d_shape_thickness);
    fill_out(gr.d_shape_stroke_width, out->d_shape_stroke_width);
    fill_out(gr.d_shape_transform, out->d_shape_transform);
    fill_out(gr.d_group_transform, out->d_group_transform);
    fill_out(gr.d_group_data, out->d_group_data);
    fill_out(gr.d_gradient_data, out->d_gradient_data);
    fill_out(gr.d_stop_offsets, out->d_stop_offsets);
    fill_out(gr.d_stop_colors, out->d_stop_colors);
    if (out->d_filter_radius) out->d_filter_radius[0] = (float)gr.d_filter_radius[0];
    if (out->d_background)
        for (int i = 0; i < 4; ++i) out->d_background[i] = (float)gr.d_background[i];
    fill_out(gr.d_background_image, out->d_background_image);
    fill_out(gr.d_translation, out->d_translation);
    return DVG_OK;
}

// Reference CPU render_backward into the flat gradient layout.
int oracle_render_backward(const dvg_scene_desc *desc, const dvg_render_options *opt,
                           const dvg_backward_options *bopt, const float *d_render, const dvg_grads *out) {
    return oracle_render_backward_rows(desc, opt, bopt, d_render, out, 0, 0xffffffffu);
}

// ---- primitives, for pinning the restatement against closed forms ----------

void oracle_pcg32(uint64_t idx, uint64_t seed, int n, uint32_t *out_u32, float *out_f32) {
    Pcg32 a(idx, seed), b(idx, seed);
    for (int i = 0; i < n; ++i) {
        out_u32[i] = a.next_u32();
        out_f32[i] = b.next_f32();
    }
}

int oracle_solve_cubic_f64(double a, double b, double c, double d, double *roots) {
    return cpu::solve_cubic(a, b, c, d, roots);
}
int oracle_solve_cubic_f32(float a, float b, float c, float d, float *roots) {
    return gpu::solve_cubic(a, b, c, d, roots);
}

// Closest point on a segment: kind 0/1/2; pts = 8 floats; out = (x, y, t, dist).
void oracle_closest_f64(uint32_t kind, const float *pts, float px, float py, int approx, float *out) {
    Seg s;
    s.kind = kind;
    s.p0 = V2(pts[0], pts[1]); s.p1 = V2(pts[2], pts[3]); s.p2 = V2(pts[4], pts[5]); s.p3 = V2(pts[6], pts[7]);
    V2 cp;
    float t, d;
    cpu::closest_on_seg(s, V2(px, py), approx != 0, cp, t, d);
    out[0] = cp.x; out[1] = cp.y; out[2] = t; out[3] = d;
}
void oracle_closest_f32(uint32_t kind, const float *pts, float px, float py, int approx, float *out) {
    float d = 0.f, t = 0.f;
    if (kind == 0) gpu::distance_to_segment_with_t(px, py, pts[0], pts[1], pts[2], pts[3], d, t);
    else if (kind == 1) gpu::closest_quadratic(px, py, pts[0], pts[1], pts[2], pts[3], pts[4], pts[5], approx != 0, d, t);
    else gpu::closest_cubic(px, py, pts[0], pts[1], pts[2], pts[3], pts[4], pts[5], pts[6], pts[7], approx != 0, d, t);
    out[0] = d; out[1] = t;
}

// Winding of one segment for a +x ray: CPU (integer) and GPU (float) forms.
int oracle_winding_f64(uint32_t kind, const float *pts, float px, float py) {
    return cpu::winding_number_segment(kind, V2(pts[0], pts[1]), V2(pts[2], pts[3]), V2(pts[4], pts[5]),
                                       V2(pts[6], pts[7]), V2(px, py));
}
float oracle_winding_f32(uint32_t kind, const float *pts, float px, float py) {
    float w = 0.f, c = 0.f;
    if (kind == 0) gpu::winding_line(px, py, pts[0], pts[1], pts[2], pts[3], 0, w, c);
    else if (kind == 1) gpu::winding_quadratic(px, py, pts[0], pts[1], pts[2], pts[3], pts[4], pts[5], 0, w, c);
    else gpu::winding_cubic(px, py, pts[0], pts[1], pts[2], pts[3], pts[4], pts[5], pts[6], pts[7], 0, w, c);
    return w;
}

// CPU scene queries (src/distance/api.rs) for the reference's own unit tests.
int oracle_within_distance(const dvg_scene_desc *desc, uint32_t group, float px, float py) {
    Scene sc = read_scene(desc);
    Ctx cx;
    make_ctx(sc, cx);
    return cpu::within_distance_group(cx, group, V2(px, py)) ? 1 : 0;
}
int oracle_is_inside(const dvg_scene_desc *desc, uint32_t group, float px, float py) {
    Scene sc = read_scene(desc);
    Ctx cx;
    make_ctx(sc, cx);
    return cpu::is_inside_group(cx, group, V2(px, py)) ? 1 : 0;
}
// CPU baseline slice: the forward (GPU semantics) and the reference CPU
// backward (interior + boundary) over sample indices [i0, i1) of the full
// workload.  Returns the seconds spent on the per-sample work (the weight
// image, built once over all samples, is excluded and timed separately).
double oracle_time_slice(const dvg_scene_desc *desc, const dvg_render_options *opt, const float *d_render,
                         uint64_t i0, uint64_t i1, double *weight_seconds) {
    Scene sc = read_scene(desc);
    auto t0 = std::chrono::steady_clock::now();
    Ctx cx;
    make_ctx(sc, cx);
    std::vector<float> wimg = weight_image(sc, *opt);
    auto t1 = std::chrono::steady_clock::now();
    Grads gr;
    grads_zero(sc, false, gr);
    forward_samples(sc, *opt, i0, i1);
    uint32_t nsx = std::max(opt->samples_x, 1u), nsy = std::max(opt->samples_y, 1u);
    backward_samples(cx, *opt, false, d_render, wimg, gr, i0, i1);
    cpu::boundary_sampling(cx, nsx, nsy, opt->seed, d_render, wimg, gr, false, i0, i1);
    auto t2 = std::chrono::steady_clock::now();
    if (weight_seconds) *weight_seconds = std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count() * 1e-9;
    return std::chrono::duration_cast<std::chrono::nanoseconds>(t2 - t1).count() * 1e-9;
}

// The scalar whose gradient render_backward returns for a prefiltered scene:
// J = sum over samples of dot(colour, gather_d_color) with the reference CPU
// prefiltered colour (sampling.rs:227-482, no adjoint).  J is linear in the
// per-sample colours with weights independent of the geometry, so central
// differences of J pin the geometry and paint gradients.
double oracle_prefiltered_objective(const dvg_scene_desc *desc, const dvg_render_options *opt,
                                    const float *d_render) {
    Scene sc = read_scene(desc);
    Ctx cx;
    make_ctx(sc, cx);
    std::vector<float> wimg = weight_image(sc, *opt);
    uint32_t nsx = std::max(opt->samples_x, 1u), nsy = std::max(opt->samples_y, 1u);
    size_t W = sc.width, H = sc.height;
    double J = 0.0;
    for (size_t idx = 0; idx < W * H * nsx * nsy; ++idx) {
        size_t sx = idx % nsx, sy = (idx / nsx) % nsy, x = (idx / (nsx * nsy)) % W, y = idx / ((size_t)nsx * nsy * W);
        float px = (float)x + ((float)sx + 0.5f) / (float)nsx;
        float py = (float)y + ((float)sy + 0.5f) / (float)nsy;
        V2 npt(px / (float)W, py / (float)H);
        V4 bg = cpu::sample_background(sc, y * W + x);
        V4 dc = cpu::gather_d_color(sc, d_render, wimg, V2(px, py));
        V4 c = cpu::sample_color_prefiltered(cx, npt, &bg, nullptr, nullptr, -1, y * W + x);
        J += (double)c.x * dc.x + (double)c.y * dc.y + (double)c.z * dc.z + (double)c.w * dc.w;
    }
    return J;
}

float oracle_filter_weight(uint32_t type, float radius, float dx, float dy) {
    return cpu::compute_filter_weight(type, radius, dx, dy);
}

}  // extern "C"
