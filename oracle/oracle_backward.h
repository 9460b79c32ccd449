// Oracle (test infrastructure only): restatement of the reference CPU backward
// (src/backward/*.rs over src/distance/*.rs).
#pragma once
#include "oracle_common.h"

namespace oracle {

// Flat gradient arrays in the layout of src/renderer/backward_gpu.rs:599-801
// (scene shape / group order).
struct Grads {
    std::vector<float> d_shape_params, d_shape_points, d_shape_thickness, d_shape_stroke_width, d_shape_transform,
        d_group_transform, d_group_data, d_gradient_data, d_stop_offsets, d_stop_colors, d_background_image,
        d_translation;
    // Sums over every sample of the image.  The reference adds them up in f32
    // in sample order (4.2M terms at 1024^2 with 4 samples a pixel, with
    // cancellation between signs of the adjoint), so its rounding reaches
    // ~1e-2 of the result; the oracle keeps them in f64 and rounds once, the
    // value the reference's sum approximates (each sample's own terms are
    // formed in f32 as in the reference).
    double d_filter_radius[1] = {0.0};
    double d_background[4] = {0.0, 0.0, 0.0, 0.0};
};

struct Ctx {
    const Scene *sc = nullptr;
    std::vector<std::vector<Seg>> segs;  // per scene shape, default radius = stroke_width
    std::vector<float> group_bounds;     // 4 per group: union of CPU shape bounds
    std::vector<uint8_t> group_has_bounds;
    // operation counting only: 16x16 tile range of every group (4 ints a
    // group, the reference's tile lists); empty = count every group
    std::vector<int> count_tiles;
};

// Operation counting: counts only while group gi is in the tile list of the
// canvas point's 16x16 tile (no-op in the checking build).
inline void count_gate(const Ctx &cx, bool counting, V2 pt, uint32_t gi) {
    if (!counting || cx.count_tiles.empty()) return;
    [[maybe_unused]] int tx = (int)std::floor(pt.x / 16.f), ty = (int)std::floor(pt.y / 16.f);
    [[maybe_unused]] const int *t = &cx.count_tiles[4 * (size_t)gi];
    COUNT_ON(tx >= t[0] && tx <= t[2] && ty >= t[1] && ty <= t[3]);
}

namespace cpu {
bool solve_quadratic(double a, double b, double c, double t[2]);
int solve_cubic(double a, double b, double c, double d, double t[3]);
V2 transform_point_inverse(const M3 &m, V2 pt);
void distance_to_segment(V2 pt, V2 a, V2 b, float &dist, V2 &cp, float &t);
void closest_point_quadratic(V2 pt, V2 p0, V2 p1, V2 p2, bool approx, V2 &cp, float &t, float &dist);
void closest_point_cubic(V2 pt, V2 p0, V2 p1, V2 p2, V2 p3, bool approx, V2 &cp, float &t, float &dist);
void closest_on_seg(const Seg &s, V2 pt, bool approx, V2 &cp, float &t, float &dist);
V2 closest_point_circle(V2 c, float r, V2 pt);
V2 closest_point_ellipse(V2 c, V2 radius, V2 pt);
V2 closest_point_rect(V2 mn, V2 mx, V2 pt);
float rect_signed_distance(V2 pt, V2 mn, V2 mx);
float ellipse_signed_distance(V2 pt, V2 c, V2 radius);
int winding_number_segment(uint32_t kind, V2 p0, V2 p1, V2 p2, V2 p3, V2 pt);
int winding_number_shape(const Ctx &cx, uint32_t si, V2 pt);
bool within_distance_shape(const Ctx &cx, uint32_t si, V2 pt, float r);
bool within_distance_group(const Ctx &cx, uint32_t gi, V2 pt);
bool is_inside_group(const Ctx &cx, uint32_t gi, V2 pt);
V4 paint_color(const Paint &p, V2 pt);
void d_sample_paint(const Paint &p, Grads &gr, float *d_paint_solid, V2 pt, V4 d, V2 &d_trans);
float compute_filter_weight(uint32_t type, float radius, float dx, float dy);
V4 sample_background(const Scene &sc, size_t pix);
void build_group_bounds(Ctx &cx);
bool cpu_shape_bounds(const Shape &sh, float b[4]);
V2 d_length(V2 v, float d_l);
V2 normalize(V2 v);
void d_xform_pt(const M3 &m, V2 pt, V2 d_out, M3 &d_m, V2 &d_pt);
void accumulate_background_grad(const Scene &sc, Grads &gr, size_t pix, V4 d);
void accumulate_translation(Grads &gr, long pix, V2 d);
std::vector<float> build_weight_image(const Scene &sc, uint32_t nsx, uint32_t nsy, uint32_t seed, bool prefilter,
                                      bool jitter);
V4 gather_d_color(const Scene &sc, const float *d_render, const std::vector<float> &wimg, V2 pt);
void accumulate_filter_gradient(const Scene &sc, V4 color, const float *d_render, const std::vector<float> &wimg,
                                V2 pt, float &d_radius);
void finalize_background_gradients(const Scene &sc, Grads &gr);
// render_backward_positions (backward/render.rs:213-280): gradients of the
// signed distances at n points (x, y pairs) for the adjoint d_sdf.  Defined
// in oracle_prefilter.cpp.
void backward_positions(const Scene &sc, bool compute_translation, const float *pos, size_t n, const float *d_sdf,
                        Grads &gr);
// The d_sdf image term of render_backward (backward/render.rs:171-186).
void backward_sdf_image(const Scene &sc, const dvg_render_options &opt, bool compute_translation, const float *d_sdf,
                        Grads &gr);

// Prefiltered colour of one sample with its adjoint (sampling.rs:227-482);
// gr == nullptr skips the adjoint.  Defined in oracle_prefilter.cpp.
V4 sample_color_prefiltered(const Ctx &cx, V2 screen_pt, const V4 *bg, const V4 *d_color, Grads *gr, long tidx,
                            size_t pix);
void boundary_query_points(const Ctx &cx, uint32_t nsx, uint32_t nsy, uint32_t seed, size_t i0, size_t i1,
                           std::vector<float> &pts);
// Samples whose pixel row lies outside [row_lo, row_hi) are skipped.
void boundary_sampling(const Ctx &cx, uint32_t nsx, uint32_t nsy, uint32_t seed, const float *d_render,
                       const std::vector<float> &wimg, Grads &gr, bool compute_translation, size_t i0, size_t i1,
                       long row_lo = -1, long row_hi = 1L << 40);

}  // namespace cpu

// render_backward (backward/render.rs:35-210) with d_render present.
// Gradients are finalized (background.rs:29-61).
// With a row window [y0, y1) d_render must be zero outside those pixel rows:
// the samples that then gather a zero adjoint (interior samples of rows
// outside [y0 - ri, y1 + ri), boundary samples whose pixel row is more than
// ri + 1 rows away, ri = ceil(filter radius)) add exact zeros and are skipped.
int render_backward(const Scene &sc, const dvg_render_options &opt, bool compute_translation, const float *d_render,
                    Grads &gr, const char **err, uint32_t y0 = 0, uint32_t y1 = 0xffffffffu);

void grads_zero(const Scene &sc, bool with_translation, Grads &gr);
void make_ctx(const Scene &sc, Ctx &cx);
std::vector<float> weight_image(const Scene &sc, const dvg_render_options &opt);
void backward_samples(const Ctx &cx, const dvg_render_options &opt, bool compute_translation, const float *d_render,
                      const std::vector<float> &wimg, Grads &gr, size_t i0, size_t i1);
}  // namespace oracle
