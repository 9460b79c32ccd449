// Oracle (test infrastructure only): f32 restatement of the reference's GPU
// forward kernels and curve helpers (src/gpu/kernels/*.rs).
#pragma once
#include "oracle_common.h"

namespace oracle {
namespace gpu {
float smoothstep_unit(float d);
float rect_signed_distance(float px, float py, float x0, float y0, float x1, float y1);
void distance_to_segment_with_t(float px, float py, float ax, float ay, float bx, float by, float &dist, float &t);
float distance_to_segment(float px, float py, float ax, float ay, float bx, float by);
int solve_quadratic(float a, float b, float c, float out[2]);
int solve_cubic(float a, float b, float c, float d, float out[3]);
void winding_line(float px, float py, float x0, float y0, float x1, float y1, uint32_t rule, float &w, float &cr);
void winding_quadratic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                       uint32_t rule, float &w, float &cr);
void winding_cubic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                   float y3, uint32_t rule, float &w, float &cr);
void closest_quadratic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, bool approx,
                       float &dist, float &t);
void closest_cubic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                   float y3, bool approx, float &dist, float &t);
V4 paint_color(const Paint &p, float px, float py);
float filter_weight(uint32_t type, float dx, float dy, float radius);
void jitter_xy(uint32_t x, uint32_t y, uint32_t sx, uint32_t sy, uint32_t width, uint32_t nsx, uint32_t nsy,
               uint32_t seed, bool jitter, float &rx, float &ry);
}  // namespace gpu

// The 16x16 tile range of every group (4 ints a group; the reference's tile
// lists), for operation counting.
std::vector<int> group_tile_ranges(const Scene &sc, const dvg_render_options &opt);
void forward_samples(const Scene &sc, const dvg_render_options &opt, size_t i0, size_t i1);
// Forward render; sample_colors (optional, whole image only) receives the
// premultiplied per-sample colours in (pixel, sy, sx) order.  Pixel rows
// [y0, y1) are resolved into out_img (row y0 first) from the samples of rows
// [y0 - ri, y1 + ri), ri = ceil(filter radius): every sample whose filter
// reaches those rows, visited in the whole-image order, so the rows are
// bit-identical to those of the whole-image render.
int render_forward(const Scene &sc, const dvg_render_options &opt, float *out_img, float *sample_colors,
                   const char **err, uint32_t y0 = 0, uint32_t y1 = 0xffffffffu);
// Renderer::render_sdf (W*H mean signed distances) and eval_positions (n
// points, x y pairs), the reference GPU's f32 arithmetic.
int render_sdf(const Scene &sc, const dvg_render_options &opt, float *out);
int eval_positions(const Scene &sc, const dvg_render_options &opt, const float *pos, size_t n, float *out);
}  // namespace oracle
