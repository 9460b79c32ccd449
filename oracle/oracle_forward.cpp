// Oracle (test infrastructure only): f32 restatement of the reference's GPU
// forward, the only forward renderer it has.
//   rasterize_weights_f32  src/gpu/kernels/forward.rs:1506-1565
//   rasterize_splat_f32    src/gpu/kernels/forward.rs:1568-1735
//   eval_scene_tiled       src/gpu/kernels/forward.rs:1133-1290
//   resolve_splat_f32      src/gpu/kernels/forward.rs:295-334
// Tile binning and both BVH levels only cull conservatively, so they are
// replaced here by loops over groups, shapes and segments; the per-shape
// bounds test (forward.rs:1442-1452) is kept because it changes prefiltered
// stroke results.
#include "oracle_forward.h"

namespace oracle {

float g_forward_sink = 0.f;

namespace gpu {

// ---- math.rs ---------------------------------------------------------------
static inline float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }
static inline float min_f(float a, float b) { return a < b ? a : b; }  // math.rs:28
static inline float max_f(float a, float b) { return a > b ? a : b; }  // math.rs:34
static inline float abs_f(float a) { return a < 0.f ? -a : a; }        // math.rs:40
float smoothstep_unit(float d) {                                          // math.rs:7-10
    float t = clamp01((d + 1.0f) * 0.5f);
    return t * t * (3.0f - 2.0f * t);
}
float rect_signed_distance(float px, float py, float x0, float y0, float x1, float y1) {  // math.rs:47-64
    float dx = max_f(max_f(x0 - px, 0.f), px - x1);
    float dy = max_f(max_f(y0 - py, 0.f), py - y1);
    float outside = sqrt_f(dx * dx + dy * dy);
    float inside = min_f(min_f(px - x0, x1 - px), min_f(py - y0, y1 - py));
    return outside > 0.f ? outside : -inside;
}

// ---- curves.rs ---------------------------------------------------------------
static inline float dist_pt(float px, float py, float ax, float ay) {  // curves.rs:30-35
    float dx = px - ax, dy = py - ay;
    return sqrt_f(dx * dx + dy * dy);
}

void distance_to_segment_with_t(float px, float py, float ax, float ay, float bx, float by, float &dist,
                                float &t) {  // curves.rs:39-66
    float vx = bx - ax, vy = by - ay, wx = px - ax, wy = py - ay;
    float c1 = vx * wx + vy * wy;
    float c2 = vx * vx + vy * vy;
    t = 0.f;
    if (c2 > 0.f && c1 > 0.f) {
        t = c1 / c2;
        t = clamp01(t);
    }
    float qx = ax + t * vx, qy = ay + t * vy;
    float d2 = (px - qx) * (px - qx) + (py - qy) * (py - qy);
    dist = sqrt_f(d2);
}

float distance_to_segment(float px, float py, float ax, float ay, float bx, float by) {  // curves.rs:70-93
    float vx = bx - ax, vy = by - ay, wx = px - ax, wy = py - ay;
    float c1 = vx * wx + vy * wy;
    float c2 = vx * vx + vy * vy;
    float d2 = (px - ax) * (px - ax) + (py - ay) * (py - ay);
    if (c1 > 0.f) {
        if (c2 <= c1) {
            d2 = (px - bx) * (px - bx) + (py - by) * (py - by);
        } else {
            float t = c1 / c2;
            float qx = ax + t * vx, qy = ay + t * vy;
            d2 = (px - qx) * (px - qx) + (py - qy) * (py - qy);
        }
    }
    return sqrt_f(d2);
}

int solve_quadratic(float a, float b, float c, float out[2]) {  // curves.rs:865-893
    float disc = b * b - 4.0f * a * c;
    if (disc < 0.f) { out[0] = out[1] = 0.f; return 0; }
    float root = sqrt_f(disc);
    float q = (b < 0.f) ? -0.5f * (b - root) : -0.5f * (b + root);
    float t0 = q / a, t1 = c / q;
    float lo = t0, hi = t1;
    if (lo > hi) { float tmp = lo; lo = hi; hi = tmp; }
    out[0] = lo; out[1] = hi;
    return 2;
}

static float cbrt_gpu(float x) {  // curves.rs:949-959
    const float third = 1.0f / 3.0f;
    if (x > 0.f) return pow_f(x, third);
    if (x < 0.f) return -pow_f(-x, third);
    return 0.f;
}

int solve_cubic(float a, float b, float c, float d, float out[3]) {  // curves.rs:897-946
    const float eps = 1.0e-6f;
    if (abs_f(a) < eps) {
        float r[2];
        int n = solve_quadratic(b, c, d, r);
        out[0] = r[0]; out[1] = r[1]; out[2] = 0.f;
        return n;
    }
    float bb = b / a, cc = c / a, dd = d / a;
    const float third = 1.0f / 3.0f;
    float q = (bb * bb - 3.0f * cc) / 9.0f;
    float r = (2.0f * bb * bb * bb - 9.0f * bb * cc + 27.0f * dd) / 54.0f;
    float r2 = r * r, q3 = q * q * q;
    if (r2 < q3) {
        float theta = acos_f(r / sqrt_f(q3));
        float sqrt_q = sqrt_f(q);
        const float two_pi = 2.0f * 3.14159265f;
        out[0] = -2.0f * sqrt_q * cos_f(theta * third) - bb * third;
        out[1] = -2.0f * sqrt_q * cos_f((theta + two_pi) * third) - bb * third;
        out[2] = -2.0f * sqrt_q * cos_f((theta - two_pi) * third) - bb * third;
        return 3;
    }
    float a_root = (r > 0.f) ? -cbrt_gpu(r + sqrt_f(r2 - q3)) : cbrt_gpu(-r + sqrt_f(r2 - q3));
    float b_root = (abs_f(a_root) > eps) ? q / a_root : 0.f;
    out[0] = (a_root + b_root) - bb * third;
    out[1] = 0.f; out[2] = 0.f;
    return 1;
}

void winding_line(float px, float py, float x0, float y0, float x1, float y1, uint32_t rule, float &w,
                  float &cr) {  // curves.rs:97-123
    bool a = y0 <= py, b = y1 <= py;
    if (a != b) {
        float t = (py - y0) / (y1 - y0);
        float xi = x0 + t * (x1 - x0);
        if (xi > px) {
            if (rule == 1) cr = 1.f - cr;
            else w += (y1 > y0) ? 1.f : -1.f;
        }
    }
}

void winding_quadratic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                       uint32_t rule, float &w, float &cr) {  // curves.rs:127-166
    float ay = y0 - 2.0f * y1 + y2;
    float by = -2.0f * y0 + 2.0f * y1;
    float cy = y0 - py;
    float roots[2];
    int n = solve_quadratic(ay, by, cy, roots);
    for (int i = 0; i < n; ++i) {
        float t = roots[i];
        if (t >= 0.f && t <= 1.f) {
            float tt = 1.f - t;
            float x = (tt * tt) * x0 + (2.0f * tt * t) * x1 + (t * t) * x2;
            if (x > px) {
                if (rule == 1) cr = 1.f - cr;
                else {
                    float dy = 2.0f * ay * t + by;
                    w += (dy > 0.f) ? 1.f : -1.f;
                }
            }
        }
    }
}

void winding_cubic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                   float y3, uint32_t rule, float &w, float &cr) {  // curves.rs:170-214
    float a = -y0 + 3.0f * y1 - 3.0f * y2 + y3;
    float b = 3.0f * y0 - 6.0f * y1 + 3.0f * y2;
    float c = -3.0f * y0 + 3.0f * y1;
    float d = y0 - py;
    float roots[3];
    int n = solve_cubic(a, b, c, d, roots);
    for (int i = 0; i < n; ++i) {
        float t = roots[i];
        if (t >= 0.f && t <= 1.f) {
            float tt = 1.f - t;
            float x = (tt * tt * tt) * x0 + (3.0f * tt * tt * t) * x1 + (3.0f * tt * t * t) * x2 + (t * t * t) * x3;
            if (x > px) {
                if (rule == 1) cr = 1.f - cr;
                else {
                    float dy = 3.0f * a * t * t + 2.0f * b * t + c;
                    w += (dy > 0.f) ? 1.f : -1.f;
                }
            }
        }
    }
}

static inline float det2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

static void quadratic_closest_pt_approx(float px, float py, float x0, float y0, float x1, float y1, float x2,
                                        float y2, float &ox, float &oy, float &ot) {  // curves.rs:792-857
    float b0x = x0 - px, b0y = y0 - py, b1x = x1 - px, b1y = y1 - py, b2x = x2 - px, b2y = y2 - py;
    float a = det2(b0x, b0y, b2x, b2y);
    float b = 2.0f * det2(b1x, b1y, b0x, b0y);
    float d = 2.0f * det2(b2x, b2y, b1x, b1y);
    float f = b * d - a * a;
    float d21x = b2x - b1x, d21y = b2y - b1y, d10x = b1x - b0x, d10y = b1y - b0y, d20x = b2x - b0x,
          d20y = b2y - b0y;
    float gfx = 2.0f * (b * d21x + d * d10x + a * d20x);
    float gfy = 2.0f * (b * d21y + d * d10y + a * d20y);
    float tmp = gfx;
    gfx = gfy;
    gfy = -tmp;
    float gf_dot = gfx * gfx + gfy * gfy;
    float t = 0.f;
    if (gf_dot > 1.0e-8f) {
        float ppx = -f * gfx / gf_dot, ppy = -f * gfy / gf_dot;
        float d0px = b0x - ppx, d0py = b0y - ppy;
        float ap = det2(d0px, d0py, d20x, d20y);
        float bp = 2.0f * det2(d10x, d10y, d0px, d0py);
        float denom = 2.0f * a + b + d;
        if (abs_f(denom) > 1.0e-8f) t = clamp01((ap + bp) / denom);
        else t = clamp01((ap + bp) * 0.5f);
    }
    float tt = 1.f - t;
    ox = (tt * tt) * x0 + (2.0f * tt * t) * x1 + (t * t) * x2;
    oy = (tt * tt) * y0 + (2.0f * tt * t) * y1 + (t * t) * y2;
    ot = t;
}

void closest_quadratic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                       bool approx, float &dist, float &tout) {  // curves.rs:218-279
    float best_t = 0.f;
    float best = dist_pt(px, py, x0, y0);
    if (approx) {
        float ax, ay, at;
        quadratic_closest_pt_approx(px, py, x0, y0, x1, y1, x2, y2, ax, ay, at);
        best = dist_pt(px, py, ax, ay);
        best_t = at;
    } else {
        float de = dist_pt(px, py, x2, y2);
        if (de < best) { best = de; best_t = 1.f; }
        float ax = x0 - 2.0f * x1 + x2, ay = y0 - 2.0f * y1 + y2;
        float bx = -x0 + x1, by = -y0 + y1;
        float cx = x0 - px, cy = y0 - py;
        float a = ax * ax + ay * ay;
        float b = 3.0f * (ax * bx + ay * by);
        float c = 2.0f * (bx * bx + by * by) + (ax * cx + ay * cy);
        float d = bx * cx + by * cy;
        float roots[3];
        int n = solve_cubic(a, b, c, d, roots);
        for (int i = 0; i < n; ++i) {
            float t = roots[i];
            if (t >= 0.f && t <= 1.f) {
                float tt = 1.f - t;
                float x = (tt * tt) * x0 + (2.0f * tt * t) * x1 + (t * t) * x2;
                float y = (tt * tt) * y0 + (2.0f * tt * t) * y1 + (t * t) * y2;
                float dd = dist_pt(px, py, x, y);
                if (dd < best) { best = dd; best_t = t; }
            }
        }
    }
    dist = best;
    tout = best_t;
}

static inline float eval_quintic(float t, float b, float c, float d, float e, float f) {  // curves.rs:777
    return ((((t + b) * t + c) * t + d) * t + e) * t + f;
}
static inline float eval_quintic_deriv(float t, float b, float c, float d, float e) {  // curves.rs:783
    return (((5.0f * t + 4.0f * b) * t + 3.0f * c) * t + 2.0f * d) * t + e;
}

static void cubic_approx(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                         float y3, float &dist, float &tout) {  // curves.rs:507-556
    const uint32_t steps = 8;
    float inv = 1.0f / (float)steps;
    float prev_x = x0, prev_y = y0, prev_t = 0.f;
    float best = dist_pt(px, py, x0, y0), best_t = prev_t;
    for (uint32_t i = 1; i <= steps; ++i) {
        float t = (float)i * inv;
        float tt = 1.0f - t, tt2 = tt * tt, t2 = t * t;
        float a = tt2 * tt, b = 3.0f * tt2 * t, c = 3.0f * tt * t2, d = t2 * t;
        float cx = a * x0 + b * x1 + c * x2 + d * x3;
        float cy = a * y0 + b * y1 + c * y2 + d * y3;
        float sd, st;
        distance_to_segment_with_t(px, py, prev_x, prev_y, cx, cy, sd, st);
        float lt = prev_t + st * (t - prev_t);
        if (sd < best) { best = sd; best_t = lt; }
        prev_x = cx; prev_y = cy; prev_t = t;
    }
    dist = best;
    tout = best_t;
}

// Shared body of distance_to_cubic (curves.rs:341-505) and
// closest_point_cubic_with_t (curves.rs:561-731): identical arithmetic.
void closest_cubic(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                   float y3, bool approx, float &dist, float &tout) {
    if (approx) {
        cubic_approx(px, py, x0, y0, x1, y1, x2, y2, x3, y3, dist, tout);
        return;
    }
    float best_t = 0.f;
    float best = dist_pt(px, py, x0, y0);
    float de = dist_pt(px, py, x3, y3);
    if (de < best) { best = de; best_t = 1.f; }
    float ax = -x0 + 3.0f * x1 - 3.0f * x2 + x3;
    float ay = -y0 + 3.0f * y1 - 3.0f * y2 + y3;
    float bx = 3.0f * x0 - 6.0f * x1 + 3.0f * x2;
    float by = 3.0f * y0 - 6.0f * y1 + 3.0f * y2;
    float cx = -3.0f * x0 + 3.0f * x1;
    float cy = -3.0f * y0 + 3.0f * y1;
    float dx = x0 - px, dy = y0 - py;
    float a = 3.0f * (ax * ax + ay * ay);
    if (abs_f(a) > 1.0e-8f) {
        float b = 5.0f * (ax * bx + ay * by);
        float c = 4.0f * (ax * cx + ay * cy) + 2.0f * (bx * bx + by * by);
        float d = 3.0f * ((bx * cx + by * cy) + (ax * dx + ay * dy));
        float e = (cx * cx + cy * cy) + 2.0f * (dx * bx + dy * by);
        float f = dx * cx + dy * cy;
        b = b / a; c = c / a; d = d / a; e = e / a; f = f / a;
        float p1a = (2.0f / 5.0f) * c - (4.0f / 25.0f) * b * b;
        float p1b = (3.0f / 5.0f) * d - (3.0f / 25.0f) * b * c;
        float p1c = (4.0f / 5.0f) * e - (2.0f / 25.0f) * b * d;
        float p1d = f - b * e / 25.0f;
        float q_root = -b / 5.0f;
        float p_roots[3];
        int num_p = solve_cubic(p1a, p1b, p1c, p1d, p_roots);
        float iv[4];
        int ni = 0;
        if (q_root >= 0.f && q_root <= 1.f) iv[ni++] = q_root;
        for (int i = 0; i < num_p; ++i) iv[ni++] = p_roots[i];
        for (int j = 1; j < ni; ++j) {  // insertion sort, curves.rs:630-645
            int k = j;
            while (k > 0) {
                float u = iv[k - 1], v = iv[k];
                if (u <= v) break;
                iv[k - 1] = v; iv[k] = u;
                --k;
            }
        }
        float lower = 0.f;
        int idx = 0;
        while (idx <= ni) {
            if (idx < ni && iv[idx] < 0.f) { ++idx; continue; }
            float upper = (idx < ni) ? min_f(iv[idx], 1.f) : 1.f;
            float lb = lower, ub = upper;
            float lbe = eval_quintic(lb, b, c, d, e, f);
            float ube = eval_quintic(ub, b, c, d, e, f);
            if (lbe * ube > 0.f) {
                lower = upper;
                ++idx;
                continue;
            }
            if (lbe > ube) {
                float tmp = lb; lb = ub; ub = tmp;
                tmp = lbe; lbe = ube; ube = tmp;
            }
            float t = (lb + ub) * 0.5f;
            for (uint32_t it = 0; it < 20; ++it) {
                if (t < lb || t > ub) t = (lb + ub) * 0.5f;
                float value = eval_quintic(t, b, c, d, e, f);
                if (abs_f(value) < 1.0e-5f || it == 19) break;
                if (value > 0.f) ub = t; else lb = t;
                float der = eval_quintic_deriv(t, b, c, d, e);
                t = t - value / der;
            }
            if (t >= 0.f && t <= 1.f) {
                float tt = 1.f - t;
                float x = (tt * tt * tt) * x0 + (3.0f * tt * tt * t) * x1 + (3.0f * tt * t * t) * x2 + (t * t * t) * x3;
                float y = (tt * tt * tt) * y0 + (3.0f * tt * tt * t) * y1 + (3.0f * tt * t * t) * y2 + (t * t * t) * y3;
                float dd = dist_pt(px, py, x, y);
                if (dd < best) { best = dd; best_t = t; }
            }
            if (upper >= 1.f) break;
            lower = upper;
            ++idx;
        }
    }
    dist = best;
    tout = best_t;
}

// ---- sampling.rs -----------------------------------------------------------
V4 paint_color(const Paint &p, float px, float py) {  // sampling.rs:11-43, 47-139
    V4 out;
    if (p.kind == DVG_PAINT_SOLID) return p.color;
    if (p.kind != DVG_PAINT_LINEAR && p.kind != DVG_PAINT_RADIAL) return out;
    float p0 = p.p0.x, p1 = p.p0.y, p2 = p.p1.x, p3 = p.p1.y;
    float t;
    if (!p.radial) {
        float vx = p2 - p0, vy = p3 - p1;
        float denom = max_f(vx * vx + vy * vy, 1.0e-3f);
        t = ((px - p0) * vx + (py - p1) * vy) / denom;
    } else {
        float dx = (px - p0) / max_f(p2, 1.0e-3f);
        float dy = (py - p1) / max_f(p3, 1.0e-3f);
        t = sqrt_f(dx * dx + dy * dy);
    }
    size_t n = p.stops.size();
    if (n == 0) return out;
    float first = p.stops[0].offset, last = p.stops[n - 1].offset;
    if (t <= first) return p.stops[0].color;
    if (t >= last) return p.stops[n - 1].color;
    for (size_t i = 0; i + 1 < n; ++i) {
        float curr = p.stops[i].offset, next = p.stops[i + 1].offset;
        if (t >= curr && t < next) {
            float tt = (t - curr) / max_f(next - curr, 1.0e-5f);
            const V4 &c0 = p.stops[i].color, &c1 = p.stops[i + 1].color;
            return V4(c0.x * (1.f - tt) + c1.x * tt, c0.y * (1.f - tt) + c1.y * tt, c0.z * (1.f - tt) + c1.z * tt,
                      c0.w * (1.f - tt) + c1.w * tt);
        }
    }
    return out;
}

float filter_weight(uint32_t type, float dx, float dy, float radius) {  // sampling.rs:143-181
    float adx = abs_f(dx), ady = abs_f(dy);
    if (radius <= 0.f) return 1.f;
    if (adx > radius || ady > radius) return 0.f;
    if (type == DVG_FILTER_BOX) {
        float denom = (2.0f * radius) * (2.0f * radius);
        return 1.0f / denom;
    }
    if (type == DVG_FILTER_TENT) {
        float fx = radius - adx, fy = radius - ady;
        float norm = 1.0f / (radius * radius);
        return fx * fy * norm * norm;
    }
    if (type == DVG_FILTER_RADIAL_PARABOLIC) {
        float rx = 1.0f - (dx / radius) * (dx / radius);
        float ry = 1.0f - (dy / radius) * (dy / radius);
        return (4.0f / 3.0f) * rx * (4.0f / 3.0f) * ry;
    }
    if (type == DVG_FILTER_HANN) {
        const float pi = 3.14159265f;
        float ndx = (dx / (2.0f * radius)) + 0.5f;
        float ndy = (dy / (2.0f * radius)) + 0.5f;
        float fx = 0.5f * (1.0f - cos_f(2.0f * pi * ndx));
        float fy = 0.5f * (1.0f - cos_f(2.0f * pi * ndy));
        float norm = 1.0f / (radius * radius);
        return fx * fy * norm;
    }
    return 1.f;
}

// ---- rng.rs ----------------------------------------------------------------
void jitter_xy(uint32_t x, uint32_t y, uint32_t sx, uint32_t sy, uint32_t width, uint32_t nsx, uint32_t nsy,
               uint32_t seed, bool jitter, float &rx, float &ry) {  // rng.rs:167-197
    rx = 0.5f;
    ry = 0.5f;
    if (jitter) {
        uint32_t idx = ((y * width + x) * nsy + sy) * nsx + sx;  // u32 wrap as on the GPU
        Pcg32 rng((uint64_t)idx, (uint64_t)seed);
        rx = rng.next_f32();
        ry = rng.next_f32();
    }
}

}  // namespace gpu

// ---------------------------------------------------------------------------
// Prepared per-(group, shape) data (prepare.rs:114-357).
struct PreparedInstance {
    uint32_t shape;        // scene index
    bool bounds_valid;
    float b[4];            // group-local bounds, padded
    float inv[6];          // inverse shape transform, rows 0-1
    std::vector<Seg> segs; // curve segments in shape space
};
// A node of a group BVH (renderer/bvh.rs:44-59); items are positions in
// PreparedGroup::shapes.
struct BvhNode {
    float b[4];
    uint32_t left, right, skip, start, count;
};
struct PreparedGroup {
    int tile[4] = {1, 1, 0, 0};  // 16-px tile range of the canvas AABB (count mode)
    float xform[6];        // canvas_to_shape rows 0-1
    float inv_scale;
    std::vector<PreparedInstance> shapes;
    std::vector<BvhNode> bvh;       // empty: no shape with bounds
    std::vector<uint32_t> bvh_items;
};

static const uint32_t BVH_NONE = 0xffffffffu;

// build_bvh_node, renderer/bvh.rs:149-190: median split by box centre on the
// longer axis (a stable sort, as sort_by), leaves of at most 8 items
// (renderer/constants.rs:23).
static uint32_t bvh_build(std::vector<BvhNode> &nodes, const std::vector<const float *> &boxes,
                          std::vector<uint32_t> &idx, size_t start, size_t end) {
    BvhNode n;
    n.b[0] = INFINITY; n.b[1] = INFINITY; n.b[2] = -INFINITY; n.b[3] = -INFINITY;
    for (size_t i = start; i < end; ++i) {
        const float *o = boxes[idx[i]];
        n.b[0] = rmin(n.b[0], o[0]); n.b[1] = rmin(n.b[1], o[1]);
        n.b[2] = rmax(n.b[2], o[2]); n.b[3] = rmax(n.b[3], o[3]);
    }
    n.left = n.right = n.skip = BVH_NONE;
    n.start = (uint32_t)start;
    n.count = (uint32_t)(end - start);
    uint32_t id = (uint32_t)nodes.size();
    nodes.push_back(n);
    if (end - start <= 8) return id;
    int axis = (n.b[2] - n.b[0]) >= (n.b[3] - n.b[1]) ? 0 : 1;
    std::stable_sort(idx.begin() + start, idx.begin() + end, [&](uint32_t a, uint32_t b) {
        return (boxes[a][axis] + boxes[a][axis + 2]) * 0.5f < (boxes[b][axis] + boxes[b][axis + 2]) * 0.5f;
    });
    size_t mid = start + (end - start) / 2;
    uint32_t l = bvh_build(nodes, boxes, idx, start, mid);
    uint32_t r = bvh_build(nodes, boxes, idx, mid, end);
    nodes[id].left = l;
    nodes[id].right = r;
    nodes[id].count = 0;
    return id;
}

// append_group_bvh, renderer/bvh.rs:62-109, over the shapes with bounds, with
// the skip pointers of assign_bvh_skip (:192-215).
static void group_bvh(PreparedGroup &pg) {
    std::vector<const float *> boxes;
    std::vector<uint32_t> ids;
    for (size_t i = 0; i < pg.shapes.size(); ++i)
        if (pg.shapes[i].bounds_valid) {
            boxes.push_back(pg.shapes[i].b);
            ids.push_back((uint32_t)i);
        }
    if (boxes.empty()) return;
    std::vector<uint32_t> idx(boxes.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = (uint32_t)i;
    uint32_t root = bvh_build(pg.bvh, boxes, idx, 0, idx.size());
    std::vector<std::pair<uint32_t, uint32_t>> st = {{root, BVH_NONE}};
    while (!st.empty()) {
        auto [id, sk] = st.back();
        st.pop_back();
        if (id == BVH_NONE) continue;
        BvhNode &n = pg.bvh[id];
        n.skip = sk;
        if (n.count == 0) {
            if (n.right != BVH_NONE) st.push_back({n.right, sk});
            if (n.left != BVH_NONE) st.push_back({n.left, n.right});
        }
    }
    for (uint32_t i : idx) pg.bvh_items.push_back(ids[i]);
}

static void prepare(const Scene &sc, const dvg_render_options &opt, std::vector<PreparedGroup> &out) {
    out.clear();
    for (const Group &g : sc.groups) {
        PreparedGroup pg;
        float group_scale = rmax(g.shape_to_canvas.average_scale(), 1.0e-6f);
        float inv_scale = 1.0f / group_scale;
        float base_tol = rmax(opt.path_tolerance, 0.01f);
        float aa_local = opt.use_prefiltering ? inv_scale : 0.0f;
        const M3 &x = g.canvas_to_shape;
        float xf[6] = {x.m[0][0], x.m[0][1], x.m[0][2], x.m[1][0], x.m[1][1], x.m[1][2]};
        std::memcpy(pg.xform, xf, sizeof(xf));
        pg.inv_scale = inv_scale;
        for (uint32_t si : g.shape_indices) {
            const Shape &sh = sc.shapes[si];
            PreparedInstance pi;
            pi.shape = si;
            float total_scale = rmax(group_scale * sh.transform.average_scale(), 1.0e-6f);
            float tol = rmax(base_tol / total_scale, 1.0e-3f);
            float max_r = std::fabs(sh.stroke_width);
            bool has_bounds = true;
            V2 bmin, bmax;
            if (sh.kind == DVG_SHAPE_CIRCLE) {
                float r = std::fabs(sh.radius_c);
                bmin = V2(sh.center.x - r, sh.center.y - r);
                bmax = V2(sh.center.x + r, sh.center.y + r);
            } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
                float rx = std::fabs(sh.radius_e.x), ry = std::fabs(sh.radius_e.y);
                bmin = V2(sh.center.x - rx, sh.center.y - ry);
                bmax = V2(sh.center.x + rx, sh.center.y + ry);
            } else if (sh.kind == DVG_SHAPE_RECT) {
                bmin = sh.rmin;
                bmax = sh.rmax;
            } else {
                max_r = rmax(max_r, flattened_max_radius(sh.path, tol, sh.stroke_width));
                pi.segs = path_segments(sh.path, sh.stroke_width);
                if (sh.path.points.empty()) {
                    has_bounds = false;  // flattened bounds of an empty path are None
                } else {
                    bmin = V2(INFINITY, INFINITY);
                    bmax = V2(-INFINITY, -INFINITY);
                    for (const V2 &p : sh.path.points) {
                        bmin = bmin.vmin(p);
                        bmax = bmax.vmax(p);
                    }
                }
            }
            M3 inv = sh.transform.inverse_or_identity();
            float iv[6] = {inv.m[0][0], inv.m[0][1], inv.m[0][2], inv.m[1][0], inv.m[1][1], inv.m[1][2]};
            std::memcpy(pi.inv, iv, sizeof(iv));
            float stroke_pad = std::fabs(max_r) * rmax(sh.transform.max_scale(), 0.0f);
            if (sh.join == DVG_JOIN_MITER) stroke_pad *= rmax(sh.miter_limit, 1.0f);
            float pad = rmax(aa_local + stroke_pad, 0.0f);
            pi.bounds_valid = has_bounds;
            if (has_bounds) {
                V2 c[4] = {V2(bmin.x, bmin.y), V2(bmax.x, bmin.y), V2(bmax.x, bmax.y), V2(bmin.x, bmax.y)};
                V2 mn(INFINITY, INFINITY), mx(-INFINITY, -INFINITY);
                for (int k = 0; k < 4; ++k) {
                    V2 t = sh.transform.transform_point(c[k]);
                    mn = mn.vmin(t);
                    mx = mx.vmax(t);
                }
                pi.b[0] = mn.x - pad; pi.b[1] = mn.y - pad;
                pi.b[2] = mx.x + pad; pi.b[3] = mx.y + pad;
            } else {
                pi.b[0] = 1.f; pi.b[1] = 1.f; pi.b[2] = 0.f; pi.b[3] = 0.f;
            }
            pg.shapes.push_back(pi);
        }
        group_bvh(pg);
        // canvas AABB -> tile range, as bin_tiles_count (binning.rs:30-110)
        {
            float u[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
            bool any = false;
            for (const PreparedInstance &pi : pg.shapes)
                if (pi.bounds_valid) {
                    u[0] = rmin(u[0], pi.b[0]); u[1] = rmin(u[1], pi.b[1]);
                    u[2] = rmax(u[2], pi.b[2]); u[3] = rmax(u[3], pi.b[3]);
                    any = true;
                }
            if (any) {
                const M3 &m = g.shape_to_canvas;
                float xs[4], ys[4];
                float cx[4] = {u[0], u[2], u[2], u[0]}, cy[4] = {u[1], u[1], u[3], u[3]};
                for (int k = 0; k < 4; ++k) {
                    xs[k] = m.m[0][0] * cx[k] + m.m[0][1] * cy[k] + m.m[0][2];
                    ys[k] = m.m[1][0] * cx[k] + m.m[1][1] * cy[k] + m.m[1][2];
                }
                float mnx = std::min({xs[0], xs[1], xs[2], xs[3]}), mxx = std::max({xs[0], xs[1], xs[2], xs[3]});
                float mny = std::min({ys[0], ys[1], ys[2], ys[3]}), mxy = std::max({ys[0], ys[1], ys[2], ys[3]});
                pg.tile[0] = (int)std::floor(mnx / 16.f); pg.tile[1] = (int)std::floor(mny / 16.f);
                pg.tile[2] = (int)std::floor(mxx / 16.f); pg.tile[3] = (int)std::floor(mxy / 16.f);
            }
        }
        out.push_back(pg);
    }
}

using namespace gpu;

struct GroupAccum {
    float fill_min = 1.0e20f, winding = 0.f, crossings = 0.f;
    float stroke_min = 1.0e20f, stroke_radius = 0.f, stroke_hit = 0.f;
};

// accumulate_shape_fill (forward.rs:409-560) for one prepared instance.
static void shape_fill(const Shape &sh, const PreparedInstance &pi, uint32_t rule, bool prefilter, float px,
                       float py, GroupAccum &acc) {
    if (sh.kind == DVG_SHAPE_CIRCLE || sh.kind == DVG_SHAPE_ELLIPSE || sh.kind == DVG_SHAPE_RECT) {
        float dist;
        if (sh.kind == DVG_SHAPE_CIRCLE) {
            float dx = px - sh.center.x, dy = py - sh.center.y;
            dist = sqrt_f(dx * dx + dy * dy) - std::fabs(sh.radius_c);
        } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
            float p2 = std::fabs(sh.radius_e.x), p3 = std::fabs(sh.radius_e.y);
            float dx = (px - sh.center.x) / max_f(abs_f(p2), 1.0e-3f);
            float dy = (py - sh.center.y) / max_f(abs_f(p3), 1.0e-3f);
            float len = sqrt_f(dx * dx + dy * dy);
            dist = (len - 1.0f) * min_f(abs_f(p2), abs_f(p3));
        } else {
            dist = rect_signed_distance(px, py, sh.rmin.x, sh.rmin.y, sh.rmax.x, sh.rmax.y);
        }
        float ad = abs_f(dist);
        if (ad < acc.fill_min) acc.fill_min = ad;
        if (dist <= 0.f) {
            if (rule == 1) acc.crossings = 1.f - acc.crossings;
            else acc.winding += 1.f;
        }
        return;
    }
    // path
    float local_min = 1.0e20f, lw = 0.f, lc = 0.f;
    bool approx = sh.path.use_distance_approx;
    for (const Seg &s : pi.segs) {
        if (prefilter) {
            float d, t;
            if (s.kind == 0) d = distance_to_segment(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y);
            else if (s.kind == 1) closest_quadratic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, approx, d, t);
            else closest_cubic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, s.p3.x, s.p3.y, approx, d, t);
            if (d < local_min) local_min = d;
        }
        if (s.kind == 0) winding_line(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, rule, lw, lc);
        else if (s.kind == 1) winding_quadratic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, rule, lw, lc);
        else winding_cubic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, s.p3.x, s.p3.y, rule, lw, lc);
    }
    if (local_min < acc.fill_min) acc.fill_min = local_min;
    if (rule == 1) {
        if (lc > 0.f) acc.crossings = 1.f - acc.crossings;
    } else {
        acc.winding += lw;
    }
}

static inline float seg_radius(const Seg &s, float t) {  // forward.rs:1092-1103
    if (s.kind == 0) return s.r0 + t * (s.r1 - s.r0);
    float tt = 1.f - t;
    if (s.kind == 1) return tt * tt * s.r0 + 2.0f * tt * t * s.r1 + t * t * s.r2;
    return tt * tt * tt * s.r0 + 3.0f * tt * tt * t * s.r1 + 3.0f * tt * t * t * s.r2 + t * t * t * s.r3;
}

// accumulate_shape_stroke (forward.rs:565-709) for one prepared instance.
static void shape_stroke(const Shape &sh, const PreparedInstance &pi, bool prefilter, float px, float py,
                         GroupAccum &acc) {
    float sw = sh.stroke_width;
    uint32_t curve_count = (uint32_t)pi.segs.size();
    if (sw <= 0.f && curve_count == 0) return;
    if (sh.kind != DVG_SHAPE_PATH) {
        float dist;
        if (sh.kind == DVG_SHAPE_CIRCLE) {
            float dx = px - sh.center.x, dy = py - sh.center.y;
            dist = abs_f(sqrt_f(dx * dx + dy * dy) - std::fabs(sh.radius_c));
        } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
            float p2 = std::fabs(sh.radius_e.x), p3 = std::fabs(sh.radius_e.y);
            float dx = (px - sh.center.x) / max_f(abs_f(p2), 1.0e-3f);
            float dy = (py - sh.center.y) / max_f(abs_f(p3), 1.0e-3f);
            float len = sqrt_f(dx * dx + dy * dy);
            dist = abs_f((len - 1.0f) * min_f(abs_f(p2), abs_f(p3)));
        } else {
            dist = abs_f(rect_signed_distance(px, py, sh.rmin.x, sh.rmin.y, sh.rmax.x, sh.rmax.y));
        }
        if (prefilter) {
            if (dist < acc.stroke_min) { acc.stroke_min = dist; acc.stroke_radius = sw; }
        } else if (dist < sw) {
            acc.stroke_hit = 1.f;
        }
        return;
    }
    float local_min = 1.0e20f, local_r = 0.f, local_hit = 0.f;
    bool approx = sh.path.use_distance_approx;
    for (const Seg &s : pi.segs) {
        float d, t;
        if (s.kind == 0) distance_to_segment_with_t(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, d, t);
        else if (s.kind == 1) closest_quadratic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, approx, d, t);
        else closest_cubic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, s.p3.x, s.p3.y, approx, d, t);
        float r = seg_radius(s, t);
        if (prefilter) {
            if (d < local_min) { local_min = d; local_r = r; }
        } else if (d < r) {
            local_hit = 1.f;
            break;
        }
    }
    if (prefilter) {
        if (local_min < acc.stroke_min) { acc.stroke_min = local_min; acc.stroke_radius = local_r; }
    } else if (local_hit > 0.f) {
        acc.stroke_hit = 1.f;
    }
}

// blend_group, forward.rs:339-404.
static void blend_group(float out[4], const Paint &fill, const Paint &stroke, uint32_t rule, V4 fc, V4 scol,
                        float fill_min, float winding, float crossings, float stroke_min, float stroke_r,
                        float stroke_hit, bool prefilter) {
    const float big = 1.0e20f;
    if (stroke.kind != DVG_PAINT_NONE) {
        float cov = 0.f;
        if (prefilter) {
            if (stroke_min < big && stroke_r > 0.f)
                cov = smoothstep_unit(stroke_min + stroke_r) - smoothstep_unit(stroke_min - stroke_r);
        } else if (stroke_hit > 0.f) {
            cov = 1.f;
        }
        if (cov > 0.f) {
            float a = cov * scol.w, inv = 1.f - a;
            out[0] = scol.x * a + out[0] * inv;
            out[1] = scol.y * a + out[1] * inv;
            out[2] = scol.z * a + out[2] * inv;
            out[3] = a + out[3] * inv;
        }
    }
    if (fill.kind != DVG_PAINT_NONE) {
        bool inside = (rule == 1) ? crossings > 0.f : winding != 0.f;
        float cov = 0.f;
        if (prefilter) {
            if (fill_min < big) {
                float sd = inside ? fill_min : -fill_min;
                cov = smoothstep_unit(sd);
            }
        } else if (inside) {
            cov = 1.f;
        }
        if (cov > 0.f) {
            float a = cov * fc.w, inv = 1.f - a;
            out[0] = fc.x * a + out[0] * inv;
            out[1] = fc.y * a + out[1] * inv;
            out[2] = fc.z * a + out[2] * inv;
            out[3] = a + out[3] * inv;
        }
    }
}

// eval_scene_tiled (forward.rs:1133-1290) at canvas point (px, py); returns
// premultiplied RGBA starting from the premultiplied background.
static void eval_scene(const Scene &sc, const std::vector<PreparedGroup> &pgs, bool prefilter, float px, float py,
                       const float bg[4], float out[4]) {
    out[0] = bg[0] * bg[3];
    out[1] = bg[1] * bg[3];
    out[2] = bg[2] * bg[3];
    out[3] = bg[3];
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    [[maybe_unused]] int tx = (int)std::floor(px / 16.f), ty = (int)std::floor(py / 16.f);
    for (size_t gi = 0; gi < sc.groups.size(); ++gi) {
        const Group &g = sc.groups[gi];
        const PreparedGroup &pg = pgs[gi];
        // counting: only the groups of this tile's list (the reference's
        // culling; the loop over every group is the restatement's)
        COUNT_ON(counting && tx >= pg.tile[0] && tx <= pg.tile[2] && ty >= pg.tile[1] && ty <= pg.tile[3]);
        const float *x = pg.xform;
        float lx = x[0] * px + x[1] * py + x[2];
        float ly = x[3] * px + x[4] * py + x[5];
        float gscale = pg.inv_scale > 0.f ? 1.f / pg.inv_scale : 1.f;
        V4 fc = paint_color(g.fill, px, py);
        V4 scol = paint_color(g.stroke, px, py);
        GroupAccum acc;
        if (g.fill.kind != DVG_PAINT_NONE || g.stroke.kind != DVG_PAINT_NONE) {
            for (const PreparedInstance &pi : pg.shapes) {
                // accumulate_shape_in_group, forward.rs:1420-1503
                bool inb = pi.b[0] <= pi.b[2] && pi.b[1] <= pi.b[3];
                if (inb && (lx < pi.b[0] || lx > pi.b[2] || ly < pi.b[1] || ly > pi.b[3])) inb = false;
                if (!inb) continue;
                const float *iv = pi.inv;
                float spx = iv[0] * lx + iv[1] * ly + iv[2];
                float spy = iv[3] * lx + iv[4] * ly + iv[5];
                const Shape &sh = sc.shapes[pi.shape];
                if (g.fill.kind != DVG_PAINT_NONE) shape_fill(sh, pi, g.fill_rule, prefilter, spx, spy, acc);
                if (g.stroke.kind != DVG_PAINT_NONE) shape_stroke(sh, pi, prefilter, spx, spy, acc);
            }
        }
        blend_group(out, g.fill, g.stroke, g.fill_rule, fc, scol, acc.fill_min * gscale, acc.winding, acc.crossings,
                    acc.stroke_min * gscale, acc.stroke_radius * gscale, acc.stroke_hit, prefilter);
    }
    COUNT_ON(counting);
}

// The 16x16 tile range of every group's canvas box (4 ints a group,
// min > max when empty): the groups of each tile's list (binning.rs:30-110).
std::vector<int> group_tile_ranges(const Scene &sc, const dvg_render_options &opt) {
    std::vector<PreparedGroup> pgs;
    prepare(sc, opt, pgs);
    std::vector<int> out;
    for (const PreparedGroup &pg : pgs) out.insert(out.end(), pg.tile, pg.tile + 4);
    return out;
}

// Scene evaluation (no splat) of the samples [i0, i1), for CPU timing.
void forward_samples(const Scene &sc, const dvg_render_options &opt, size_t i0, size_t i1) {
    std::vector<PreparedGroup> pgs;
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    COUNT_ON(false);  // scene preparation is per step, not per sample
    prepare(sc, opt, pgs);
    uint32_t W = sc.width, H = sc.height;
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    size_t total = (size_t)W * H * nsx * nsy;
    if (i1 > total) i1 = total;
    bool prefilter = opt.use_prefiltering != 0;
    bool jitter = !prefilter && opt.jitter != 0;
    float inv_sx = 1.0f / (float)nsx, inv_sy = 1.0f / (float)nsy;
    float sink = 0.f;
    COUNT_ON(counting);
    for (size_t idx = i0; idx < i1; ++idx) {
        uint32_t sx = idx % nsx, sy = (idx / nsx) % nsy, x = (idx / (nsx * nsy)) % W, y = idx / ((size_t)nsx * nsy * W);
        float rx, ry;
        jitter_xy(x, y, sx, sy, W, nsx, nsy, opt.seed, jitter, rx, ry);
        float px = (float)x + ((float)sx + rx) * inv_sx, py = (float)y + ((float)sy + ry) * inv_sy;
        float bg[4] = {sc.background.x, sc.background.y, sc.background.z, sc.background.w};
        float col[4];
        eval_scene(sc, pgs, prefilter, px, py, bg, col);
        sink = sink + col[0];
    }
    g_forward_sink = sink;  // keeps the evaluations observable
}

int render_forward(const Scene &sc, const dvg_render_options &opt, float *out_img, float *sample_colors,
                   const char **err, uint32_t y0, uint32_t y1) {
    uint32_t W = sc.width, H = sc.height;
    y1 = std::min(y1, H);
    if (y0 >= y1) return DVG_OK;
    if (sample_colors && (y0 != 0 || y1 != H)) {
        *err = "sample colours need the whole image";
        return DVG_INVALID_SCENE;
    }
    if (sc.has_bg_image && sc.bg_image.size() != (size_t)W * H * 4) {
        *err = "background image size mismatch";
        return DVG_INVALID_SCENE;
    }
    if (W == 0 || H == 0) return DVG_OK;
    if (sc.groups.empty()) {
        for (size_t i = (size_t)y0 * W; i < (size_t)y1 * W; ++i) {
            float *o = out_img + (i - (size_t)y0 * W) * 4;
            if (sc.has_bg_image) {
                for (int c = 0; c < 4; ++c) o[c] = sc.bg_image[i * 4 + c];
            } else {
                o[0] = sc.background.x;
                o[1] = sc.background.y;
                o[2] = sc.background.z;
                o[3] = sc.background.w;
            }
        }
        return DVG_OK;
    }
    std::vector<PreparedGroup> pgs;
    prepare(sc, opt, pgs);
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    bool prefilter = opt.use_prefiltering != 0;
    bool jitter = !prefilter && opt.jitter != 0;
    int ri = (int)rmax(std::ceil(sc.filter_radius), 0.0f);
    float inv_sx = 1.0f / (float)nsx, inv_sy = 1.0f / (float)nsy;
    size_t npix = (size_t)W * H;
    std::vector<float> weight(npix, 0.f), color(npix * 4, 0.f);
    // sample rows whose filter footprint reaches rows [y0, y1)
    uint32_t ys = y0 > (uint32_t)ri ? y0 - (uint32_t)ri : 0u;
    uint32_t ye = std::min(H, y1 + (uint32_t)ri);
    // rasterize_weights_f32
    for (uint32_t y = ys; y < ye; ++y)
        for (uint32_t x = 0; x < W; ++x)
            for (uint32_t sy = 0; sy < nsy; ++sy)
                for (uint32_t sx = 0; sx < nsx; ++sx) {
                    float rx, ry;
                    jitter_xy(x, y, sx, sy, W, nsx, nsy, opt.seed, jitter, rx, ry);
                    float px = (float)x + ((float)sx + rx) * inv_sx;
                    float py = (float)y + ((float)sy + ry) * inv_sy;
                    for (int dy = -ri; dy <= ri; ++dy)
                        for (int dx = -ri; dx <= ri; ++dx) {
                            int xx = (int)x + dx, yy = (int)y + dy;
                            if (xx < 0 || yy < 0 || xx >= (int)W || yy >= (int)H) continue;
                            float w = filter_weight(sc.filter_type, ((float)xx + 0.5f) - px, ((float)yy + 0.5f) - py,
                                                    sc.filter_radius);
                            if (w > 0.f) weight[(size_t)yy * W + xx] += w;
                        }
                }
    // rasterize_splat_f32
    uint32_t spp = nsx * nsy;
    for (uint32_t y = ys; y < ye; ++y)
        for (uint32_t x = 0; x < W; ++x)
            for (uint32_t sy = 0; sy < nsy; ++sy)
                for (uint32_t sx = 0; sx < nsx; ++sx) {
                    float rx, ry;
                    jitter_xy(x, y, sx, sy, W, nsx, nsy, opt.seed, jitter, rx, ry);
                    float px = (float)x + ((float)sx + rx) * inv_sx;
                    float py = (float)y + ((float)sy + ry) * inv_sy;
                    float bg[4] = {sc.background.x, sc.background.y, sc.background.z, sc.background.w};
                    if (sc.has_bg_image) {
                        size_t i4 = ((size_t)y * W + x) * 4;
                        for (int c = 0; c < 4; ++c) bg[c] = sc.bg_image[i4 + c];
                    }
                    float col[4];
                    eval_scene(sc, pgs, prefilter, px, py, bg, col);
                    if (sample_colors) {
                        size_t si = (((size_t)y * W + x) * spp + (size_t)sy * nsx + sx) * 4;
                        for (int c = 0; c < 4; ++c) sample_colors[si + c] = col[c];
                    }
                    for (int dy = -ri; dy <= ri; ++dy)
                        for (int dx = -ri; dx <= ri; ++dx) {
                            int xx = (int)x + dx, yy = (int)y + dy;
                            if (xx < 0 || yy < 0 || xx >= (int)W || yy >= (int)H) continue;
                            size_t base = (size_t)yy * W + xx;
                            float ws = weight[base];
                            if (ws > 0.f) {
                                float w = filter_weight(sc.filter_type, ((float)xx + 0.5f) - px,
                                                        ((float)yy + 0.5f) - py, sc.filter_radius);
                                if (w > 0.f) {
                                    float wn = w / ws;
                                    for (int c = 0; c < 4; ++c) color[base * 4 + c] += col[c] * wn;
                                }
                            }
                        }
                }
    // resolve_splat_f32 with the premultiplied background of renderer.rs:499-515
    float bgp[4] = {sc.background.x * sc.background.w, sc.background.y * sc.background.w,
                    sc.background.z * sc.background.w, sc.background.w};
    for (size_t i = (size_t)y0 * W; i < (size_t)y1 * W; ++i) {
        float r = bgp[0], g = bgp[1], b = bgp[2], a = bgp[3];
        if (weight[i] > 0.f) {
            r = color[i * 4]; g = color[i * 4 + 1]; b = color[i * 4 + 2]; a = color[i * 4 + 3];
        }
        if (a > 1.0e-8f) { r = r / a; g = g / a; b = b / a; }
        float *o = out_img + (i - (size_t)y0 * W) * 4;
        o[0] = r; o[1] = g; o[2] = b; o[3] = a;
    }
    return DVG_OK;
}


// ---------------------------------------------------------------------------
// Signed distance (Renderer::render_sdf / eval_positions, renderer.rs:586-848,
// over src/gpu/kernels/sdf.rs:9-262 and distance.rs:9-513), f32 as the
// reference GPU computes it.

// closest_point_ellipse, distance.rs:83-139
static void closest_point_ellipse(float cx, float cy, float rx_in, float ry_in, float px, float py, float &ox,
                                  float &oy) {
    float rx = abs_f(rx_in), ry = abs_f(ry_in);
    const float eps = 1.0e-6f;
    if (rx < eps && ry < eps) { ox = cx; oy = cy; return; }
    if (rx < eps) { ox = cx; oy = cy + std::min(std::max(py - cy, -ry), ry); return; }
    if (ry < eps) { ox = cx + std::min(std::max(px - cx, -rx), rx); oy = cy; return; }
    float dx = px - cx, dy = py - cy;
    if (abs_f(dx) < eps && abs_f(dy) < eps) { ox = cx + rx; oy = cy; return; }
    float sgx = dx < 0.f ? -1.f : 1.f, sgy = dy < 0.f ? -1.f : 1.f;
    dx = abs_f(dx);
    dy = abs_f(dy);
    float t = std::atan2(dy * rx, dx * ry);
    for (int i = 0; i < 20; ++i) {
        float s = std::sin(t), c = std::cos(t);
        float g = rx * dx * s - ry * dy * c + (ry * ry - rx * rx) * s * c;
        float gt = rx * dx * c + ry * dy * s + (ry * ry - rx * rx) * (c * c - s * s);
        if (abs_f(gt) < 1.0e-12f) break;
        float next = std::min<float>(std::max<float>(t - g / gt, 0.0f), 1.57079633f);
        if (abs_f(next - t) < 1.0e-6f) { t = next; break; }
        t = next;
    }
    float s = std::sin(t), c = std::cos(t);
    ox = cx + sgx * rx * c;
    oy = cy + sgy * ry * s;
}

// closest_point_shape, distance.rs:235-296 (shape space); false when a path
// has no segment.
static bool closest_point_shape(const Shape &sh, const PreparedInstance &pi, float px, float py, float &ox,
                                float &oy) {
    if (sh.kind == DVG_SHAPE_CIRCLE) {  // closest_point_circle, distance.rs:33-49
        float cx = sh.center.x, cy = sh.center.y, r = std::fabs(sh.radius_c);
        float dx = px - cx, dy = py - cy;
        float len = sqrt_f(dx * dx + dy * dy);
        if (len <= 1.0e-8f) { ox = cx + r; oy = cy; }
        else { float inv = r / len; ox = cx + dx * inv; oy = cy + dy * inv; }
        return true;
    }
    if (sh.kind == DVG_SHAPE_ELLIPSE) {
        closest_point_ellipse(sh.center.x, sh.center.y, sh.radius_e.x, sh.radius_e.y, px, py, ox, oy);
        return true;
    }
    if (sh.kind == DVG_SHAPE_RECT) {  // closest_point_rect, distance.rs:52-80
        float x0 = sh.rmin.x, y0 = sh.rmin.y, x1 = sh.rmax.x, y1 = sh.rmax.y;
        bool inside = px >= x0 && px <= x1 && py >= y0 && py <= y1;
        if (!inside) {
            ox = std::min(std::max(px, x0), x1);
            oy = std::min(std::max(py, y0), y1);
        } else {
            float dl = px - x0, dr = x1 - px, db = py - y0, dt = y1 - py;
            if (dl <= dr && dl <= db && dl <= dt) { ox = x0; oy = py; }
            else if (dr <= db && dr <= dt) { ox = x1; oy = py; }
            else if (db <= dt) { ox = px; oy = y0; }
            else { ox = px; oy = y1; }
        }
        return true;
    }
    // closest_point_path, distance.rs:142-231: the first segment reaching the
    // smallest distance; the point from its parameter
    bool approx = sh.path.use_distance_approx;
    float best = 1.0e20f;
    for (const Seg &s : pi.segs) {
        float d, t;
        if (s.kind == 0) distance_to_segment_with_t(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, d, t);
        else if (s.kind == 1) closest_quadratic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, approx, d, t);
        else closest_cubic(px, py, s.p0.x, s.p0.y, s.p1.x, s.p1.y, s.p2.x, s.p2.y, s.p3.x, s.p3.y, approx, d, t);
        if (d < best) {
            best = d;
            float tt = 1.0f - t;
            if (s.kind == 0) {
                ox = s.p0.x + t * (s.p1.x - s.p0.x);
                oy = s.p0.y + t * (s.p1.y - s.p0.y);
            } else if (s.kind == 1) {
                ox = tt * tt * s.p0.x + 2.0f * tt * t * s.p1.x + t * t * s.p2.x;
                oy = tt * tt * s.p0.y + 2.0f * tt * t * s.p1.y + t * t * s.p2.y;
            } else {
                ox = tt * tt * tt * s.p0.x + 3.0f * tt * tt * t * s.p1.x + 3.0f * tt * t * t * s.p2.x +
                     t * t * t * s.p3.x;
                oy = tt * tt * tt * s.p0.y + 3.0f * tt * tt * t * s.p1.y + 3.0f * tt * t * t * s.p2.y +
                     t * t * t * s.p3.y;
            }
        }
    }
    return best < 1.0e20f;
}

// compute_signed_distance, sdf.rs:9-117: the nearest shape boundary over the
// groups from last to first (compute_distance_group, distance.rs:298-513): a
// group's BVH walked by its skip pointers, a node skipped when its group-local
// box lies farther than the best canvas distance so far, every shape of a
// visited leaf measured (a group without shapes with bounds scans its shapes);
// negative inside the fill of the nearest shape's group (fill_inside_group,
// backward.rs:1529-1606: the forward's fill test).
static float signed_distance(const Scene &sc, const std::vector<PreparedGroup> &pgs, float px, float py) {
    float best = 1.0e20f;
    size_t best_g = 0;
    bool found = false;
    for (size_t gi = sc.groups.size(); gi-- > 0;) {
        const Group &g = sc.groups[gi];
        const PreparedGroup &pg = pgs[gi];
        const float *x = pg.xform;
        float lx = x[0] * px + x[1] * py + x[2];
        float ly = x[3] * px + x[4] * py + x[5];
        auto measure = [&](const PreparedInstance &pi) {
            const float *iv = pi.inv;
            float spx = iv[0] * lx + iv[1] * ly + iv[2];
            float spy = iv[3] * lx + iv[4] * ly + iv[5];
            const Shape &sh = sc.shapes[pi.shape];
            float cx = 0.f, cy = 0.f;
            if (!closest_point_shape(sh, pi, spx, spy, cx, cy)) return;
            const M3 &T = sh.transform;
            float gx = T.m[0][0] * cx + T.m[0][1] * cy + T.m[0][2];
            float gy = T.m[1][0] * cx + T.m[1][1] * cy + T.m[1][2];
            const M3 &Gc = g.shape_to_canvas;
            float ccx = Gc.m[0][0] * gx + Gc.m[0][1] * gy + Gc.m[0][2];
            float ccy = Gc.m[1][0] * gx + Gc.m[1][1] * gy + Gc.m[1][2];
            float dx = ccx - px, dy = ccy - py;
            float d = sqrt_f(dx * dx + dy * dy);
            if (d < best) {
                best = d;
                best_g = gi;
                found = true;
            }
        };
        if (!pg.bvh.empty()) {
            uint32_t node = 0;
            while (node != BVH_NONE) {
                const BvhNode &n = pg.bvh[node];
                // bounds_distance, distance.rs:9-15
                float ddx = max_f(max_f(n.b[0] - lx, 0.0f), lx - n.b[2]);
                float ddy = max_f(max_f(n.b[1] - ly, 0.0f), ly - n.b[3]);
                if (sqrt_f(ddx * ddx + ddy * ddy) > best) {
                    node = n.skip;
                } else if (n.count > 0) {
                    for (uint32_t i = 0; i < n.count; ++i) measure(pg.shapes[pg.bvh_items[n.start + i]]);
                    node = n.skip;
                } else {
                    node = n.left;
                }
            }
        } else {
            for (const PreparedInstance &pi : pg.shapes) measure(pi);
        }
    }
    if (!found) return 0.0f;
    const Group &g = sc.groups[best_g];
    bool inside = false;
    if (g.fill.kind != DVG_PAINT_NONE) {
        const PreparedGroup &pg = pgs[best_g];
        const float *x = pg.xform;
        float lx = x[0] * px + x[1] * py + x[2];
        float ly = x[3] * px + x[4] * py + x[5];
        GroupAccum acc;
        for (const PreparedInstance &pi : pg.shapes) {
            bool inb = pi.b[0] <= pi.b[2] && pi.b[1] <= pi.b[3];
            if (inb && (lx < pi.b[0] || lx > pi.b[2] || ly < pi.b[1] || ly > pi.b[3])) inb = false;
            if (!inb) continue;
            const float *iv = pi.inv;
            float spx = iv[0] * lx + iv[1] * ly + iv[2];
            float spy = iv[3] * lx + iv[4] * ly + iv[5];
            shape_fill(sc.shapes[pi.shape], pi, g.fill_rule, false, spx, spy, acc);
        }
        inside = g.fill_rule == 1 ? acc.crossings > 0.f : acc.winding != 0.f;
    }
    return inside ? -best : best;
}

// render_sdf_kernel, sdf.rs:119-200: the mean signed distance of each pixel's
// samples (the forward's jittered sample positions).
int render_sdf(const Scene &sc, const dvg_render_options &opt, float *out) {
    uint32_t W = sc.width, H = sc.height;
    size_t npix = (size_t)W * H;
    for (size_t i = 0; i < npix; ++i) out[i] = 0.0f;
    if (npix == 0 || sc.groups.empty()) return DVG_OK;
    std::vector<PreparedGroup> pgs;
    prepare(sc, opt, pgs);
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    bool jitter = !opt.use_prefiltering && opt.jitter != 0;
    float inv_sx = 1.0f / (float)nsx, inv_sy = 1.0f / (float)nsy;
    float inv_n = 1.0f / (float)(nsx * nsy);
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            float acc = 0.0f;
            for (uint32_t sy = 0; sy < nsy; ++sy)
                for (uint32_t sx = 0; sx < nsx; ++sx) {
                    float rx, ry;
                    jitter_xy(x, y, sx, sy, W, nsx, nsy, opt.seed, jitter, rx, ry);
                    float px = (float)x + ((float)sx + rx) * inv_sx;
                    float py = (float)y + ((float)sy + ry) * inv_sy;
                    acc += signed_distance(sc, pgs, px, py);
                }
            out[(size_t)y * W + x] = acc * inv_n;
        }
    return DVG_OK;
}

// eval_positions_kernel, sdf.rs:207-262: the signed distance at each canvas
// point (x, y pairs).
int eval_positions(const Scene &sc, const dvg_render_options &opt, const float *pos, size_t n, float *out) {
    for (size_t i = 0; i < n; ++i) out[i] = 0.0f;
    if (n == 0 || sc.groups.empty()) return DVG_OK;
    std::vector<PreparedGroup> pgs;
    prepare(sc, opt, pgs);
    for (size_t i = 0; i < n; ++i) out[i] = signed_distance(sc, pgs, pos[2 * i], pos[2 * i + 1]);
    return DVG_OK;
}

}  // namespace oracle
