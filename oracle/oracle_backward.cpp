// Oracle (test infrastructure only): restatement of the reference CPU backward
// pass, the gradient oracle.
//   render_backward            src/backward/render.rs:35-210
//   sample_color               src/backward/sampling.rs:76-224
//   sample_color_prefiltered   src/backward/sampling.rs:227-482
//   filters                    src/backward/filters.rs
//   paint adjoints             src/backward/paint.rs
//   background                 src/backward/background.rs
//   boundary sampling          src/backward/boundary.rs
// over the CPU queries of src/distance/*.rs (f64 root solvers).  The group
// and path BVHs of src/distance/bvh.rs only cull conservatively for the
// boolean and winding queries, so those are restated as loops; the
// closest-point query keeps the BVH's visiting order (distance/api.rs:138-232)
// because ties there pick the last-visited shape.
#include "oracle_backward.h"

namespace oracle {
namespace cpu {

// ---- distance/utils.rs -----------------------------------------------------
bool solve_quadratic(double a, double b, double c, double t[2]) {  // utils.rs:44-79
    const double eps = 1.0e-12;
    if (std::fabs(a) < eps) {
        if (std::fabs(b) < eps) return false;
        double root = -c / b;
        t[0] = root; t[1] = root;
        return true;
    }
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return false;
    double root = std::sqrt(disc);
    if (std::fabs(root) < eps) {
        double r = -0.5 * b / a;
        t[0] = r; t[1] = r;
        return true;
    }
    double q = (b < 0.0) ? -0.5 * (b - root) : -0.5 * (b + root);
    if (std::fabs(q) < eps) {
        double inv = 0.5 / a;
        t[0] = (-b - root) * inv;
        t[1] = (-b + root) * inv;
    } else {
        t[0] = q / a;
        t[1] = c / q;
    }
    if (t[0] > t[1]) std::swap(t[0], t[1]);
    return true;
}

static double cbrt64(double x) {  // utils.rs:115-123
    if (x > 0.0) return std::pow(x, 1.0 / 3.0);
    if (x < 0.0) return -std::pow(-x, 1.0 / 3.0);
    return 0.0;
}

int solve_cubic(double a, double b, double c, double d, double t[3]) {  // utils.rs:82-112
    if (std::fabs(a) < 1.0e-6) {
        double r[2];
        if (solve_quadratic(b, c, d, r)) {
            t[0] = r[0]; t[1] = r[1];
            return 2;
        }
        return 0;
    }
    b = b / a; c = c / a; d = d / a;
    double q = (b * b - 3.0 * c) / 9.0;
    double r = (2.0 * b * b * b - 9.0 * b * c + 27.0 * d) / 54.0;
    if (r * r < q * q * q) {
        double theta = std::acos(r / std::sqrt(q * q * q));
        const double pi = 3.14159265358979323846;
        t[0] = -2.0 * std::sqrt(q) * std::cos(theta / 3.0) - b / 3.0;
        t[1] = -2.0 * std::sqrt(q) * std::cos((theta + 2.0 * pi) / 3.0) - b / 3.0;
        t[2] = -2.0 * std::sqrt(q) * std::cos((theta - 2.0 * pi) / 3.0) - b / 3.0;
        return 3;
    }
    double a_root = (r > 0.0) ? -cbrt64(r + std::sqrt(r * r - q * q * q)) : cbrt64(-r + std::sqrt(r * r - q * q * q));
    double b_root = (std::fabs(a_root) > 1.0e-6) ? q / a_root : 0.0;
    t[0] = (a_root + b_root) - b / 3.0;
    return 1;
}

V2 transform_point_inverse(const M3 &m, V2 pt) {  // distance/utils.rs:9-17
    if (m.is_identity()) return pt;
    M3 inv;
    if (m.inverse(inv)) return inv.transform_point(pt);
    return pt;
}

// ---- distance/curve.rs ------------------------------------------------------
void distance_to_segment(V2 pt, V2 a, V2 b, float &dist, V2 &cp, float &t) {  // curve.rs:10-21
    V2 ab = b - a;
    float denom = ab.dot(ab);
    if (denom <= 1.0e-10f) {
        dist = (pt - a).length();
        cp = a;
        t = 0.f;
        return;
    }
    t = rclamp((pt - a).dot(ab) / denom, 0.f, 1.f);
    cp = a + ab * t;
    dist = (pt - cp).length();
}

static void quad_approx(V2 p0, V2 p1, V2 p2, V2 pt, V2 &cp, float &tout) {  // curve.rs:262-297
    V2 b0 = p0 - pt, b1 = p1 - pt, b2 = p2 - pt;
    float a = b0.cross(b2);
    float b = 2.0f * b1.cross(b0);
    float d = 2.0f * b2.cross(b1);
    float f = b * d - a * a;
    V2 d21 = b2 - b1, d10 = b1 - b0, d20 = b2 - b0;
    V2 gf = ((d21 * b) + (d10 * d)) + (d20 * a);
    gf = gf * 2.0f;
    gf = V2(gf.y, -gf.x);
    float t = 0.f;
    float denom = gf.dot(gf);
    if (denom > 1.0e-8f) {
        V2 pp = gf * (-f / denom);
        V2 d0p = b0 - pp;
        float ap = d0p.cross(d20);
        float bp = 2.0f * d10.cross(d0p);
        float denom2 = 2.0f * a + b + d;
        if (std::fabs(denom2) > 1.0e-8f) t = rclamp((ap + bp) / denom2, 0.f, 1.f);
    }
    float tt = 1.0f - t;
    cp = ((p0 * (tt * tt)) + (p1 * (2.0f * tt * t))) + (p2 * (t * t));
    tout = t;
}

void closest_point_quadratic(V2 pt, V2 p0, V2 p1, V2 p2, bool approx, V2 &cp, float &tout,
                             float &dist) {  // curve.rs:26-74
    if (approx) {
        quad_approx(p0, p1, p2, pt, cp, tout);
        dist = (pt - cp).length();
        return;
    }
    float best_t = 0.f;
    V2 best_pt = p0;
    float best = (pt - p0).length();
    float de = (pt - p2).length();
    if (de < best) { best = de; best_pt = p2; best_t = 1.f; }
    double ax = (double)(p0.x - 2.0f * p1.x + p2.x), ay = (double)(p0.y - 2.0f * p1.y + p2.y);
    double bx = (double)(-p0.x + p1.x), by = (double)(-p0.y + p1.y);
    double cx = (double)(p0.x - pt.x), cy = (double)(p0.y - pt.y);
    double a = ax * ax + ay * ay;
    double b = 3.0 * (ax * bx + ay * by);
    double c = 2.0 * (bx * bx + by * by) + (ax * cx + ay * cy);
    double d = bx * cx + by * cy;
    double roots[3];
    int n = solve_cubic(a, b, c, d, roots);
    for (int i = 0; i < n; ++i) {
        double t = roots[i];
        if (t >= 0.0 && t <= 1.0) {
            float t32 = (float)t;
            float tt = 1.0f - t32;
            V2 q = ((p0 * (tt * tt)) + (p1 * (2.0f * tt * t32))) + (p2 * (t32 * t32));
            float dd = (pt - q).length();
            if (dd < best) { best = dd; best_pt = q; best_t = t32; }
        }
    }
    cp = best_pt;
    tout = best_t;
    dist = best;
}

static void cubic_approx(V2 pt, V2 p0, V2 p1, V2 p2, V2 p3, V2 &cp, float &tout, float &dist) {  // curve.rs:77-117
    const size_t steps = 8;
    float inv = 1.0f / (float)steps;
    V2 prev = p0;
    float prev_t = 0.f;
    V2 best_pt = p0;
    float best_t = 0.f;
    float best = (pt - p0).length();
    for (size_t i = 1; i <= steps; ++i) {
        float t = (float)i * inv;
        float tt = 1.0f - t, tt2 = tt * tt, t2 = t * t;
        float a = tt2 * tt, b = 3.0f * tt2 * t, c = 3.0f * tt * t2, d = t2 * t;
        V2 curr(a * p0.x + b * p1.x + c * p2.x + d * p3.x, a * p0.y + b * p1.y + c * p2.y + d * p3.y);
        float sd, st;
        V2 scp;
        distance_to_segment(pt, prev, curr, sd, scp, st);
        float lt = prev_t + st * (t - prev_t);
        if (sd < best) { best = sd; best_pt = scp; best_t = lt; }
        prev = curr;
        prev_t = t;
    }
    cp = best_pt;
    tout = best_t;
    dist = best;
}

void closest_point_cubic(V2 pt, V2 p0, V2 p1, V2 p2, V2 p3, bool approx, V2 &cp, float &tout,
                         float &dist) {  // curve.rs:122-259
    if (approx) {
        cubic_approx(pt, p0, p1, p2, p3, cp, tout, dist);
        return;
    }
    float best_t = 0.f;
    V2 best_pt = p0;
    float best = (pt - p0).length();
    float de = (pt - p3).length();
    if (de < best) { best = de; best_pt = p3; best_t = 1.f; }
    double ax = (double)(-p0.x + 3.0f * p1.x - 3.0f * p2.x + p3.x);
    double ay = (double)(-p0.y + 3.0f * p1.y - 3.0f * p2.y + p3.y);
    double bx = (double)(3.0f * p0.x - 6.0f * p1.x + 3.0f * p2.x);
    double by = (double)(3.0f * p0.y - 6.0f * p1.y + 3.0f * p2.y);
    double cx = (double)(-3.0f * p0.x + 3.0f * p1.x);
    double cy = (double)(-3.0f * p0.y + 3.0f * p1.y);
    double dx = (double)(p0.x - pt.x), dy = (double)(p0.y - pt.y);
    double a = 3.0 * (ax * ax + ay * ay);
    if (std::fabs(a) < 1.0e-10) {
        cp = best_pt; tout = best_t; dist = best;
        return;
    }
    double b = 5.0 * (ax * bx + ay * by);
    double c = 4.0 * (ax * cx + ay * cy) + 2.0 * (bx * bx + by * by);
    double d = 3.0 * ((bx * cx + by * cy) + (ax * dx + ay * dy));
    double e = (cx * cx + cy * cy) + 2.0 * (dx * bx + dy * by);
    double f = dx * cx + dy * cy;
    b = b / a; c = c / a; d = d / a; e = e / a; f = f / a;
    double p1a = (2.0 / 5.0) * c - (4.0 / 25.0) * b * b;
    double p1b = (3.0 / 5.0) * d - (3.0 / 25.0) * b * c;
    double p1c = (4.0 / 5.0) * e - (2.0 / 25.0) * b * d;
    double p1d = f - b * e / 25.0;
    double q_root = -b / 5.0;
    double p_roots[3];
    int num_p = solve_cubic(p1a, p1b, p1c, p1d, p_roots);
    double iv[4];
    int ni = 0;
    if (q_root >= 0.0 && q_root <= 1.0) iv[ni++] = q_root;
    for (int i = 0; i < num_p; ++i) iv[ni++] = p_roots[i];
    for (int j = 1; j < ni; ++j) {
        int k = j;
        while (k > 0 && iv[k - 1] > iv[k]) {
            std::swap(iv[k - 1], iv[k]);
            --k;
        }
    }
    auto ev = [&](double t) { return ((((t + b) * t + c) * t + d) * t + e) * t + f; };
    auto evd = [&](double t) { return (((5.0 * t + 4.0 * b) * t + 3.0 * c) * t + 2.0 * d) * t + e; };
    double lower = 0.0;
    for (int j = 0; j <= ni; ++j) {
        if (j < ni && iv[j] < 0.0) continue;
        double upper = (j < ni) ? std::fmin(iv[j], 1.0) : 1.0;
        double lb = lower, ub = upper;
        double lbe = ev(lb), ube = ev(ub);
        if (lbe * ube > 0.0) {
            lower = upper;
            continue;
        }
        if (lbe > ube) {
            std::swap(lb, ub);
            std::swap(lbe, ube);
        }
        double t = 0.5 * (lb + ub);
        for (int it = 0; it < 20; ++it) {
            if (!(t >= lb && t <= ub)) t = 0.5 * (lb + ub);
            double value = ev(t);
            if (std::fabs(value) < 1.0e-5 || it == 19) break;
            if (value > 0.0) ub = t; else lb = t;
            double der = evd(t);
            t -= value / der;
        }
        if (t >= 0.0 && t <= 1.0) {
            float t32 = (float)t;
            float tt = 1.0f - t32;
            V2 q = (((p0 * (tt * tt * tt)) + (p1 * (3.0f * tt * tt * t32))) + (p2 * (3.0f * tt * t32 * t32))) +
                   (p3 * (t32 * t32 * t32));
            float dd = (pt - q).length();
            if (dd < best) { best = dd; best_pt = q; best_t = t32; }
        }
        if (upper >= 1.0) break;
        lower = upper;
    }
    cp = best_pt;
    tout = best_t;
    dist = best;
}

// ---- distance/shape.rs -------------------------------------------------------
V2 closest_point_circle(V2 c, float r, V2 pt) {  // shape.rs:107-114
    V2 d = pt - c;
    float len = d.length();
    if (len <= 1.0e-8f) return V2(c.x + r, c.y);
    return c + d * (r / len);
}

V2 closest_point_ellipse(V2 c, V2 radius, V2 pt) {  // shape.rs:117-164
    float rx = std::fabs(radius.x), ry = std::fabs(radius.y);
    const float eps = 1.0e-6f;
    if (rx < eps && ry < eps) return c;
    if (rx < eps) return V2(c.x, c.y + rclamp(pt.y - c.y, -ry, ry));
    if (ry < eps) return V2(c.x + rclamp(pt.x - c.x, -rx, rx), c.y);
    float dx = pt.x - c.x, dy = pt.y - c.y;
    if (std::fabs(dx) < eps && std::fabs(dy) < eps) return V2(c.x + rx, c.y);
    float sx = dx < 0.f ? -1.f : 1.f, sy = dy < 0.f ? -1.f : 1.f;
    dx = std::fabs(dx);
    dy = std::fabs(dy);
    float t = atan2_f(dy * rx, dx * ry);
    for (int i = 0; i < 20; ++i) {
        float s = sin_f(t), co = cos_f(t);
        float g = rx * dx * s - ry * dy * co + (ry * ry - rx * rx) * s * co;
        float gt = rx * dx * co + ry * dy * s + (ry * ry - rx * rx) * (co * co - s * s);
        if (std::fabs(gt) < 1.0e-12f) break;
        float next = rclamp(t - g / gt, 0.f, 1.57079632679489661923f);
        if (std::fabs(next - t) < 1.0e-6f) { t = next; break; }
        t = next;
    }
    float s = sin_f(t), co = cos_f(t);
    return V2(c.x + sx * rx * co, c.y + sy * ry * s);
}

V2 closest_point_rect(V2 mn, V2 mx, V2 pt) {  // shape.rs:167-185
    bool inside = pt.x >= mn.x && pt.x <= mx.x && pt.y >= mn.y && pt.y <= mx.y;
    if (!inside) return V2(rclamp(pt.x, mn.x, mx.x), rclamp(pt.y, mn.y, mx.y));
    float dl = pt.x - mn.x, dr = mx.x - pt.x, db = pt.y - mn.y, dt = mx.y - pt.y;
    if (dl <= dr && dl <= db && dl <= dt) return V2(mn.x, pt.y);
    if (dr <= db && dr <= dt) return V2(mx.x, pt.y);
    if (db <= dt) return V2(pt.x, mn.y);
    return V2(pt.x, mx.y);
}

float rect_signed_distance(V2 pt, V2 mn, V2 mx) {  // shape.rs:188-201
    float dx = rmax(rmax(mn.x - pt.x, 0.f), pt.x - mx.x);
    float dy = rmax(rmax(mn.y - pt.y, 0.f), pt.y - mx.y);
    float outside = sqrt_f(dx * dx + dy * dy);
    float inside = rmin(rmin(rmin(pt.x - mn.x, mx.x - pt.x), pt.y - mn.y), mx.y - pt.y);
    return outside > 0.f ? outside : -inside;
}

float ellipse_signed_distance(V2 pt, V2 c, V2 radius) {  // shape.rs:204-212
    float rx = rmax(std::fabs(radius.x), 1.0e-6f), ry = rmax(std::fabs(radius.y), 1.0e-6f);
    V2 l = pt - c;
    bool inside = (l.x / rx) * (l.x / rx) + (l.y / ry) * (l.y / ry) <= 1.0f;
    V2 cl = closest_point_ellipse(c, V2(rx, ry), pt);
    float d = (pt - cl).length();
    return inside ? -d : d;
}

// ---- distance/winding.rs -----------------------------------------------------
int winding_number_segment(uint32_t kind, V2 p0, V2 p1, V2 p2, V2 p3, V2 pt) {  // winding.rs:118-192
    if (kind == 0) {
        float dy = p1.y - p0.y;
        if (std::fabs(dy) <= 1.0e-8f) return 0;
        float t = (pt.y - p0.y) / dy;
        if (t >= 0.f && t <= 1.f) {
            float tp = p0.x - pt.x + t * (p1.x - p0.x);
            if (tp >= 0.f) return dy > 0.f ? 1 : -1;
        }
        return 0;
    }
    if (kind == 1) {
        double a = (double)(p0.y - 2.0f * p1.y + p2.y);
        double b = (double)(-2.0f * p0.y + 2.0f * p1.y);
        double c = (double)(p0.y - pt.y);
        double roots[2] = {0.0, 0.0};
        int n = 0;
        if (std::fabs(a) < 1.0e-12) {
            if (std::fabs(b) > 1.0e-12) { roots[0] = -c / b; n = 1; }
        } else if (solve_quadratic(a, b, c, roots)) {
            n = 2;
        }
        int w = 0;
        double ax = (double)(p0.x - 2.0f * p1.x + p2.x);
        double bx = (double)(-2.0f * p0.x + 2.0f * p1.x);
        double cx = (double)(p0.x - pt.x);
        for (int i = 0; i < n; ++i) {
            double t = roots[i];
            if (t >= 0.0 && t <= 1.0) {
                double tp = ax * t * t + bx * t + cx;
                if (tp > 0.0) {
                    double der = 2.0 * a * t + b;
                    w += der > 0.0 ? 1 : -1;
                }
            }
        }
        return w;
    }
    if (kind == 2) {
        double a = (double)(-p0.y + 3.0f * p1.y - 3.0f * p2.y + p3.y);
        double b = (double)(3.0f * p0.y - 6.0f * p1.y + 3.0f * p2.y);
        double c = (double)(-3.0f * p0.y + 3.0f * p1.y);
        double d = (double)(p0.y - pt.y);
        double roots[3];
        int n = solve_cubic(a, b, c, d, roots);
        double ax = (double)(-p0.x + 3.0f * p1.x - 3.0f * p2.x + p3.x);
        double bx = (double)(3.0f * p0.x - 6.0f * p1.x + 3.0f * p2.x);
        double cx = (double)(-3.0f * p0.x + 3.0f * p1.x);
        double dx = (double)(p0.x - pt.x);
        int w = 0;
        for (int i = 0; i < n; ++i) {
            double t = roots[i];
            if (t >= 0.0 && t <= 1.0) {
                double tp = ax * t * t * t + bx * t * t + cx * t + dx;
                if (tp > 0.0) {
                    double der = 3.0 * a * t * t + 2.0 * b * t + c;
                    w += der > 0.0 ? 1 : -1;
                }
            }
        }
        return w;
    }
    return 0;
}

// Segments of a path shape as distance/bvh.rs:241-400 builds them (default
// radius = stroke_width).
static const std::vector<Seg> &shape_segs(const Ctx &cx, uint32_t si) { return cx.segs[si]; }

int winding_number_shape(const Ctx &cx, uint32_t si, V2 pt) {  // winding.rs:13-44
    const Shape &sh = cx.sc->shapes[si];
    switch (sh.kind) {
        case DVG_SHAPE_CIRCLE: {
            float r = std::fabs(sh.radius_c);
            return (pt - sh.center).length_squared() < r * r ? 1 : 0;
        }
        case DVG_SHAPE_ELLIPSE: {
            float rx = rmax(std::fabs(sh.radius_e.x), 1.0e-6f), ry = rmax(std::fabs(sh.radius_e.y), 1.0e-6f);
            float dx = (pt.x - sh.center.x) / rx, dy = (pt.y - sh.center.y) / ry;
            return dx * dx + dy * dy < 1.0f ? 1 : 0;
        }
        case DVG_SHAPE_RECT:
            return (pt.x > sh.rmin.x && pt.x < sh.rmax.x && pt.y > sh.rmin.y && pt.y < sh.rmax.y) ? 1 : 0;
        default: {
            int w = 0;
            for (const Seg &s : shape_segs(cx, si)) w += winding_number_segment(s.kind, s.p0, s.p1, s.p2, s.p3, pt);
            return w;
        }
    }
}

static inline float seg_radius(const Seg &s, float t) {  // closest.rs:68-81
    if (s.kind == 0) return s.r0 + t * (s.r1 - s.r0);
    float tt = 1.0f - t;
    if (s.kind == 1) return tt * tt * s.r0 + 2.0f * tt * t * s.r1 + t * t * s.r2;
    return tt * tt * tt * s.r0 + 3.0f * tt * tt * t * s.r1 + 3.0f * tt * t * t * s.r2 + t * t * t * s.r3;
}

void closest_on_seg(const Seg &s, V2 pt, bool approx, V2 &cp, float &t, float &dist) {
    if (s.kind == 0) distance_to_segment(pt, s.p0, s.p1, dist, cp, t);
    else if (s.kind == 1) closest_point_quadratic(pt, s.p0, s.p1, s.p2, approx, cp, t, dist);
    else closest_point_cubic(pt, s.p0, s.p1, s.p2, s.p3, approx, cp, t, dist);
}

// within_distance_shape (closest.rs:207-357) for the shape's own stroke width
// (the caller has checked stroke_width > 0); pt is group-local.
bool within_distance_shape(const Ctx &cx, uint32_t si, V2 pt, float r) {
    if (r <= 0.f) return false;
    const Shape &sh = cx.sc->shapes[si];
    V2 lp = transform_point_inverse(sh.transform, pt);
    switch (sh.kind) {
        case DVG_SHAPE_CIRCLE: {
            float d = (lp - sh.center).length() - std::fabs(sh.radius_c);
            return std::fabs(d) < r;
        }
        case DVG_SHAPE_ELLIPSE: {
            float d = ellipse_signed_distance(lp, sh.center, V2(std::fabs(sh.radius_e.x), std::fabs(sh.radius_e.y)));
            return std::fabs(d) < r;
        }
        case DVG_SHAPE_RECT: {
            float d = rect_signed_distance(lp, sh.rmin, sh.rmax);
            return std::fabs(d) < r;
        }
        default: {
            // The path segments carry radii from thickness or the default r,
            // which equals the shape's stroke width at every call site here.
            for (const Seg &s : shape_segs(cx, si)) {
                V2 cp;
                float t, d;
                closest_on_seg(s, lp, sh.path.use_distance_approx, cp, t, d);
                if (d < seg_radius(s, t)) return true;
            }
            return false;
        }
    }
}

// shape_bounds (distance/shape.rs:9-105): geometry bounds under the shape
// transform, inflated by the max stroke radius.
bool cpu_shape_bounds(const Shape &sh, float b[4]) {
    const M3 &T = sh.transform;
    V2 mn, mx;
    if (sh.kind == DVG_SHAPE_CIRCLE || sh.kind == DVG_SHAPE_ELLIPSE) {
        float rx = sh.kind == DVG_SHAPE_CIRCLE ? std::fabs(sh.radius_c) : std::fabs(sh.radius_e.x);
        float ry = sh.kind == DVG_SHAPE_CIRCLE ? std::fabs(sh.radius_c) : std::fabs(sh.radius_e.y);
        V2 c = T.transform_point(sh.center);
        float ex = sqrt_f((T.m[0][0] * rx) * (T.m[0][0] * rx) + (T.m[0][1] * ry) * (T.m[0][1] * ry));
        float ey = sqrt_f((T.m[1][0] * rx) * (T.m[1][0] * rx) + (T.m[1][1] * ry) * (T.m[1][1] * ry));
        mn = V2(c.x - ex, c.y - ey);
        mx = V2(c.x + ex, c.y + ey);
    } else if (sh.kind == DVG_SHAPE_RECT) {
        V2 cs[4] = {V2(sh.rmin.x, sh.rmin.y), V2(sh.rmax.x, sh.rmin.y), V2(sh.rmax.x, sh.rmax.y), V2(sh.rmin.x, sh.rmax.y)};
        mn = V2(INFINITY, INFINITY);
        mx = V2(-INFINITY, -INFINITY);
        for (V2 c : cs) { V2 t = T.transform_point(c); mn = mn.vmin(t); mx = mx.vmax(t); }
    } else {
        if (sh.path.points.empty()) return false;
        mn = V2(INFINITY, INFINITY);
        mx = V2(-INFINITY, -INFINITY);
        for (V2 p : sh.path.points) { V2 t = T.transform_point(p); mn = mn.vmin(t); mx = mx.vmax(t); }
    }
    float scale = rmax(T.max_scale(), 0.f);
    float r = std::fabs(sh.stroke_width) * scale;
    if (sh.kind == DVG_SHAPE_PATH && sh.path.has_thickness) {
        float mt = 0.f;
        for (float v : sh.path.thickness) mt = rmax(mt, std::fabs(v));
        r = rmax(r, mt * scale);
    }
    if (r > 0.f && sh.join == DVG_JOIN_MITER) r *= rmax(sh.miter_limit, 1.0f);
    float pad = rmax(r, 0.f);
    b[0] = mn.x - pad; b[1] = mn.y - pad; b[2] = mx.x + pad; b[3] = mx.y + pad;
    return true;
}

void build_group_bounds(Ctx &cx) {
    const Scene &sc = *cx.sc;
    cx.group_bounds.assign(sc.groups.size() * 4, 0.f);
    cx.group_has_bounds.assign(sc.groups.size(), 0);
    for (size_t gi = 0; gi < sc.groups.size(); ++gi) {
        float u[4] = {INFINITY, INFINITY, -INFINITY, -INFINITY};
        bool any = false;
        for (uint32_t si : sc.groups[gi].shape_indices) {
            float b[4];
            if (si >= sc.shapes.size() || !cpu_shape_bounds(sc.shapes[si], b)) continue;
            u[0] = rmin(u[0], b[0]); u[1] = rmin(u[1], b[1]); u[2] = rmax(u[2], b[2]); u[3] = rmax(u[3], b[3]);
            any = true;
        }
        std::memcpy(&cx.group_bounds[gi * 4], u, sizeof(u));
        cx.group_has_bounds[gi] = any ? 1 : 0;
    }
}

// The root-node test of the group BVH (api.rs:101-103, 282-284): a point
// outside the union of the shape bounds meets no shape of the group.
static bool outside_group(const Ctx &cx, uint32_t gi, V2 lp) {
    if (cx.group_has_bounds.empty() || !cx.group_has_bounds[gi]) return false;
    const float *b = &cx.group_bounds[(size_t)gi * 4];
    return !(lp.x >= b[0] && lp.x <= b[2] && lp.y >= b[1] && lp.y <= b[3]);
}

// within_distance_bvh (api.rs:79-135)
bool within_distance_group(const Ctx &cx, uint32_t gi, V2 pt) {
    const Group &g = cx.sc->groups[gi];
    V2 lp = g.canvas_to_shape.transform_point(pt);
    if (outside_group(cx, gi, lp)) return false;
    for (uint32_t si : g.shape_indices) {
        const Shape &sh = cx.sc->shapes[si];
        if (sh.stroke_width <= 0.f) continue;
        if (within_distance_shape(cx, si, lp, sh.stroke_width)) return true;
    }
    return false;
}

// is_inside_bvh (api.rs:259-309)
bool is_inside_group(const Ctx &cx, uint32_t gi, V2 pt) {
    const Group &g = cx.sc->groups[gi];
    V2 lp = g.canvas_to_shape.transform_point(pt);
    if (outside_group(cx, gi, lp)) return false;
    int w = 0;
    for (uint32_t si : g.shape_indices) {
        V2 sp = transform_point_inverse(cx.sc->shapes[si].transform, lp);
        w += winding_number_shape(cx, si, sp);
    }
    if (g.fill_rule == DVG_FILL_EVENODD) return (std::abs(w) % 2) == 1;
    return w != 0;
}

// ---- backward/paint.rs -----------------------------------------------------
V4 paint_color(const Paint &p, V2 pt) {  // paint.rs:8-74
    if (p.kind == DVG_PAINT_SOLID) return p.color;
    if (p.kind != DVG_PAINT_LINEAR && p.kind != DVG_PAINT_RADIAL) return V4();
    float t;
    if (!p.radial) {
        V2 beg = p.p0, end = p.p1;
        float denom = rmax((end - beg).dot(end - beg), 1.0e-3f);
        t = (pt - beg).dot(end - beg) / denom;
    } else {
        V2 off = pt - p.p0;
        V2 n(off.x / p.p1.x, off.y / p.p1.y);
        t = n.length();
    }
    size_t ns = p.stops.size();
    if (ns == 0) return V4();
    if (t < p.stops[0].offset) return p.stops[0].color;
    for (size_t i = 0; i + 1 < ns; ++i) {
        const Stop &cu = p.stops[i], &nx = p.stops[i + 1];
        if (t >= cu.offset && t < nx.offset) {
            float tt = (t - cu.offset) / (nx.offset - cu.offset);
            const V4 &c0 = cu.color, &c1 = nx.color;
            return V4(c0.x * (1.0f - tt) + c1.x * tt, c0.y * (1.0f - tt) + c1.y * tt, c0.z * (1.0f - tt) + c1.z * tt,
                      c0.w * (1.0f - tt) + c1.w * tt);
        }
    }
    return p.stops[ns - 1].color;
}

V2 d_length(V2 v, float d_l) {  // backward/math.rs:33-41
    float l = sqrt_f(v.dot(v));
    if (l == 0.f) return V2();
    float d_l_sq = 0.5f * d_l / l;
    return v * (2.0f * d_l_sq);
}

static void add_stop_color(Grads &gr, uint32_t stop, V4 d) {
    float *c = &gr.d_stop_colors[(size_t)stop * 4];
    c[0] += d.x; c[1] += d.y; c[2] += d.z; c[3] += d.w;
}

// d_sample_paint (paint.rs:77-280).  d_paint_solid points at the group's
// packed colour slot.
void d_sample_paint(const Paint &p, Grads &gr, float *d_paint_solid, V2 pt, V4 d, V2 &d_trans) {
    if (p.kind == DVG_PAINT_SOLID) {
        d_paint_solid[0] += d.x; d_paint_solid[1] += d.y; d_paint_solid[2] += d.z; d_paint_solid[3] += d.w;
        return;
    }
    if (p.kind != DVG_PAINT_LINEAR && p.kind != DVG_PAINT_RADIAL) return;
    size_t ns = p.stops.size();
    if (ns == 0) return;
    float *dg = &gr.d_gradient_data[(size_t)p.grad_index * 8];
    uint32_t so = p.stop_offset;
    float t, denom = 0.f;
    V2 beg = p.p0, end = p.p1, off, nrm;
    if (!p.radial) {
        denom = rmax((end - beg).dot(end - beg), 1.0e-3f);
        t = (pt - beg).dot(end - beg) / denom;
    } else {
        off = pt - p.p0;
        nrm = V2(off.x / p.p1.x, off.y / p.p1.y);
        t = nrm.length();
    }
    if (t < p.stops[0].offset) {
        add_stop_color(gr, so, d);
        return;
    }
    for (size_t i = 0; i + 1 < ns; ++i) {
        const Stop &cu = p.stops[i], &nx = p.stops[i + 1];
        if (t >= cu.offset && t < nx.offset) {
            float tt = (t - cu.offset) / (nx.offset - cu.offset);
            V4 cc = cu.color, cn = nx.color;
            add_stop_color(gr, so + (uint32_t)i, V4(d.x * (1.0f - tt), d.y * (1.0f - tt), d.z * (1.0f - tt), d.w * (1.0f - tt)));
            add_stop_color(gr, so + (uint32_t)i + 1, V4(d.x * tt, d.y * tt, d.z * tt, d.w * tt));
            float d_tt = d.x * (cn.x - cc.x) + d.y * (cn.y - cc.y) + d.z * (cn.z - cc.z) + d.w * (cn.w - cc.w);
            float denom_off = nx.offset - cu.offset;
            if (std::fabs(denom_off) > 0.f) {
                float d_next = -d_tt * tt / denom_off;
                float d_curr = d_tt * (tt - 1.0f) / denom_off;
                gr.d_stop_offsets[so + i] += d_curr;
                gr.d_stop_offsets[so + i + 1] += d_next;
            }
            float d_t = d_tt / denom_off;
            if (!p.radial) {
                V2 d_beg(d_t * (-(pt.x - beg.x) - (end.x - beg.x)) / denom, d_t * (-(pt.y - beg.y) - (end.y - beg.y)) / denom);
                V2 d_end(d_t * (pt.x - beg.x) / denom, d_t * (pt.y - beg.y) / denom);
                float d_l = -d_t * t / denom;
                if ((end - beg).dot(end - beg) > 1.0e-3f) {
                    V2 adj(beg.x - end.x, beg.y - end.y);
                    dg[1] += d_beg.x + 2.0f * d_l * adj.x;
                    dg[2] += d_beg.y + 2.0f * d_l * adj.y;
                    dg[3] += d_end.x + 2.0f * d_l * (end.x - beg.x);
                    dg[4] += d_end.y + 2.0f * d_l * (end.y - beg.y);
                } else {
                    dg[1] += d_beg.x; dg[2] += d_beg.y; dg[3] += d_end.x; dg[4] += d_end.y;
                }
                d_trans += d_beg + d_end;
            } else {
                V2 dn = d_length(nrm, d_t);
                V2 d_off(dn.x / p.p1.x, dn.y / p.p1.y);
                V2 d_rad(-dn.x * off.x / (p.p1.x * p.p1.x), -dn.y * off.y / (p.p1.y * p.p1.y));
                V2 d_center(-d_off.x, -d_off.y);
                dg[1] += d_center.x; dg[2] += d_center.y;
                dg[3] += d_rad.x; dg[4] += d_rad.y;
                d_trans += d_center;
            }
            return;
        }
    }
    add_stop_color(gr, so + (uint32_t)ns - 1, d);
}

// ---- backward/filters.rs ---------------------------------------------------
static inline float powi2(float x) { return x * x; }
static inline float powi3(float x) { return (x * x) * x; }
static inline float powi5(float x) { float x2 = x * x; return (x2 * x2) * x; }

float compute_filter_weight(uint32_t type, float radius, float dx, float dy) {  // filters.rs:159-188
    if (radius <= 0.f) return 1.f;
    if (std::fabs(dx) > radius || std::fabs(dy) > radius) return 0.f;
    switch (type) {
        case DVG_FILTER_BOX: return 1.0f / powi2(2.0f * radius);
        case DVG_FILTER_TENT: {
            float fx = radius - std::fabs(dx), fy = radius - std::fabs(dy);
            float norm = 1.0f / (radius * radius);
            return fx * fy * norm * norm;
        }
        case DVG_FILTER_RADIAL_PARABOLIC: {
            float rx = 1.0f - (dx / radius) * (dx / radius);
            float ry = 1.0f - (dy / radius) * (dy / radius);
            return (4.0f / 3.0f) * rx * (4.0f / 3.0f) * ry;
        }
        default: {
            float ndx = (dx / (2.0f * radius)) + 0.5f, ndy = (dy / (2.0f * radius)) + 0.5f;
            float fx = 0.5f * (1.0f - cos_f(2.0f * PI_F * ndx));
            float fy = 0.5f * (1.0f - cos_f(2.0f * PI_F * ndy));
            float norm = 1.0f / (radius * radius);
            return fx * fy * norm;
        }
    }
}

static void d_compute_filter_weight(uint32_t type, float radius, float dx, float dy, float d_ret,
                                    float &d_radius) {  // filters.rs:190-239
    switch (type) {
        case DVG_FILTER_BOX: {
            float denom = 2.0f * radius;
            if (denom != 0.f) d_radius += d_ret * (-2.0f) * denom / powi3(denom);
            break;
        }
        case DVG_FILTER_TENT: {
            float fx = radius - std::fabs(dx), fy = radius - std::fabs(dy);
            float norm = 1.0f / (radius * radius);
            float d_fx = d_ret * fy * norm, d_fy = d_ret * fx * norm, d_norm = d_ret * fx * fy;
            if (radius != 0.f) d_radius += d_fx + d_fy + (-4.0f) * d_norm / powi5(radius);
            break;
        }
        case DVG_FILTER_RADIAL_PARABOLIC: {
            float r3 = radius * radius * radius;
            if (r3 != 0.f) d_radius += -(2.0f * dx * dx + 2.0f * dy * dy) / r3;
            break;
        }
        default: {
            float ndx = (dx / (2.0f * radius)) + 0.5f, ndy = (dy / (2.0f * radius)) + 0.5f;
            float fx = 0.5f * (1.0f - cos_f(2.0f * PI_F * ndx));
            float fy = 0.5f * (1.0f - cos_f(2.0f * PI_F * ndy));
            float norm = 1.0f / (radius * radius);
            float d_fx = d_ret * fy * norm, d_fy = d_ret * fx * norm, d_norm = d_ret * fx * fy;
            float d_ndx = d_fx * 0.5f * sin_f(2.0f * PI_F * ndx) * (2.0f * PI_F);
            float d_ndy = d_fy * 0.5f * sin_f(2.0f * PI_F * ndy) * (2.0f * PI_F);
            d_radius += d_ndx * (-2.0f * dx / powi2(2.0f * radius)) + d_ndy * (-2.0f * dy / powi2(2.0f * radius)) +
                        (-2.0f) * d_norm / powi3(radius);
            break;
        }
    }
}

std::vector<float> build_weight_image(const Scene &sc, uint32_t nsx, uint32_t nsy, uint32_t seed, bool prefilter,
                                      bool jitter) {  // filters.rs:11-67
    uint32_t W = sc.width, H = sc.height;
    std::vector<float> img((size_t)W * H, 0.f);
    size_t total = (size_t)W * H * nsx * nsy;
    int ri = (int)std::ceil(sc.filter_radius);
    for (size_t idx = 0; idx < total; ++idx) {
        size_t sx = idx % nsx, sy = (idx / nsx) % nsy, x = (idx / (nsx * nsy)) % W, y = idx / ((size_t)nsx * nsy * W);
        float rx = 0.5f, ry = 0.5f;
        if (jitter) {
            Pcg32 rng(idx, seed);
            rx = rng.next_f32();
            ry = rng.next_f32();
        }
        if (prefilter) { rx = 0.5f; ry = 0.5f; }
        float px = (float)x + ((float)sx + rx) / (float)nsx;
        float py = (float)y + ((float)sy + ry) / (float)nsy;
        for (int dy = -ri; dy <= ri; ++dy)
            for (int dx = -ri; dx <= ri; ++dx) {
                int xx = (int)x + dx, yy = (int)y + dy;
                if (xx >= 0 && yy >= 0 && xx < (int)W && yy < (int)H) {
                    float w = compute_filter_weight(sc.filter_type, sc.filter_radius, ((float)xx + 0.5f) - px,
                                                    ((float)yy + 0.5f) - py);
                    img[(size_t)yy * W + xx] += w;
                }
            }
    }
    return img;
}

V4 gather_d_color(const Scene &sc, const float *d_render, const std::vector<float> &wimg, V2 pt) {  // filters.rs:70-103
    int W = (int)sc.width, H = (int)sc.height;
    int x = (int)std::floor(pt.x), y = (int)std::floor(pt.y);
    int ri = (int)std::ceil(sc.filter_radius);
    V4 out;
    for (int dy = -ri; dy <= ri; ++dy)
        for (int dx = -ri; dx <= ri; ++dx) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && yy >= 0 && xx < W && yy < H) {
                float w = compute_filter_weight(sc.filter_type, sc.filter_radius, ((float)xx + 0.5f) - pt.x,
                                                ((float)yy + 0.5f) - pt.y);
                float ws = wimg[(size_t)yy * W + xx];
                if (ws > 0.f) {
                    size_t b = ((size_t)yy * W + xx) * 4;
                    out.x += (w / ws) * d_render[b];
                    out.y += (w / ws) * d_render[b + 1];
                    out.z += (w / ws) * d_render[b + 2];
                    out.w += (w / ws) * d_render[b + 3];
                }
            }
        }
    return out;
}

void accumulate_filter_gradient(const Scene &sc, V4 color, const float *d_render, const std::vector<float> &wimg,
                                V2 pt, float &d_radius) {  // filters.rs:106-156
    int W = (int)sc.width, H = (int)sc.height;
    int x = (int)std::floor(pt.x), y = (int)std::floor(pt.y);
    int ri = (int)std::ceil(sc.filter_radius);
    for (int dy = -ri; dy <= ri; ++dy)
        for (int dx = -ri; dx <= ri; ++dx) {
            int xx = x + dx, yy = y + dy;
            if (xx >= 0 && yy >= 0 && xx < W && yy < H) {
                float ws = wimg[(size_t)yy * W + xx];
                if (ws <= 0.f) continue;
                float xc = (float)xx + 0.5f, yc = (float)yy + 0.5f;
                float w = compute_filter_weight(sc.filter_type, sc.filter_radius, xc - pt.x, yc - pt.y);
                if (w <= 0.f) continue;
                size_t b = ((size_t)yy * W + xx) * 4;
                float dot = d_render[b] * color.x + d_render[b + 1] * color.y + d_render[b + 2] * color.z +
                            d_render[b + 3] * color.w;
                float denom = ws * ws;
                float d_w = denom > 0.f ? (dot * ws - w * dot * (ws - w)) / denom : 0.f;
                if (d_w != 0.f) d_compute_filter_weight(sc.filter_type, sc.filter_radius, xc - pt.x, yc - pt.y, d_w, d_radius);
            }
        }
}

// ---- backward/background.rs ------------------------------------------------
V4 sample_background(const Scene &sc, size_t pix) {  // background.rs:7-26
    if (sc.has_bg_image) {
        size_t b = pix * 4;
        if (b + 3 >= sc.bg_image.size()) return V4();
        float r = sc.bg_image[b], g = sc.bg_image[b + 1], bl = sc.bg_image[b + 2], a = sc.bg_image[b + 3];
        return V4(r * a, g * a, bl * a, a);
    }
    return V4(sc.background.x * sc.background.w, sc.background.y * sc.background.w, sc.background.z * sc.background.w,
              sc.background.w);
}

void accumulate_background_grad(const Scene &sc, Grads &gr, size_t pix, V4 d) {  // background.rs:64-91
    if (sc.has_bg_image) {
        if (!gr.d_background_image.empty()) {
            size_t b = pix * 4;
            if (b + 3 < gr.d_background_image.size()) {
                gr.d_background_image[b] += d.x; gr.d_background_image[b + 1] += d.y;
                gr.d_background_image[b + 2] += d.z; gr.d_background_image[b + 3] += d.w;
            }
        }
        return;
    }
    gr.d_background[0] += (double)d.x; gr.d_background[1] += (double)d.y;
    gr.d_background[2] += (double)d.z; gr.d_background[3] += (double)d.w;
}

void finalize_background_gradients(const Scene &sc, Grads &gr) {  // background.rs:29-61
    if (sc.has_bg_image) {
        for (size_t i = 0; i + 3 < gr.d_background_image.size(); i += 4) {
            float r = sc.bg_image[i], g = sc.bg_image[i + 1], b = sc.bg_image[i + 2], a = sc.bg_image[i + 3];
            float *c = &gr.d_background_image[i];
            float dr = c[0] * a, dg = c[1] * a, db = c[2] * a, da = c[0] * r + c[1] * g + c[2] * b + c[3];
            c[0] = dr; c[1] = dg; c[2] = db; c[3] = da;
        }
    } else {
        float r = sc.background.x, g = sc.background.y, b = sc.background.z, a = sc.background.w;
        float d[4] = {(float)gr.d_background[0], (float)gr.d_background[1], (float)gr.d_background[2],
                      (float)gr.d_background[3]};
        float nr = d[0] * a, ng = d[1] * a, nb = d[2] * a, na = d[0] * r + d[1] * g + d[2] * b + d[3];
        gr.d_background[0] = nr; gr.d_background[1] = ng; gr.d_background[2] = nb; gr.d_background[3] = na;
    }
}

void accumulate_translation(Grads &gr, long pix, V2 d) {  // sampling.rs:552-563
    if (pix < 0 || gr.d_translation.empty()) return;
    gr.d_translation[(size_t)pix * 2] += d.x;
    gr.d_translation[(size_t)pix * 2 + 1] += d.y;
}

// ---- backward/sampling.rs ----------------------------------------------------
struct Fragment {
    V4 color;  // rgb + alpha
    uint32_t group;
    bool stroke;
};

struct EdgeQuery {
    uint32_t group, shape;
    bool hit;
};

// within_distance_edge / is_inside_edge (sampling.rs:485-549)
static bool within_edge(const Ctx &cx, uint32_t gi, V2 pt, EdgeQuery *eq) {
    bool hit = within_distance_group(cx, gi, pt);
    if (!eq || eq->group != gi) return hit;
    const Group &g = cx.sc->groups[gi];
    if (eq->shape >= cx.sc->shapes.size()) return hit;
    const Shape &sh = cx.sc->shapes[eq->shape];
    if (sh.stroke_width <= 0.f) return hit;
    V2 lp = g.canvas_to_shape.transform_point(pt);
    if (within_distance_shape(cx, eq->shape, lp, sh.stroke_width)) eq->hit = true;
    return hit;
}

static bool inside_edge(const Ctx &cx, uint32_t gi, V2 pt, EdgeQuery *eq) {
    bool inside = is_inside_group(cx, gi, pt);
    if (!eq || eq->group != gi) return inside;
    const Group &g = cx.sc->groups[gi];
    if (eq->shape >= cx.sc->shapes.size()) return inside;
    V2 lp = g.canvas_to_shape.transform_point(pt);
    // transform_point_inverse of backward/math.rs:123 (no identity shortcut)
    V2 sp = cx.sc->shapes[eq->shape].transform.inverse_or_identity().transform_point(lp);
    int w = winding_number_shape(cx, eq->shape, sp);
    bool hit = (g.fill_rule == DVG_FILL_EVENODD) ? (std::abs(w) % 2) == 1 : w != 0;
    if (hit) eq->hit = true;
    return inside;
}

V4 sample_color(const Ctx &cx, V2 screen_pt, const V4 *bg, const V4 *d_color, EdgeQuery *eq, Grads *gr, long tidx,
                size_t pix) {
    const Scene &sc = *cx.sc;
    if (eq) eq->hit = false;
    V2 pt(screen_pt.x * (float)sc.width, screen_pt.y * (float)sc.height);
    std::vector<Fragment> frags;
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    for (uint32_t gi = 0; gi < sc.groups.size(); ++gi) {
        const Group &g = sc.groups[gi];
        count_gate(cx, counting, pt, gi);
        if (g.stroke.present() && within_edge(cx, gi, pt, eq)) {
            V4 c = paint_color(g.stroke, pt);
            frags.push_back({c, gi, true});
        }
        if (g.fill.present() && inside_edge(cx, gi, pt, eq)) {
            V4 c = paint_color(g.fill, pt);
            frags.push_back({c, gi, false});
        }
    }
    COUNT_ON(counting);
    if (frags.empty()) {
        if (bg) {
            if (d_color && gr) accumulate_background_grad(sc, *gr, pix, *d_color);
            return *bg;
        }
        return V4();
    }
    std::vector<V4> acc;
    acc.reserve(frags.size());
    float pr = 0.f, pg = 0.f, pb = 0.f, pa = 0.f;
    if (bg) { pr = bg->x; pg = bg->y; pb = bg->z; pa = bg->w; }
    for (const Fragment &f : frags) {
        float na = f.color.w;
        if (eq) {
            if (na >= 1.0f && eq->hit) eq->hit = false;
            if (eq->group == f.group) eq->hit = true;
        }
        float br = pr * (1.0f - na) + f.color.x * na;
        float bgc = pg * (1.0f - na) + f.color.y * na;
        float bb = pb * (1.0f - na) + f.color.z * na;
        float ba = pa * (1.0f - na) + na;
        acc.push_back(V4(br, bgc, bb, ba));
        pr = br; pg = bgc; pb = bb; pa = ba;
    }
    V4 fin = acc.back();
    float fa = fin.w;
    float fr = fin.x, fg = fin.y, fb = fin.z;
    if (fa > 1.0e-6f) {
        float s = 1.0f / fa;
        fr = fr * s; fg = fg * s; fb = fb * s;
    }
    if (d_color && gr) {
        float dr = d_color->x, dgc = d_color->y, db = d_color->z, da = d_color->w;
        if (fa > 1.0e-6f) {
            float s = 1.0f / fa;
            float ddot = dr * fr + dgc * fg + db * fb;
            dr = dr * s; dgc = dgc * s; db = db * s;
            da -= ddot / fa;
        }
        for (size_t ii = frags.size(); ii-- > 0;) {
            float pa2 = (ii > 0) ? acc[ii - 1].w : (bg ? bg->w : 0.f);
            V4 pc = (ii > 0) ? acc[ii - 1] : (bg ? *bg : V4());
            const Fragment &f = frags[ii];
            float d_prev_a = da * (1.0f - f.color.w);
            float d_alpha = da * (1.0f - pa2);
            d_alpha += dr * (f.color.x - pc.x) + dgc * (f.color.y - pc.y) + db * (f.color.z - pc.z);
            float dpr = dr * (1.0f - f.color.w), dpg = dgc * (1.0f - f.color.w), dpb = db * (1.0f - f.color.w);
            V4 dv(dr * f.color.w, dgc * f.color.w, db * f.color.w, d_alpha);
            V2 dtrans;
            const Group &g = sc.groups[f.group];
            float *slot = &gr->d_group_data[(size_t)f.group * 16 + (f.stroke ? 12 : 8)];
            d_sample_paint(f.stroke ? g.stroke : g.fill, *gr, slot, pt, dv, dtrans);
            accumulate_translation(*gr, tidx, dtrans);
            dr = dpr; dgc = dpg; db = dpb; da = d_prev_a;
        }
        if (bg) accumulate_background_grad(sc, *gr, pix, V4(dr, dgc, db, da));
    }
    return V4(fr, fg, fb, fa);
}

// ---- backward/boundary.rs ------------------------------------------------------
static float seg_len_polyline(const Seg &s) {  // boundary.rs:230-309 (per segment)
    if (s.kind == 0) return (s.p1 - s.p0).length();
    if (s.kind == 1) {
        auto ev = [&](float t) {
            float tt = 1.0f - t;
            return ((s.p0 * (tt * tt)) + (s.p1 * (2.0f * tt * t))) + (s.p2 * (t * t));
        };
        V2 v0 = s.p0, v1 = ev(0.5f), v2 = s.p2;
        return (v1 - v0).length() + (v2 - v1).length();
    }
    auto ev = [&](float t) {
        float tt = 1.0f - t;
        return (((s.p0 * (tt * tt * tt)) + (s.p1 * (3.0f * tt * tt * t))) + (s.p2 * (3.0f * tt * t * t))) +
               (s.p3 * (t * t * t));
    };
    V2 v0 = s.p0, v1 = ev(1.0f / 3.0f), v2 = ev(2.0f / 3.0f), v3 = s.p3;
    return ((v1 - v0).length() + (v2 - v1).length()) + (v3 - v2).length();
}

static std::vector<float> compute_shape_lengths(const Ctx &cx) {  // boundary.rs:200-227
    const Scene &sc = *cx.sc;
    std::vector<float> L(sc.shapes.size(), 0.f);
    for (size_t i = 0; i < sc.shapes.size(); ++i) {
        const Shape &sh = sc.shapes[i];
        float len = 0.f;
        if (sh.kind == DVG_SHAPE_CIRCLE) {
            len = 2.0f * PI_F * std::fabs(sh.radius_c);
        } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
            float a = std::fabs(sh.radius_e.x), b = std::fabs(sh.radius_e.y);
            if (a == 0.f || b == 0.f) len = 0.f;
            else len = PI_F * (3.0f * (a + b) - sqrt_f((3.0f * a + b) * (a + 3.0f * b)));
        } else if (sh.kind == DVG_SHAPE_RECT) {
            float w = std::fabs(sh.rmax.x - sh.rmin.x), h = std::fabs(sh.rmax.y - sh.rmin.y);
            len = 2.0f * (w + h);
        } else {
            // path_length walks segments with path_point (no radius); the
            // segment list of path_segments has the same extent.
            if (!sh.path.points.empty())
                for (const Seg &s : cx.segs[i]) len += seg_len_polyline(s);
        }
        L[i] = len;
    }
    return L;
}

struct ShapeCdf {
    std::vector<uint32_t> shape_ids, group_ids;
    std::vector<float> cdf, pmf;
};

static bool build_shape_cdf(const Scene &sc, const std::vector<float> &L, ShapeCdf &out) {  // boundary.rs:312-346
    float accum = 0.f;
    for (uint32_t gi = 0; gi < sc.groups.size(); ++gi)
        for (uint32_t si : sc.groups[gi].shape_indices) {
            float len = si < L.size() ? L[si] : 0.f;
            out.shape_ids.push_back(si);
            out.group_ids.push_back(gi);
            accum += len;
            out.cdf.push_back(accum);
            out.pmf.push_back(len);
        }
    if (accum <= 0.f || !std::isfinite(accum) || out.cdf.empty()) return false;
    for (float &v : out.cdf) v /= accum;
    for (float &v : out.pmf) v /= accum;
    return true;
}

struct PathCdf {
    bool valid = false;
    std::vector<float> cdf, pmf;
    std::vector<uint32_t> point_ids;
    float length = 0.f;
};

static std::vector<PathCdf> build_path_cdfs(const Ctx &cx, const std::vector<float> &L) {  // boundary.rs:349-451
    const Scene &sc = *cx.sc;
    std::vector<PathCdf> out(sc.shapes.size());
    for (size_t i = 0; i < sc.shapes.size(); ++i) {
        const Shape &sh = sc.shapes[i];
        if (sh.kind != DVG_SHAPE_PATH) continue;
        float len = L[i];
        if (len <= 0.f || !std::isfinite(len)) continue;
        size_t nb = sh.path.num_control_points.size();
        if (nb == 0) continue;
        float inv = 1.0f / len;
        PathCdf pc;
        pc.valid = true;
        pc.length = len;
        // Every base segment gets an entry; a segment with missing points has
        // length 0 (boundary.rs:372-441).
        size_t pid = 0;
        const std::vector<Seg> &segs = cx.segs[i];
        for (size_t k = 0; k < nb; ++k) {
            uint8_t nc = sh.path.num_control_points[k];
            pc.point_ids.push_back((uint32_t)pid);
            float sl = 0.f;
            if (k < segs.size() && nc <= 2) sl = seg_len_polyline(segs[k]);
            else if (nc <= 2) {
                // points missing: path_point lookups fail -> 0
                Seg s;
                s.kind = nc;
                V2 a, b, c, d;
                bool ok = path_point(sh.path, pid, a) && path_point(sh.path, pid + 1, b);
                if (ok && nc >= 1) ok = path_point(sh.path, pid + 2, c);
                if (ok && nc == 2) ok = path_point(sh.path, pid + 3, d);
                if (ok) {
                    s.p0 = a; s.p1 = b; s.p2 = (nc >= 1) ? c : b; s.p3 = (nc == 2) ? d : s.p2;
                    sl = seg_len_polyline(s);
                }
            }
            if (nc == 0) pid += 1;
            else if (nc == 1) pid += 2;
            else if (nc == 2) pid += 3;
            float sn = sl * inv;
            pc.pmf.push_back(sn);
            float prev = pc.cdf.empty() ? 0.f : pc.cdf.back();
            pc.cdf.push_back(sn + prev);
        }
        out[i] = pc;
    }
    return out;
}

static void sample_cdf(const std::vector<float> &cdf, float u, size_t &idx_out, float &t_out) {  // boundary.rs:454-469
    if (cdf.empty()) { idx_out = 0; t_out = 0.f; return; }
    size_t idx = 0;
    while (idx < cdf.size() && u > cdf[idx]) ++idx;
    if (idx >= cdf.size()) idx = cdf.size() - 1;
    float prev = idx == 0 ? 0.f : cdf[idx - 1];
    float denom = rmax(cdf[idx] - prev, 1.0e-6f);
    t_out = rclamp((u - prev) / denom, 0.f, 1.f);
    idx_out = idx;
}

V2 normalize(V2 v) {  // backward/math.rs:44-51
    float len = v.length();
    if (len == 0.f) return V2();
    return v / len;
}

struct BData {
    uint32_t base_point_id = 0, point_id = 0;
    float t = 0.f;
    bool is_stroke = false;
};

static V2 sample_boundary_path(const Path &path, const PathCdf &pc, float path_length, float t, V2 &normal,
                               float &pdf, BData &data, float dir, float stroke_radius) {  // boundary.rs:586-843
    size_t np = path.points.size(), nb = path.num_control_points.size();
    if (np == 0 || nb == 0) { pdf = 0.f; return V2(); }
    if (dir != 0.f && !path.is_closed) {
        float cap_length = 0.f;
        if (path.has_thickness) {
            if (!path.thickness.empty()) cap_length = PI_F * (path.thickness[0] + path.thickness.back());
        } else {
            cap_length = 2.0f * PI_F * stroke_radius;
        }
        float denom = cap_length + path_length;
        if (denom > 0.f) {
            float cap_prob = cap_length / denom;
            if (t < cap_prob) {
                t /= cap_prob;
                pdf *= cap_prob;
                float r0 = stroke_radius, r1 = stroke_radius;
                if (path.has_thickness && !path.thickness.empty()) {
                    r0 = path.thickness[0];
                    r1 = path.thickness.back();
                }
                float angle = 2.0f * PI_F * t;
                if (dir < 0.f) {
                    V2 p0 = path.points[0];
                    V2 off(r0 * cos_f(angle), r0 * sin_f(angle));
                    normal = normalize(off);
                    pdf /= 2.0f * PI_F * r0;
                    data.base_point_id = 0; data.point_id = 0; data.t = 0.f;
                    return p0 + off;
                } else {
                    V2 p0 = path.points[np - 1];
                    V2 off(r1 * cos_f(angle), r1 * sin_f(angle));
                    normal = normalize(off);
                    pdf /= 2.0f * PI_F * r1;
                    data.base_point_id = (uint32_t)(nb - 1);
                    size_t controls = path.num_control_points[nb - 1];
                    size_t sub = 2 + controls;
                    data.point_id = (uint32_t)(np >= sub ? np - sub : 0);
                    data.t = 1.f;
                    return p0 + off;
                }
            } else {
                t = (t - cap_prob) / (1.0f - cap_prob);
                pdf *= 1.0f - cap_prob;
            }
        }
    }
    size_t sid;
    float lt;
    sample_cdf(pc.cdf, t, sid, lt);
    if (sid >= pc.point_ids.size()) { pdf = 0.f; return V2(); }
    size_t pid = pc.point_ids[sid];
    data.base_point_id = (uint32_t)sid;
    data.point_id = (uint32_t)pid;
    data.t = lt;
    if (lt < -1.0e-3f || lt > 1.0f + 1.0e-3f) { pdf = 0.f; return V2(); }
    auto next_index = [&](size_t idx) -> size_t {
        if (path.is_closed) return idx % np;
        return std::min(idx, np ? np - 1 : 0);
    };
    uint8_t nc = sid < nb ? path.num_control_points[sid] : 0;
    float pmf = sid < pc.pmf.size() ? pc.pmf[sid] : 0.f;
    if (nc == 0) {
        V2 p0, p1;
        float r0, r1;
        if (!path_point_radius(path, pid, stroke_radius, p0, r0) ||
            !path_point_radius(path, next_index(pid + 1), stroke_radius, p1, r1)) { pdf = 0.f; return V2(); }
        V2 tan = p1 - p0;
        float tl = tan.length();
        if (tl == 0.f) { pdf = 0.f; return V2(); }
        normal = V2(-tan.y, tan.x) / tl;
        pdf *= pmf / tl;
        V2 out = p0 + tan * lt;
        if (dir != 0.f) {
            float r = r0 + lt * (r1 - r0);
            out += normal * (dir * r);
            if (dir < 0.f) normal = normal * -1.0f;
        }
        return out;
    }
    if (nc == 1) {
        V2 p0, p1, p2;
        float r0, r1, r2;
        if (!path_point_radius(path, pid, stroke_radius, p0, r0) ||
            !path_point_radius(path, pid + 1, stroke_radius, p1, r1) ||
            !path_point_radius(path, next_index(pid + 2), stroke_radius, p2, r2)) { pdf = 0.f; return V2(); }
        float tt = 1.0f - lt;
        V2 ev = ((p0 * (tt * tt)) + (p1 * (2.0f * tt * lt))) + (p2 * (lt * lt));
        V2 tan = ((p1 - p0) * (2.0f * tt)) + ((p2 - p1) * (2.0f * lt));
        float tl = tan.length();
        if (tl == 0.f) { pdf = 0.f; return V2(); }
        normal = V2(-tan.y, tan.x) / tl;
        pdf *= pmf / tl;
        V2 out = ev;
        if (dir != 0.f) {
            float r = ((r0 * (tt * tt)) + (r1 * (2.0f * tt * lt))) + (r2 * (lt * lt));
            out += normal * (dir * r);
            if (dir < 0.f) normal = normal * -1.0f;
        }
        return out;
    }
    if (nc == 2) {
        V2 p0, p1, p2, p3;
        float r0, r1, r2, r3;
        if (!path_point_radius(path, pid, stroke_radius, p0, r0) ||
            !path_point_radius(path, pid + 1, stroke_radius, p1, r1) ||
            !path_point_radius(path, pid + 2, stroke_radius, p2, r2) ||
            !path_point_radius(path, next_index(pid + 3), stroke_radius, p3, r3)) { pdf = 0.f; return V2(); }
        float tt = 1.0f - lt;
        V2 ev = (((p0 * (tt * tt * tt)) + (p1 * (3.0f * tt * tt * lt))) + (p2 * (3.0f * tt * lt * lt))) +
                (p3 * (lt * lt * lt));
        V2 tan = (((p1 - p0) * (3.0f * tt * tt)) + ((p2 - p1) * (6.0f * tt * lt))) + ((p3 - p2) * (3.0f * lt * lt));
        float tl = tan.length();
        if (tl == 0.f) { pdf = 0.f; return V2(); }
        normal = V2(-tan.y, tan.x) / tl;
        pdf *= pmf / tl;
        V2 out = ev;
        if (dir != 0.f) {
            float r = (((r0 * (tt * tt * tt)) + (r1 * (3.0f * tt * tt * lt))) + (r2 * (3.0f * tt * lt * lt))) +
                      (r3 * (lt * lt * lt));
            out += normal * (dir * r);
            if (dir < 0.f) normal = normal * -1.0f;
        }
        return out;
    }
    pdf = 0.f;
    return V2();
}

static bool sample_boundary_point(const Scene &sc, uint32_t si, uint32_t gi, float t, float shape_length,
                                  const PathCdf *pc, V2 &pt_out, V2 &normal, float &pdf,
                                  BData &data) {  // boundary.rs:472-583
    const Group &g = sc.groups[gi];
    const Shape &sh = sc.shapes[si];
    if (!g.fill.present() && !g.stroke.present()) return false;
    pdf = 1.0f;
    float lt = t;
    bool stroke = false;
    if (g.fill.present() && g.stroke.present()) {
        if (lt < 0.5f) { stroke = false; lt *= 2.0f; pdf = 0.5f; }
        else { stroke = true; lt = 2.0f * (lt - 0.5f); pdf = 0.5f; }
    } else if (g.stroke.present()) {
        stroke = true;
    }
    float dir = 0.f;
    if (stroke) {
        if (lt < 0.5f) { dir = -1.0f; lt *= 2.0f; pdf *= 0.5f; }
        else { dir = 1.0f; lt = 2.0f * (lt - 0.5f); pdf *= 0.5f; }
    }
    data = BData();
    data.is_stroke = stroke;
    normal = V2();
    V2 pt;
    if (sh.kind == DVG_SHAPE_CIRCLE) {
        float r = std::fabs(sh.radius_c);
        if (r <= 0.f) return false;
        float angle = 2.0f * PI_F * lt;
        V2 off(r * cos_f(angle), r * sin_f(angle));
        normal = normalize(off);
        pdf /= 2.0f * PI_F * r;
        pt = sh.center + off;
        if (dir != 0.f) {
            pt += normal * (dir * sh.stroke_width);
            if (dir < 0.f) normal = normal * -1.0f;
        }
    } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
        float rx = std::fabs(sh.radius_e.x), ry = std::fabs(sh.radius_e.y);
        if (rx <= 0.f || ry <= 0.f) return false;
        float angle = 2.0f * PI_F * lt;
        float s = sin_f(angle), c = cos_f(angle);
        V2 off(rx * c, ry * s);
        float dxdt = -rx * s * 2.0f * PI_F;
        float dydt = ry * c * 2.0f * PI_F;
        normal = normalize(V2(dydt, -dxdt));
        pdf /= sqrt_f(dxdt * dxdt + dydt * dydt);
        pt = sh.center + off;
        if (dir != 0.f) {
            pt += normal * (dir * sh.stroke_width);
            if (dir < 0.f) normal = normal * -1.0f;
        }
    } else if (sh.kind == DVG_SHAPE_RECT) {
        V2 mn = sh.rmin, mx = sh.rmax;
        float w = mx.x - mn.x, h = mx.y - mn.y;
        if (w == 0.f && h == 0.f) return false;
        pdf /= 2.0f * (w + h);
        V2 out;
        if (lt <= w / (w + h)) {
            lt *= (w + h) / w;
            if (lt < 0.5f) { normal = V2(0.f, -1.f); out = mn + V2(2.0f * lt * (mx.x - mn.x), 0.f); }
            else { normal = V2(0.f, 1.f); out = V2(mn.x, mx.y) + V2(2.0f * (lt - 0.5f) * (mx.x - mn.x), 0.f); }
        } else {
            lt = (lt - w / (w + h)) * (w + h) / h;
            if (lt < 0.5f) { normal = V2(-1.f, 0.f); out = mn + V2(0.f, 2.0f * lt * (mx.y - mn.y)); }
            else { normal = V2(1.f, 0.f); out = V2(mx.x, mn.y) + V2(0.f, 2.0f * (lt - 0.5f) * (mx.y - mn.y)); }
        }
        if (dir != 0.f) {
            out += normal * (dir * sh.stroke_width);
            if (dir < 0.f) normal = normal * -1.0f;
        }
        pt = out;
    } else {
        if (!pc || !pc->valid) return false;
        if (pc->length <= 0.f) return false;
        pt = sample_boundary_path(sh.path, *pc, shape_length, lt, normal, pdf, data, dir, sh.stroke_width);
    }
    if (!std::isfinite(pdf) || pdf <= 0.f) return false;
    pt_out = pt;
    return true;
}

void d_xform_pt(const M3 &m, V2 pt, V2 d_out, M3 &d_m, V2 &d_pt) {  // backward/math.rs:101-120
    float t0 = m.m[0][0] * pt.x + m.m[0][1] * pt.y + m.m[0][2];
    float t1 = m.m[1][0] * pt.x + m.m[1][1] * pt.y + m.m[1][2];
    float t2 = m.m[2][0] * pt.x + m.m[2][1] * pt.y + m.m[2][2];
    V2 out(t0 / t2, t1 / t2);
    float d_t0 = d_out.x / t2, d_t1 = d_out.y / t2;
    float d_t2 = -(d_out.x * out.x + d_out.y * out.y) / t2;
    d_m.m[0][0] += d_t0 * pt.x; d_m.m[0][1] += d_t0 * pt.y; d_m.m[0][2] += d_t0;
    d_m.m[1][0] += d_t1 * pt.x; d_m.m[1][1] += d_t1 * pt.y; d_m.m[1][2] += d_t1;
    d_m.m[2][0] += d_t2 * pt.x; d_m.m[2][1] += d_t2 * pt.y; d_m.m[2][2] += d_t2;
    d_pt.x += d_t0 * m.m[0][0] + d_t1 * m.m[1][0] + d_t2 * m.m[2][0];
    d_pt.y += d_t0 * m.m[0][1] + d_t1 * m.m[1][1] + d_t2 * m.m[2][1];
}

static V2 xform_normal(const M3 &mi, V2 n) {  // backward/math.rs:128-133
    float x = mi.m[0][0] * n.x + mi.m[1][0] * n.y;
    float y = mi.m[0][1] * n.x + mi.m[1][1] * n.y;
    return normalize(V2(x, y));
}

static void accumulate_boundary_gradient(const Scene &sc, uint32_t si, uint32_t gi, float contrib, float t, V2 normal,
                                         const BData &bd, V2 local_pt, Grads &gr) {  // boundary.rs:970-1148
    if (!std::isfinite(contrib)) return;
    const Shape &sh = sc.shapes[si];
    const Group &g = sc.groups[gi];
    if (bd.is_stroke) {
        bool has_thickness = sh.kind == DVG_SHAPE_PATH && sh.path.has_thickness;
        if (has_thickness) {
            const Path &path = sh.path;
            size_t nth = path.thickness.size();
            uint32_t to = path.thickness_offset;
            uint8_t nc = bd.base_point_id < path.num_control_points.size() ? path.num_control_points[bd.base_point_id] : 0;
            size_t pid = bd.point_id, np = path.points.size();
            float tt = 1.0f - bd.t, tv = bd.t;
            auto add = [&](size_t i, float v) { if (i < nth) gr.d_shape_thickness[to + i] += v; };
            if (nc == 0) {
                size_t i1 = path.is_closed ? (pid + 1) % np : pid + 1;
                add(pid, (1.0f - tv) * contrib);
                add(i1, tv * contrib);
            } else if (nc == 1) {
                size_t i2 = path.is_closed ? (pid + 2) % np : pid + 2;
                add(pid, tt * tt * contrib);
                add(pid + 1, 2.0f * tt * tv * contrib);
                add(i2, tv * tv * contrib);
            } else if (nc == 2) {
                size_t i3 = path.is_closed ? (pid + 3) % np : pid + 3;
                add(pid, tt * tt * tt * contrib);
                add(pid + 1, 3.0f * tt * tt * tv * contrib);
                add(pid + 2, 3.0f * tt * tv * tv * contrib);
                add(i3, tv * tv * tv * contrib);
            }
        } else {
            gr.d_shape_stroke_width[si] += contrib;
        }
    }
    float *dp = &gr.d_shape_params[(size_t)si * 8];
    if (sh.kind == DVG_SHAPE_CIRCLE) {
        dp[0] += normal.x * contrib; dp[1] += normal.y * contrib;
        dp[2] += contrib;
    } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
        dp[0] += normal.x * contrib; dp[1] += normal.y * contrib;
        float angle = 2.0f * PI_F * t;
        dp[2] += cos_f(angle) * normal.x * contrib;
        dp[3] += sin_f(angle) * normal.y * contrib;
    } else if (sh.kind == DVG_SHAPE_PATH) {
        const Path &path = sh.path;
        size_t np = path.points.size();
        uint8_t nc = bd.base_point_id < path.num_control_points.size() ? path.num_control_points[bd.base_point_id] : 0;
        size_t pid = bd.point_id;
        float tv = bd.t, tt = 1.0f - tv;
        auto add = [&](size_t idx, float w) {
            size_t index = path.is_closed ? idx % np : idx;
            if (index < np) {
                float *q = &gr.d_shape_points[((size_t)path.point_offset + index) * 2];
                V2 v = normal * (w * contrib);
                q[0] += v.x; q[1] += v.y;
            }
        };
        if (nc == 0) {
            add(pid, 1.0f - tv);
            add(pid + 1, tv);
        } else if (nc == 1) {
            add(pid, tt * tt);
            add(pid + 1, 2.0f * tt * tv);
            add(pid + 2, tv * tv);
        } else if (nc == 2) {
            add(pid, tt * tt * tt);
            add(pid + 1, 3.0f * tt * tt * tv);
            add(pid + 2, 3.0f * tt * tv * tv);
            add(pid + 3, tv * tv * tv);
        }
    } else if (sh.kind == DVG_SHAPE_RECT) {
        if (normal == V2(-1.f, 0.f)) dp[0] += -contrib;
        else if (normal == V2(1.f, 0.f)) dp[2] += contrib;
        else if (normal == V2(0.f, -1.f)) dp[1] += -contrib;
        else if (normal == V2(0.f, 1.f)) dp[3] += contrib;
    }
    M3 s2c = g.shape_to_canvas.mul(sh.transform);
    M3 dm = M3::zero();
    V2 dlp;
    d_xform_pt(s2c, local_pt, normal * contrib, dm, dlp);
    M3 dgm = dm.mul(sh.transform.transpose());
    M3 dsm = g.shape_to_canvas.transpose().mul(dm);
    float *dgt = &gr.d_group_transform[(size_t)gi * 9];
    float *dst = &gr.d_shape_transform[(size_t)si * 9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            dgt[i * 3 + j] += dgm.m[i][j];
            dst[i * 3 + j] += dsm.m[i][j];
        }
}

void boundary_sampling(const Ctx &cx, uint32_t nsx, uint32_t nsy, uint32_t seed, const float *d_render,
                       const std::vector<float> &wimg, Grads &gr, bool compute_translation, size_t i0,
                       size_t i1, long row_lo, long row_hi) {  // boundary.rs:113-197, over sample indices [i0, i1)
    const Scene &sc = *cx.sc;
    if (sc.groups.empty()) return;
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    COUNT_ON(false);  // the sampling tables are per step, not per sample
    std::vector<float> L = compute_shape_lengths(cx);
    ShapeCdf cdf;
    bool have_cdf = build_shape_cdf(sc, L, cdf);
    std::vector<PathCdf> pcs;
    if (have_cdf) pcs = build_path_cdfs(cx, L);
    COUNT_ON(counting);
    if (!have_cdf) return;
    size_t total = (size_t)sc.width * sc.height * nsx * nsy;
    if (i1 > total) i1 = total;
    int W = (int)sc.width, H = (int)sc.height;
    for (size_t idx = i0; idx < i1; ++idx) {
        Pcg32 rng(idx, seed);
        float u = rng.next_f32();
        float t = rng.next_f32();
        size_t sid;
        float unused;
        sample_cdf(cdf.cdf, u, sid, unused);
        if (sid >= cdf.shape_ids.size()) continue;
        uint32_t si = cdf.shape_ids[sid], gi = cdf.group_ids[sid];
        float spmf = cdf.pmf[sid];
        if (spmf <= 0.f) continue;
        float slen = si < L.size() ? L[si] : 0.f;
        const PathCdf *pc = (si < pcs.size()) ? &pcs[si] : nullptr;
        V2 local_pt, normal;
        float bpdf;
        BData bd;
        if (!sample_boundary_point(sc, si, gi, t, slen, pc, local_pt, normal, bpdf, bd)) continue;
        if (bpdf <= 0.f) continue;
        const Group &g = sc.groups[gi];
        const Shape &sh = sc.shapes[si];
        M3 s2c = g.shape_to_canvas.mul(sh.transform);
        V2 bpt = s2c.transform_point(local_pt);
        M3 sinv = sh.transform.inverse_or_identity();
        M3 cinv = sinv.mul(g.canvas_to_shape);
        V2 ncan = xform_normal(cinv, normal);
        bpt.x /= (float)sc.width;
        bpt.y /= (float)sc.height;
        float pdf = spmf * bpdf;
        // render_edge_sample, boundary.rs:845-967
        if (W <= 0 || H <= 0 || pdf <= 0.f) continue;
        int bx = (int)(bpt.x * (float)W), by = (int)(bpt.y * (float)H);
        if (bx < 0 || bx >= W || by < 0 || by >= H) continue;
        if (by < row_lo || by >= row_hi) continue;
        size_t pix = (size_t)by * W + bx;
        V4 bg = sample_background(sc, pix);
        EdgeQuery iq{gi, si, false}, oq{gi, si, false};
        V2 n = ncan;
        V4 cin = sample_color(cx, bpt - n * 1.0e-4f, &bg, nullptr, &iq, nullptr, -1, pix);
        V4 cout = sample_color(cx, bpt + n * 1.0e-4f, &bg, nullptr, &oq, nullptr, -1, pix);
        if (!iq.hit && !oq.hit) continue;
        if (!iq.hit) {
            n = n * -1.0f;
            std::swap(cin, cout);
        }
        V2 sbp(bpt.x * (float)W, bpt.y * (float)H);
        V4 dc = gather_d_color(sc, d_render, wimg, sbp);
        float norm = (float)sc.width * (float)sc.height;
        if (norm > 0.f) { dc.x /= norm; dc.y /= norm; dc.z /= norm; dc.w /= norm; }
        float contrib = ((cin.x - cout.x) * dc.x + (cin.y - cout.y) * dc.y + (cin.z - cout.z) * dc.z +
                         (cin.w - cout.w) * dc.w) / pdf;
        if (!std::isfinite(contrib)) continue;
        accumulate_boundary_gradient(sc, si, gi, contrib, t, n, bd, local_pt, gr);
        if (compute_translation && !gr.d_translation.empty()) {
            size_t i2 = pix * 2;
            if (i2 + 1 < gr.d_translation.size()) {
                gr.d_translation[i2] += n.x * contrib;
                gr.d_translation[i2 + 1] += n.y * contrib;
            }
        }
    }
}

// Canvas positions of the two colour queries (boundary.rs:880-901) of every
// contributing boundary sample in [i0, i1); for operation counting.
void boundary_query_points(const Ctx &cx, uint32_t nsx, uint32_t nsy, uint32_t seed, size_t i0, size_t i1,
                           std::vector<float> &pts) {
    const Scene &sc = *cx.sc;
    pts.clear();
    if (sc.groups.empty()) return;
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    COUNT_ON(false);  // the sampling tables are per step, not per sample
    std::vector<float> L = compute_shape_lengths(cx);
    ShapeCdf cdf;
    bool have_cdf = build_shape_cdf(sc, L, cdf);
    std::vector<PathCdf> pcs;
    if (have_cdf) pcs = build_path_cdfs(cx, L);
    COUNT_ON(counting);
    if (!have_cdf) return;
    size_t total = (size_t)sc.width * sc.height * nsx * nsy;
    if (i1 > total) i1 = total;
    for (size_t idx = i0; idx < i1; ++idx) {
        Pcg32 rng(idx, seed);
        float u = rng.next_f32();
        float t = rng.next_f32();
        size_t sid;
        float unused;
        sample_cdf(cdf.cdf, u, sid, unused);
        if (sid >= cdf.shape_ids.size()) continue;
        uint32_t si = cdf.shape_ids[sid], gi = cdf.group_ids[sid];
        if (cdf.pmf[sid] <= 0.f) continue;
        V2 lp, nrm;
        float bpdf;
        BData bd;
        if (!sample_boundary_point(sc, si, gi, t, L[si], &pcs[si], lp, nrm, bpdf, bd)) continue;
        const Group &g = sc.groups[gi];
        const Shape &sh = sc.shapes[si];
        V2 bpt = g.shape_to_canvas.mul(sh.transform).transform_point(lp);
        V2 n = xform_normal(sh.transform.inverse_or_identity().mul(g.canvas_to_shape), nrm);
        float qx = bpt.x / (float)sc.width, qy = bpt.y / (float)sc.height;
        int bx = (int)(qx * (float)sc.width), by = (int)(qy * (float)sc.height);
        if (bx < 0 || by < 0 || bx >= (int)sc.width || by >= (int)sc.height) continue;
        for (float sgn : {-1.0f, 1.0f}) {
            pts.push_back((qx + sgn * n.x * 1.0e-4f) * (float)sc.width);
            pts.push_back((qy + sgn * n.y * 1.0e-4f) * (float)sc.height);
        }
    }
}

}  // namespace cpu

// The per-sample loop of render_backward (backward/render.rs:95-189) over
// sample indices [i0, i1).
void backward_samples(const Ctx &cx, const dvg_render_options &opt, bool compute_translation, const float *d_render,
                      const std::vector<float> &wimg, Grads &gr, size_t i0, size_t i1) {
    using namespace cpu;
    const Scene &sc = *cx.sc;
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    size_t W = sc.width, H = sc.height;
    size_t total = W * H * nsx * nsy;
    if (i1 > total) i1 = total;
    bool jitter = opt.jitter != 0 && !opt.use_prefiltering;
    for (size_t idx = i0; idx < i1; ++idx) {
        size_t sx = idx % nsx, sy = (idx / nsx) % nsy, x = (idx / (nsx * nsy)) % W, y = idx / ((size_t)nsx * nsy * W);
        float rx = 0.5f, ry = 0.5f;
        if (jitter) {
            Pcg32 rng(idx, opt.seed);
            rx = rng.next_f32();
            ry = rng.next_f32();
        }
        float px = (float)x + ((float)sx + rx) / (float)nsx;
        float py = (float)y + ((float)sy + ry) / (float)nsy;
        V2 pt(px, py);
        V2 npt(px / (float)sc.width, py / (float)sc.height);
        size_t pix = y * W + x;
        V4 bg = sample_background(sc, pix);
        V4 dc = gather_d_color(sc, d_render, wimg, pt);
        long tidx = compute_translation ? (long)pix : -1;
        V4 color = opt.use_prefiltering ? sample_color_prefiltered(cx, npt, &bg, &dc, &gr, tidx, pix)
                                        : sample_color(cx, npt, &bg, &dc, nullptr, &gr, tidx, pix);
        float d_radius = 0.f;
        accumulate_filter_gradient(sc, color, d_render, wimg, pt, d_radius);
        gr.d_filter_radius[0] += (double)d_radius;
    }
}

void make_ctx(const Scene &sc, Ctx &cx) {
    cx.sc = &sc;
    cx.segs.clear();
    for (const Shape &sh : sc.shapes)
        cx.segs.push_back(sh.kind == DVG_SHAPE_PATH ? path_segments(sh.path, sh.stroke_width) : std::vector<Seg>());
    cpu::build_group_bounds(cx);
}

std::vector<float> weight_image(const Scene &sc, const dvg_render_options &opt) {
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    bool prefilter = opt.use_prefiltering != 0;
    return cpu::build_weight_image(sc, nsx, nsy, opt.seed, prefilter, opt.jitter != 0 && !prefilter);
}

// render_backward, backward/render.rs:35-210 (d_render present, d_sdf absent).
int render_backward(const Scene &sc, const dvg_render_options &opt, bool compute_translation, const float *d_render,
                    Grads &gr, const char **err, uint32_t y0, uint32_t y1) {
    using namespace cpu;
    if (sc.has_bg_image && sc.bg_image.size() != (size_t)sc.width * sc.height * 4) {
        *err = "background image size mismatch";
        return DVG_INVALID_SCENE;
    }
    if (sc.width == 0 || sc.height == 0) return DVG_OK;
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    size_t total = (size_t)sc.width * sc.height * nsx * nsy;
    Ctx cx;
    make_ctx(sc, cx);
    std::vector<float> wimg = weight_image(sc, opt);
    long ri = (long)rmax(std::ceil(sc.filter_radius), 0.0f);
    size_t row = (size_t)sc.width * nsx * nsy;  // samples of a pixel row
    y1 = std::min(y1, sc.height);
    size_t ys = (size_t)std::max(0L, (long)y0 - ri), ye = std::min<size_t>(sc.height, (size_t)y1 + ri);
    if (ys < ye) backward_samples(cx, opt, compute_translation, d_render, wimg, gr, ys * row, ye * row);
    // boundary sampling only without prefiltering (render.rs:192-206)
    if (!opt.use_prefiltering)
        boundary_sampling(cx, nsx, nsy, opt.seed, d_render, wimg, gr, compute_translation, 0, total,
                          (long)y0 - ri - 1, (long)y1 + ri + 1);
    finalize_background_gradients(sc, gr);
    return DVG_OK;
}

}  // namespace oracle
