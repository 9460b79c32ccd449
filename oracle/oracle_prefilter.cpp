// Oracle (test infrastructure only): restatement of the reference CPU
// prefiltered (analytic-coverage) sample colour and its adjoint.
//   sample_color_prefiltered        src/backward/sampling.rs:227-482
//   compute_distance_bvh            src/distance/api.rs:138-232
//   closest_point_shape / _path_bvh src/distance/closest.rs:100-221, 364-456
//   d_compute_distance              src/backward/distance.rs:50-126
//   d_closest_point and its shapes  src/backward/distance.rs:128-561
//   smoothstep / d_smoothstep, d_distance, d_normalize
//                                   src/backward/math.rs:9-70
//
// Visiting order: the reference walks a group BVH (shapes) and a path BVH
// (segments).  For groups of <= 8 shapes and paths of <= 8 segments the BVH is
// one leaf visited in index order, which is the order used here; for larger
// ones the BVH order differs only on exact distance ties (the group query
// keeps the last shape at the minimum, `<=`; the path query the first
// segment, `<`).
#include "oracle_backward.h"

namespace oracle {
namespace cpu {

static float smoothstep(float d) {  // backward/math.rs:9-12
    float t = rclamp((d + 1.0f) * 0.5f, 0.f, 1.f);
    return t * t * (3.0f - 2.0f * t);
}

static float d_smoothstep(float d, float d_ret) {  // backward/math.rs:15-22
    if (d < -1.0f || d > 1.0f) return 0.f;
    float t = (d + 1.0f) * 0.5f;
    float d_t = d_ret * (6.0f * t - 6.0f * t * t);
    return d_t * 0.5f;
}

static V2 d_normalize(V2 v, V2 d_n) {  // backward/math.rs:54-64
    float l = v.length();
    if (l == 0.f) return V2();
    V2 n = v / l;
    V2 d_v = d_n / l;
    float d_l = -d_n.dot(n) / l;
    d_v += d_length(v, d_l);
    return d_v;
}

static M3 m3_scale(const M3 &m, float s) {
    M3 r = m;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] *= s;
    return r;
}

static void add_m3(float *dst, const M3 &m) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) dst[i * 3 + j] += m.m[i][j];
}

// Result of a closest-point query on a group (distance/types.rs:13-27).
struct Hit {
    uint32_t shape = 0;  // scene shape index
    V2 point;            // canvas
    float distance = 0.f;
    bool path = false;   // path metadata present
    uint32_t seg = 0;    // segment index in num_control_points
    float t = 0.f;
};

static float bounds_distance(const float *b, V2 p) {  // distance/bvh.rs:104-108
    float dx = rmax(rmax(b[0] - p.x, 0.f), p.x - b[2]);
    float dy = rmax(rmax(b[1] - p.y, 0.f), p.y - b[3]);
    return sqrt_f(dx * dx + dy * dy);
}

// Closest point of shape si to the group-local point lp, returned group-local
// (closest_point_shape, closest.rs:364-456, and its path-BVH twin,
// api.rs:185-192 over closest.rs:100-221).
static bool closest_local(const Ctx &cx, uint32_t si, V2 lp, V2 &out, Hit &h) {
    const Shape &sh = cx.sc->shapes[si];
    V2 sp = transform_point_inverse(sh.transform, lp);
    V2 c;
    h.path = false;
    h.seg = 0;
    h.t = 0.f;
    if (sh.kind == DVG_SHAPE_CIRCLE) {
        c = closest_point_circle(sh.center, std::fabs(sh.radius_c), sp);
    } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
        c = closest_point_ellipse(sh.center, V2(std::fabs(sh.radius_e.x), std::fabs(sh.radius_e.y)), sp);
    } else if (sh.kind == DVG_SHAPE_RECT) {
        c = closest_point_rect(sh.rmin, sh.rmax, sp);
    } else {
        float best = INFINITY;
        for (const Seg &s : cx.segs[si]) {
            V2 cp;
            float t, d;
            closest_on_seg(s, sp, sh.path.use_distance_approx, cp, t, d);
            if (d < best) {
                best = d;
                c = cp;
                h.seg = s.index;
                h.t = t;
            }
        }
        if (!std::isfinite(best)) return false;
        h.path = true;
    }
    out = sh.transform.transform_point(c);
    return true;
}

// compute_distance_bvh (api.rs:138-232): the closest boundary point of group
// gi to the canvas point pt within max_radius.
static bool compute_distance(const Ctx &cx, uint32_t gi, V2 pt, float max_radius, Hit &hit) {
    const Scene &sc = *cx.sc;
    const Group &g = sc.groups[gi];
    V2 lp = g.canvas_to_shape.transform_point(pt);
    bool bvh = !cx.group_has_bounds.empty() && cx.group_has_bounds[gi];
    // root node: group-local box distance against the canvas radius
    if (bvh && bounds_distance(&cx.group_bounds[(size_t)gi * 4], lp) > max_radius) return false;
    float best = max_radius;
    bool found = false;
    for (uint32_t si : g.shape_indices) {
        if (si >= sc.shapes.size()) continue;
        float b[4];
        if (bvh && !cpu_shape_bounds(sc.shapes[si], b)) continue;  // not in the BVH (bvh.rs:148-160)
        V2 lc;
        Hit h;
        if (!closest_local(cx, si, lp, lc, h)) continue;
        V2 c = g.shape_to_canvas.transform_point(lc);
        float d = (c - pt).length();
        bool take = bvh ? d <= best : (d <= max_radius && (!found || d < best));  // api.rs:206 / :60-63
        if (take) {
            best = d;
            h.shape = si;
            h.point = c;
            h.distance = d;
            hit = h;
            found = true;
        }
    }
    return found;
}

static size_t path_point_id(const Path &path, size_t base) {  // backward/distance.rs:36-47
    size_t id = 0, n = std::min(base, path.num_control_points.size());
    for (size_t i = 0; i < n; ++i) {
        uint8_t c = path.num_control_points[i];
        id += c == 0 ? 1 : (c == 1 ? 2 : (c == 2 ? 3 : 0));
    }
    return id;
}

// d_closest_point_path (backward/distance.rs:188-404); pt is path-local.
static void d_closest_point_path(const Path &path, V2 pt, V2 d_cp, size_t base, float min_t, Grads &gr, V2 &d_pt) {
    size_t pid = path_point_id(path, base);
    size_t n = path.points.size();
    if (n == 0) return;
    auto get = [&](size_t i) { return path.is_closed ? path.points[i % n] : path.points[std::min(i, n - 1)]; };
    auto add = [&](size_t i, V2 d) {
        size_t index = path.is_closed ? i % n : i;
        if (index < n) {
            float *q = &gr.d_shape_points[((size_t)path.point_offset + index) * 2];
            q[0] += d.x;
            q[1] += d.y;
        }
    };
    uint8_t nc = base < path.num_control_points.size() ? path.num_control_points[base] : 0;
    if (nc == 0) {
        size_t i0 = pid, i1 = (pid + 1) % n;
        V2 p0 = get(i0), p1 = get(i1);
        float t = (pt - p0).dot(p1 - p0) / (p1 - p0).dot(p1 - p0);
        if (t < 0.f) add(i0, d_cp);
        else if (t > 1.f) add(i1, d_cp);
        else {
            add(i0, d_cp * (1.0f - t));
            add(i1, d_cp * t);
        }
    } else if (nc == 1) {
        size_t i0 = pid, i1 = pid + 1, i2 = (pid + 2) % n;
        V2 p0 = get(i0), p1 = get(i1), p2 = get(i2);
        float t = min_t;
        V2 d_p0, d_p1, d_p2;
        if (t == 0.f) {
            d_p0 += d_cp;
        } else if (t == 1.f) {
            d_p2 += d_cp;
        } else {
            V2 a = (p0 - p1 * 2.0f) + p2;
            V2 b = p1 - p0;
            float A = a.dot(a);
            float B = 3.0f * a.dot(b);
            float C = 2.0f * b.dot(b) + a.dot(p0 - pt);
            float tt = 1.0f - t;
            V2 d_p = d_cp;
            float d_tt = 2.0f * tt * d_p.dot(p0) + 2.0f * t * d_p.dot(p1);
            float d_t = -d_tt + 2.0f * tt * d_p.dot(p1) + 2.0f * t * d_p.dot(p2);
            d_p0 += d_p * (tt * tt);
            d_p1 += d_p * (2.0f * tt * t);
            d_p2 += d_p * (t * t);
            float pd = 3.0f * A * t * t + 2.0f * B * t + C;
            if (std::fabs(pd) > 1.0e-6f) {
                float dA = -(d_t / pd) * t * t * t;
                float dB = -(d_t / pd) * t * t;
                float dC = -(d_t / pd) * t;
                float dD = -(d_t / pd);
                V2 q = p0 - pt;
                d_p0 += (((a * (2.0f * dA) + (b - a) * (3.0f * dB)) + (b * -4.0f) * dC) + (q + a) * dC) +
                        (b - q) * (2.0f * dD);
                d_p1 += (((a * (-4.0f * dA) + (a + b * -2.0f) * (3.0f * dB)) + (b * 4.0f) * dC) + q * (-2.0f * dC)) +
                        q * dD;
                d_p2 += (a * (2.0f * dA) + b * (3.0f * dB)) + q * dC;
                d_pt += a * (-dC) + b * dD;
            }
        }
        add(i0, d_p0);
        add(i1, d_p1);
        add(i2, d_p2);
    } else if (nc == 2) {
        size_t i0 = pid, i1 = pid + 1, i2 = pid + 2, i3 = (pid + 3) % n;
        V2 p0 = get(i0), p1 = get(i1), p2 = get(i2), p3 = get(i3);
        float t = min_t;
        V2 d_p0, d_p1, d_p2, d_p3;
        if (t == 0.f) {
            d_p0 += d_cp;
        } else if (t == 1.f) {
            d_p3 += d_cp;
        } else {
            V2 a = ((p0 * -1.0f + p1 * 3.0f) + p2 * -3.0f) + p3;
            V2 b = (p0 * 3.0f + p1 * -6.0f) + p2 * 3.0f;
            V2 c = p0 * -3.0f + p1 * 3.0f;
            float A = 3.0f * a.dot(a);
            if (std::fabs(A) < 1.0e-10f) return;
            V2 q = p0 - pt;
            float B = 5.0f * a.dot(b);
            float C = 4.0f * a.dot(c) + 2.0f * b.dot(b);
            float D = 3.0f * (b.dot(c) + a.dot(q));
            float E = c.dot(c) + 2.0f * q.dot(b);
            float F = q.dot(c);
            B = B / A; C = C / A; D = D / A; E = E / A; F = F / A;
            float tt = 1.0f - t;
            V2 d_p = d_cp;
            float d_tt = 3.0f * tt * tt * d_p.dot(p0) + 6.0f * tt * t * d_p.dot(p1) + 3.0f * t * t * d_p.dot(p2);
            float d_t = -d_tt + 3.0f * tt * tt * d_p.dot(p1) + 6.0f * tt * t * d_p.dot(p2) + 3.0f * t * t * d_p.dot(p3);
            d_p0 += d_p * (tt * tt * tt);
            d_p1 += d_p * (3.0f * tt * tt * t);
            d_p2 += d_p * (3.0f * tt * t * t);
            d_p3 += d_p * (t * t * t);
            float pd = 5.0f * t * t * t * t + 4.0f * B * t * t * t + 3.0f * C * t * t + 2.0f * D * t + E;
            if (std::fabs(pd) > 1.0e-10f) {
                float dB = -(d_t / pd) * t * t * t * t;
                float dC = -(d_t / pd) * t * t * t;
                float dD = -(d_t / pd) * t * t;
                float dE = -(d_t / pd) * t;
                float dF = -(d_t / pd);
                float dA = -dB * B / A - dC * C / A - dD * D / A - dE * E / A - dF * F / A;
                dB /= A; dC /= A; dD /= A; dE /= A; dF /= A;
                d_p0 += a * (3.0f * -1.0f * 2.0f * dA);
                d_p1 += a * (3.0f * 3.0f * 2.0f * dA);
                d_p2 += a * (3.0f * -3.0f * 2.0f * dA);
                d_p3 += a * (3.0f * 1.0f * 2.0f * dA);
                d_p0 += (b * -1.0f + a * 3.0f) * (5.0f * dB);
                d_p1 += (b * 3.0f + a * -6.0f) * (5.0f * dB);
                d_p2 += (b * -3.0f + a * 3.0f) * (5.0f * dB);
                d_p3 += b * (5.0f * dB);
                d_p0 += (c * -1.0f + a * -3.0f) * (4.0f * dC) + b * (3.0f * 2.0f * dC);
                d_p1 += (c * 3.0f + a * 3.0f) * (4.0f * dC) + b * (-6.0f * 2.0f * dC);
                d_p2 += c * (-3.0f * dC * 4.0f) + b * (3.0f * 2.0f * dC);
                d_p3 += c * (4.0f * dC);
                d_p0 += (c * 3.0f + b * -3.0f) * (3.0f * dD) + (a - q) * (3.0f * dD);
                d_p1 += (c * -6.0f + b * 3.0f) * (3.0f * dD) + q * (3.0f * 3.0f * dD);
                d_p2 += c * (3.0f * 3.0f * dD) + q * (-3.0f * 3.0f * dD);
                d_pt += a * (-1.0f * 3.0f * dD);
                d_p0 += c * (-3.0f * 2.0f * dE) + (b + q * 3.0f) * (2.0f * dE);
                d_p1 += c * (3.0f * 2.0f * dE) + q * (-6.0f * 2.0f * dE);
                d_p2 += q * (3.0f * 2.0f * dE);
                d_pt += b * (-1.0f * 2.0f * dE);
                d_p0 += c * dF + q * (-3.0f * dF);
                d_p1 += q * (3.0f * dF);
                d_pt += c * (-1.0f * dF);
            }
        }
        add(i0, d_p0);
        add(i1, d_p1);
        add(i2, d_p2);
        add(i3, d_p3);
    }
}

// d_closest_point_rect (backward/distance.rs:407-486)
static void d_closest_point_rect(V2 mn, V2 mx, V2 pt, V2 d_cp, V2 &d_min, V2 &d_max, V2 &d_pt) {
    auto dist_to_seg = [&](V2 p0, V2 p1) {
        float t = (pt - p0).dot(p1 - p0) / (p1 - p0).dot(p1 - p0);
        if (t < 0.f) return (p0 - pt).length();
        if (t > 1.f) return (p1 - pt).length();
        return ((p0 + (p1 - p0) * t) - pt).length();
    };
    V2 lt = mn, rt(mx.x, mn.y), lb(mn.x, mx.y), rb = mx;
    int id = 0;
    float md = dist_to_seg(lt, lb);
    float td = dist_to_seg(lt, rt), rd = dist_to_seg(rt, rb), bd = dist_to_seg(lb, rb);
    if (td < md) { md = td; id = 1; }
    if (rd < md) { md = rd; id = 2; }
    if (bd < md) { md = bd; id = 3; }
    auto update = [&](V2 p0, V2 p1, V2 &d_p0, V2 &d_p1) {
        float t = (pt - p0).dot(p1 - p0) / (p1 - p0).dot(p1 - p0);
        if (t < 0.f) {
            d_p0 += d_cp;
        } else if (t > 1.f) {
            d_p1 += d_cp;
        } else {
            V2 d_p = d_cp;
            d_p0 += d_p * (1.0f - t);
            d_p1 += d_p * t;
            float d_t = d_p.dot(p1 - p0);
            float denom = (p1 - p0).dot(p1 - p0);
            float d_num = d_t / denom;
            float d_den = d_t * (-t) / denom;
            d_pt += (p1 - p0) * d_num;
            d_p1 += (pt - p0) * d_num;
            d_p0 += (((p0 - p1) + p0) - pt) * d_num;
            d_p1 += (p1 - p0) * (2.0f * d_den);
            d_p0 += (p0 - p1) * (2.0f * d_den);
        }
    };
    V2 dlt, drt, dlb, drb;
    if (id == 0) update(lt, lb, dlt, dlb);
    else if (id == 1) update(lt, rt, dlt, drt);
    else if (id == 2) update(rt, rb, drt, drb);
    else update(lb, rb, dlb, drb);
    d_min += dlt;
    d_max.x += drt.x;
    d_min.y += drt.y;
    d_min.x += dlb.x;
    d_max.y += dlb.y;
    d_max += drb;
}

// d_closest_point_ellipse (backward/distance.rs:489-561)
static void d_closest_point_ellipse(V2 center, V2 radius, V2 pt, V2 d_cp, V2 &d_center, V2 &d_radius, V2 &d_pt) {
    float rx = std::fabs(radius.x), ry = std::fabs(radius.y);
    const float eps = 1.0e-6f;
    V2 local = pt - center;
    if (rx < eps && ry < eps) {
        d_center += d_cp;
        return;
    }
    if (rx < eps) {
        float hit = (local.y >= -ry && local.y <= ry) ? 1.0f : 0.0f;
        d_center += d_cp;
        d_pt.y += d_cp.y * hit;
        if (local.y > ry) d_radius.y += d_cp.y;
        else if (local.y < -ry) d_radius.y -= d_cp.y;
        return;
    }
    if (ry < eps) {
        float hit = (local.x >= -rx && local.x <= rx) ? 1.0f : 0.0f;
        d_center += d_cp;
        d_pt.x += d_cp.x * hit;
        if (local.x > rx) d_radius.x += d_cp.x;
        else if (local.x < -rx) d_radius.x -= d_cp.x;
        return;
    }
    float sx = local.x < 0.f ? -1.0f : 1.0f, sy = local.y < 0.f ? -1.0f : 1.0f;
    float x = std::fabs(local.x), y = std::fabs(local.y);
    float t = atan2_f(y * rx, x * ry);
    float g_t = 0.f;
    for (int i = 0; i < 20; ++i) {
        float s = sin_f(t), c = cos_f(t);
        float g = rx * x * s - ry * y * c + (ry * ry - rx * rx) * s * c;
        g_t = rx * x * c + ry * y * s + (ry * ry - rx * rx) * (c * c - s * s);
        if (std::fabs(g_t) < 1.0e-12f) break;
        float next = rclamp(t - g / g_t, 0.f, 1.57079632679489661923f);
        if (std::fabs(next - t) < 1.0e-6f) {
            t = next;
            break;
        }
        t = next;
    }
    float s = sin_f(t), c = cos_f(t);
    float d_t = 0.f;
    d_radius.x += d_cp.x * sx * c;
    d_t += d_cp.x * sx * (-rx * s);
    d_radius.y += d_cp.y * sy * s;
    d_t += d_cp.y * sy * (ry * c);
    float g_a = x * s - 2.0f * rx * s * c;
    float g_b = -y * c + 2.0f * ry * s * c;
    float g_x = rx * s;
    float g_y = -ry * c;
    if (std::fabs(g_t) > 1.0e-12f) {
        float inv = -d_t / g_t;
        d_radius.x += inv * g_a;
        d_radius.y += inv * g_b;
        float d_x = inv * g_x, d_y = inv * g_y;
        d_pt.x += d_x * sx;
        d_pt.y += d_y * sy;
        d_center.x -= d_x * sx;
        d_center.y -= d_y * sy;
    }
}

// d_closest_point (backward/distance.rs:128-185); pt is shape-local.
static void d_closest_point(const Shape &sh, uint32_t si, V2 pt, V2 d_cp, const Hit &h, Grads &gr, V2 &d_pt) {
    float *dp = &gr.d_shape_params[(size_t)si * 8];
    if (sh.kind == DVG_SHAPE_CIRCLE) {
        V2 v = pt - sh.center;
        V2 n = normalize(v);
        V2 d_center = d_cp;
        float d_radius = d_cp.dot(n);
        V2 d_n = d_cp * sh.radius_c;
        V2 d_v = d_normalize(v, d_n);
        d_center -= d_v;
        d_pt += d_v;
        dp[0] += d_center.x;
        dp[1] += d_center.y;
        dp[2] += d_radius;
    } else if (sh.kind == DVG_SHAPE_ELLIPSE) {
        V2 dc, dr;
        d_closest_point_ellipse(sh.center, sh.radius_e, pt, d_cp, dc, dr, d_pt);
        dp[0] += dc.x; dp[1] += dc.y; dp[2] += dr.x; dp[3] += dr.y;
    } else if (sh.kind == DVG_SHAPE_RECT) {
        V2 dmn, dmx;
        d_closest_point_rect(sh.rmin, sh.rmax, pt, d_cp, dmn, dmx, d_pt);
        dp[0] += dmn.x; dp[1] += dmn.y; dp[2] += dmx.x; dp[3] += dmx.y;
    } else {
        if (!h.path) return;
        d_closest_point_path(sh.path, pt, d_cp, h.seg, h.t, gr, d_pt);
    }
}

// d_compute_distance (backward/distance.rs:50-126); returns d(pt).
static V2 d_compute_distance(const Ctx &cx, uint32_t gi, uint32_t si, V2 pt, V2 closest, const Hit &h, float d_dist,
                             Grads &gr) {
    if ((pt - closest).length_squared() < 1.0e-10f) return V2();
    const Group &g = cx.sc->groups[gi];
    const Shape &sh = cx.sc->shapes[si];
    V2 lpg = g.canvas_to_shape.transform_point(pt);
    V2 lcg = g.canvas_to_shape.transform_point(closest);
    M3 sinv = sh.transform.inverse_or_identity();
    V2 lps = sinv.transform_point(lpg);
    V2 lcs = sinv.transform_point(lcg);
    V2 d_pt, d_closest;
    {  // d_distance(closest, pt, ...) (backward/math.rs:25-30)
        V2 d = d_length(pt - closest, d_dist);
        d_closest -= d;
        d_pt += d;
    }
    M3 d_s2c = M3::zero();
    V2 d_lcg;
    d_xform_pt(g.shape_to_canvas, lcg, d_closest, d_s2c, d_lcg);
    M3 d_st = M3::zero();
    V2 d_lcs;
    d_xform_pt(sh.transform, lcs, d_lcg, d_st, d_lcs);
    V2 d_lps;
    d_closest_point(sh, si, lps, d_lcs, h, gr, d_lps);
    M3 d_sinv = M3::zero();
    V2 d_lpg;
    d_xform_pt(sinv, lpg, d_lps, d_sinv, d_lpg);
    M3 d_c2s = M3::zero();
    d_xform_pt(g.canvas_to_shape, pt, d_lpg, d_c2s, d_pt);
    M3 tc2s = g.canvas_to_shape.transpose();
    M3 corr = m3_scale(tc2s, -1.0f).mul(d_c2s).mul(tc2s);
    float *dg = &gr.d_group_transform[(size_t)gi * 9];
    add_m3(dg, d_s2c);
    add_m3(dg, corr);
    M3 tsi = sinv.transpose();
    M3 scorr = m3_scale(tsi, -1.0f).mul(d_sinv).mul(tsi);
    float *ds = &gr.d_shape_transform[(size_t)si * 9];
    add_m3(ds, scorr);
    add_m3(ds, d_st);
    return d_pt;
}

struct PFrag {
    V4 color;  // rgb + coverage-weighted alpha
    uint32_t group;
    bool stroke;
    bool within;
    float distance;
    Hit hit;
};

V4 sample_color_prefiltered(const Ctx &cx, V2 screen_pt, const V4 *bg, const V4 *d_color, Grads *gr, long tidx,
                            size_t pix) {
    const Scene &sc = *cx.sc;
    V2 pt(screen_pt.x * (float)sc.width, screen_pt.y * (float)sc.height);
    std::vector<PFrag> frags;
    [[maybe_unused]] bool counting = COUNT_IS_ON();
    for (uint32_t gi = 0; gi < sc.groups.size(); ++gi) {
        const Group &g = sc.groups[gi];
        count_gate(cx, counting, pt, gi);
        if (g.stroke.present()) {
            Hit h;
            if (compute_distance(cx, gi, pt, INFINITY, h)) {
                const Shape &sh = sc.shapes[h.shape];
                float d = h.distance;
                float w = smoothstep(std::fabs(d) + sh.stroke_width) - smoothstep(std::fabs(d) - sh.stroke_width);
                if (w > 0.f) {
                    V4 c = paint_color(g.stroke, pt);
                    c.w *= w;
                    frags.push_back({c, gi, true, true, d, h});
                }
            }
        }
        if (g.fill.present()) {
            Hit h;
            bool hit = compute_distance(cx, gi, pt, 1.0f, h);
            bool inside = is_inside_group(cx, gi, pt);
            if (hit || inside) {
                Hit fh;  // inside beyond the band: shape 0, no path metadata
                float d = 1.0f;
                if (hit) {
                    fh = h;
                    d = h.distance;
                }
                if (!inside) d = -d;
                float w = smoothstep(d);
                if (w > 0.f) {
                    V4 c = paint_color(g.fill, pt);
                    c.w *= w;
                    frags.push_back({c, gi, false, hit, d, fh});
                }
            }
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
    // fragments are already in group order, stroke before fill (the stable
    // sort of sampling.rs:318 keeps that order)
    std::vector<V4> acc;
    acc.reserve(frags.size());
    float pr = 0.f, pg = 0.f, pb = 0.f, pa = 0.f;
    if (bg) { pr = bg->x; pg = bg->y; pb = bg->z; pa = bg->w; }
    for (const PFrag &f : frags) {
        float na = f.color.w;
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
            const PFrag &f = frags[ii];
            float d_prev_a = da * (1.0f - f.color.w);
            float d_alpha = da * (1.0f - pa2);
            d_alpha += dr * (f.color.x - pc.x) + dgc * (f.color.y - pc.y) + db * (f.color.z - pc.z);
            float dpr = dr * (1.0f - f.color.w), dpg = dgc * (1.0f - f.color.w), dpb = db * (1.0f - f.color.w);
            float dcr = dr * f.color.w, dcg = dgc * f.color.w, dcb = db * f.color.w;
            const Group &g = sc.groups[f.group];
            float *slot = &gr->d_group_data[(size_t)f.group * 16 + (f.stroke ? 12 : 8)];
            float d = f.distance;
            if (f.stroke) {
                const Shape &sh = sc.shapes[f.hit.shape];
                float ap = std::fabs(d) + sh.stroke_width, am = std::fabs(d) - sh.stroke_width;
                float w = smoothstep(ap) - smoothstep(am);
                if (w != 0.f) {
                    float d_w = w > 0.f ? (f.color.w / w) * d_alpha : 0.f;
                    float d_alpha_w = d_alpha * w;
                    V2 dtrans;
                    d_sample_paint(g.stroke, *gr, slot, pt, V4(dcr, dcg, dcb, d_alpha_w), dtrans);
                    accumulate_translation(*gr, tidx, dtrans);
                    float d_ap = d_smoothstep(ap, d_w);
                    float d_am = -d_smoothstep(am, d_w);
                    float d_d = d_ap + d_am;
                    if (d < 0.f) d_d = -d_d;
                    gr->d_shape_stroke_width[f.hit.shape] += d_ap - d_am;
                    if (std::fabs(d_d) > 1.0e-10f) {
                        V2 d_pt = d_compute_distance(cx, f.group, f.hit.shape, pt, f.hit.point, f.hit, d_d, *gr);
                        accumulate_translation(*gr, tidx, d_pt * -1.0f);
                    }
                }
            } else {
                float w = smoothstep(d);
                if (w != 0.f) {
                    float d_w = w > 0.f ? (f.color.w / w) * d_alpha : 0.f;
                    float d_alpha_w = d_alpha * w;
                    V2 dtrans;
                    d_sample_paint(g.fill, *gr, slot, pt, V4(dcr, dcg, dcb, d_alpha_w), dtrans);
                    accumulate_translation(*gr, tidx, dtrans);
                    float d_d = d_smoothstep(d, d_w);
                    if (d < 0.f) d_d = -d_d;
                    if (std::fabs(d_d) > 1.0e-10f && f.within) {
                        V2 d_pt = d_compute_distance(cx, f.group, f.hit.shape, pt, f.hit.point, f.hit, d_d, *gr);
                        accumulate_translation(*gr, tidx, d_pt * -1.0f);
                    }
                }
            }
            dr = dpr; dgc = dpg; db = dpb; da = d_prev_a;
        }
        if (bg) accumulate_background_grad(sc, *gr, pix, V4(dr, dgc, db, da));
    }
    return V4(fr, fg, fb, fa);
}

// sample_distance (backward/sampling.rs:14-68) for the adjoint d_dist of the
// signed distance at pt: the nearest boundary over the groups from last to
// first, d_compute_distance of it with the sign of the group's fill test,
// -d(pt) into translation pixel tidx (-1: none).  The reference's `weight`
// only scales a value it never reads.
static void sample_distance_grad(const Ctx &cx, V2 pt, float d_dist, long tidx, Grads &gr) {
    const Scene &sc = *cx.sc;
    float best = INFINITY;
    long best_g = -1;
    Hit bh;
    for (size_t gi = sc.groups.size(); gi-- > 0;) {
        Hit h;
        if (compute_distance(cx, (uint32_t)gi, pt, best, h) && h.distance < best) {
            best = h.distance;
            best_g = (long)gi;
            bh = h;
        }
    }
    if (best_g < 0) return;
    bool inside = sc.groups[best_g].fill.kind != DVG_PAINT_NONE && is_inside_group(cx, (uint32_t)best_g, pt);
    float d_abs = inside ? -d_dist : d_dist;
    V2 d_pt = d_compute_distance(cx, (uint32_t)best_g, bh.shape, pt, bh.point, bh, d_abs, gr);
    accumulate_translation(gr, tidx, d_pt * -1.0f);
}

// render_backward_positions (backward/render.rs:213-280): each point with a
// nonzero d_sdf; translation into the pixel holding it (`as i32` truncates
// toward zero).
void backward_positions(const Scene &sc, bool compute_translation, const float *pos, size_t n, const float *d_sdf,
                        Grads &gr) {
    Ctx cx;
    make_ctx(sc, cx);
    for (size_t i = 0; i < n; ++i) {
        float d_dist = d_sdf[i];
        if (d_dist == 0.0f) continue;
        V2 pt(pos[2 * i], pos[2 * i + 1]);
        long tidx = -1;
        if (compute_translation) {
            long x = (long)std::trunc(pt.x), y = (long)std::trunc(pt.y);
            if (x >= 0 && y >= 0 && x < (long)sc.width && y < (long)sc.height) tidx = y * (long)sc.width + x;
        }
        sample_distance_grad(cx, pt, d_dist, tidx, gr);
    }
}

// The d_sdf image term of render_backward (backward/render.rs:171-186): at
// every sample of every pixel (the CPU sample positions), sample_distance
// with the pixel's d_sdf; translation into that pixel.
void backward_sdf_image(const Scene &sc, const dvg_render_options &opt, bool compute_translation, const float *d_sdf,
                        Grads &gr) {
    Ctx cx;
    make_ctx(sc, cx);
    uint32_t nsx = std::max(opt.samples_x, 1u), nsy = std::max(opt.samples_y, 1u);
    size_t W = sc.width, H = sc.height;
    size_t total = W * H * nsx * nsy;
    bool jitter = opt.jitter != 0 && !opt.use_prefiltering;
    for (size_t idx = 0; idx < total; ++idx) {
        size_t sx = idx % nsx, sy = (idx / nsx) % nsy, x = (idx / (nsx * nsy)) % W, y = idx / ((size_t)nsx * nsy * W);
        size_t pix = y * W + x;
        float d_dist = d_sdf[pix];
        if (d_dist == 0.0f) continue;  // adds exact zeros
        float rx = 0.5f, ry = 0.5f;
        if (jitter) {
            Pcg32 rng(idx, opt.seed);
            rx = rng.next_f32();
            ry = rng.next_f32();
        }
        V2 pt((float)x + ((float)sx + rx) / (float)nsx, (float)y + ((float)sy + ry) / (float)nsy);
        sample_distance_grad(cx, pt, d_dist, compute_translation ? (long)pix : -1, gr);
    }
}

}  // namespace cpu
}  // namespace oracle
