// Device-side geometry, paint, filter and RNG helpers for the gfx950 kernels.
//
// The arithmetic follows the reference kernels operation by operation (the
// file:line of each restated function is given) so that images agree with the
// reference forward to f32 rounding; everything is f32 and branch-light.  The
// one departure: the winding solve's acos / cos / cube root use fast f32
// approximations on well-conditioned segments (see WIND_FAST_COND).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/dvg.h"
#include "dvg_kernels.h"

namespace dvgk {

#define DVG_DEV __device__ __forceinline__

constexpr float BIG = 1.0e20f;

DVG_DEV float clamp01(float v) { return v < 0.f ? 0.f : (v > 1.f ? 1.f : v); }
DVG_DEV float minf_(float a, float b) { return a < b ? a : b; }
DVG_DEV float maxf_(float a, float b) { return a > b ? a : b; }
DVG_DEV float absf_(float a) { return a < 0.f ? -a : a; }

// math.rs:7-10
DVG_DEV float smoothstep_unit(float d) {
    float t = clamp01((d + 1.0f) * 0.5f);
    return t * t * (3.0f - 2.0f * t);
}
// math.rs:168-175
DVG_DEV float d_smoothstep_unit(float d, float d_ret) {
    if (d >= -1.0f && d <= 1.0f) {
        float t = (d + 1.0f) * 0.5f;
        float dt = d_ret * (6.0f * t - 6.0f * t * t);
        return dt * 0.5f;
    }
    return 0.f;
}

// math.rs:47-64
DVG_DEV float rect_sdf(float px, float py, float x0, float y0, float x1, float y1) {
    float dx = maxf_(maxf_(x0 - px, 0.f), px - x1);
    float dy = maxf_(maxf_(y0 - py, 0.f), py - y1);
    float outside = sqrtf(dx * dx + dy * dy);
    float inside = minf_(minf_(px - x0, x1 - px), minf_(py - y0, y1 - py));
    return outside > 0.f ? outside : -inside;
}

DVG_DEV float dist_pt(float px, float py, float ax, float ay) {
    float dx = px - ax, dy = py - ay;
    return sqrtf(dx * dx + dy * dy);
}

// curves.rs:39-66
DVG_DEV void seg_dist_t(float px, float py, float ax, float ay, float bx, float by, float &dist, float &t) {
    float vx = bx - ax, vy = by - ay, wx = px - ax, wy = py - ay;
    float c1 = vx * wx + vy * wy;
    float c2 = vx * vx + vy * vy;
    t = 0.f;
    if (c2 > 0.f && c1 > 0.f) t = clamp01(c1 / c2);
    float qx = ax + t * vx, qy = ay + t * vy;
    dist = sqrtf((px - qx) * (px - qx) + (py - qy) * (py - qy));
}

// curves.rs:70-93
DVG_DEV float seg_dist(float px, float py, float ax, float ay, float bx, float by) {
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
    return sqrtf(d2);
}

// curves.rs:865-893
DVG_DEV int solve_quadratic(float a, float b, float c, float &r0, float &r1) {
    float disc = b * b - 4.0f * a * c;
    if (disc < 0.f) return 0;
    float root = sqrtf(disc);
    float q = (b < 0.f) ? -0.5f * (b - root) : -0.5f * (b + root);
    float t0 = q / a, t1 = c / q;
    if (t0 > t1) { r0 = t1; r1 = t0; }
    else { r0 = t0; r1 = t1; }
    return 2;
}

// curves.rs:949-959 (cube root through powf(x, 1/3))
DVG_DEV float cbrt_ref(float x) {
    const float third = 1.0f / 3.0f;
    if (x > 0.f) return powf(x, third);
    if (x < 0.f) return -powf(-x, third);
    return 0.f;
}

// Fast transcendentals of the winding solve on well-conditioned segments
// (record bit 3: sqrt(q) and |b / 3a| below WIND_FAST_COND, so that the root
// formulas -2 sqrt(q) cos(.) - b/3a and a + q/a - b/3a do not amplify their
// errors).  acos: Abramowitz & Stegun 4.4.46, sqrt(1 - x) times a degree-7
// polynomial, |error| <= 2e-8 in exact arithmetic; cos: the hardware's f32
// cosine (measured |error| <= 3e-7 on the solver's argument range); cube
// root: the hardware's exp2(log2(x) / 3) (measured relative error <= 1.3e-6).
// Checked against the libm solver on 16.7M random (segment, point) pairs:
// the winding decisions differ on 2, near-ties where libm implementations
// already disagree.  The reference's own GPU solver runs WGSL's pow / acos /
// cos, which the WGSL specification bounds only loosely (cos: 2^-11
// absolute).
constexpr float WIND_FAST_COND = 16.0f;
// Segments whose sqrt|q| or |b / 3a| reach WIND_PIECE_COND (record bit 4)
// lose their roots in the f32 Cardano solve: a small next to b and c inflates
// both terms, whose difference is the root.  The backward's inside tests, which
// follow the reference CPU's f64 solve, find their roots on the monotone
// pieces of y(t) instead (winding_cubic_pieces).  On the 1024^2 scene of 1000
// random closed cubic paths they are 1 % of the segments; solved by Cardano in
// f32 they cause 6 858 of the 6 879 inside decisions (over 11.8M sample /
// shape pairs) that differ from the f64 solve.  The 21 left are samples
// within rounding distance of an edge.
constexpr float WIND_PIECE_COND = 16.0f;

DVG_DEV float fast_acos(float x) {
    float ax = fabsf(x);
    float p = -0.0012624911f;
    p = p * ax + 0.0066700901f;
    p = p * ax - 0.0170881256f;
    p = p * ax + 0.0308918810f;
    p = p * ax - 0.0501743046f;
    p = p * ax + 0.0889789874f;
    p = p * ax - 0.2145988016f;
    p = p * ax + 1.5707963050f;
    float r = sqrtf(1.0f - ax) * p;
    return x < 0.f ? 3.14159265f - r : r;
}

DVG_DEV float cardano_a_fast(float rr, float r2, float q3) {
    const float third = 1.0f / 3.0f;
    float x = ((rr > 0.f) ? rr : -rr) + sqrtf(r2 - q3);
    float c = (x > 0.f) ? __builtin_amdgcn_exp2f(__builtin_amdgcn_logf(x) * third) : 0.f;
    return (rr > 0.f) ? -c : c;
}

// The Cardano term of solve_cubic (curves.rs:938-941):
//   rr > 0 ? -cbrt(rr + sqrt(r2 - q3)) : cbrt(-rr + sqrt(r2 - q3)).
// Both arms take the cube root of |rr| + sqrt(r2 - q3) >= 0, so one powf
// site serves every lane (a wave whose lanes split on the sign of rr would
// otherwise run two expansions); the value is the same bit for bit.
DVG_DEV float cardano_a(float rr, float r2, float q3) {
    const float third = 1.0f / 3.0f;
    float x = ((rr > 0.f) ? rr : -rr) + sqrtf(r2 - q3);
    float c = (x > 0.f) ? powf(x, third) : 0.f;
    return (rr > 0.f) ? -c : c;
}

// Correctly rounded f32 acos / cos / pow: the f64 function rounded once,
// which is the f32 value nearest the exact one unless that lies within f64
// rounding of a tie.  The prefiltered forward's closest-point solves use them:
// its Newton search stops at |f| < 1e-5, so the distance it returns moves with
// the last bit of the root-isolating transcendentals, and a last-bit change
// shifts a pixel by up to 7e-5 (1 % of the pixels of the 1024^2 prefiltered
// scene beyond 1e-5 with the f32 library functions, none with these).
DVG_DEV float acos_cr(float x) { return (float)acos((double)x); }
DVG_DEV float cos_cr(float x) { return (float)cos((double)x); }
DVG_DEV float pow_cr(float x, float y) { return (float)pow((double)x, (double)y); }

// curves.rs:897-946.  CR: correctly rounded transcendentals (above);
// otherwise fast ones on well-conditioned cubics and the f32 library's
// elsewhere.
template <bool CR = false>
DVG_DEV int solve_cubic(float a, float b, float c, float d, float r[3]) {
    const float eps = 1.0e-6f;
    if (absf_(a) < eps) {
        return solve_quadratic(b, c, d, r[0], r[1]);
    }
    float bb = b / a, cc = c / a, dd = d / a;
    const float third = 1.0f / 3.0f;
    float q = (bb * bb - 3.0f * cc) / 9.0f;
    float rr = (2.0f * bb * bb * bb - 9.0f * bb * cc + 27.0f * dd) / 54.0f;
    float r2 = rr * rr, q3 = q * q * q;
    if (CR) {
        if (r2 < q3) {
            float theta = acos_cr(rr / sqrtf(q3));
            float sq = sqrtf(q);
            const float two_pi = 2.0f * 3.14159265f;
            r[0] = -2.0f * sq * cos_cr(theta * third) - bb * third;
            r[1] = -2.0f * sq * cos_cr((theta + two_pi) * third) - bb * third;
            r[2] = -2.0f * sq * cos_cr((theta - two_pi) * third) - bb * third;
            return 3;
        }
        float x = ((rr > 0.f) ? rr : -rr) + sqrtf(r2 - q3);
        float cb = (x > 0.f) ? pow_cr(x, third) : 0.f;
        float ar = (rr > 0.f) ? -cb : cb;
        float br = (absf_(ar) > eps) ? q / ar : 0.f;
        r[0] = (ar + br) - bb * third;
        return 1;
    }
    // fast transcendentals when the root formulas do not amplify their errors
    bool fast = absf_(bb * third) < WIND_FAST_COND && q < WIND_FAST_COND * WIND_FAST_COND;
    if (r2 < q3) {
        float sq = sqrtf(q);
        const float two_pi = 2.0f * 3.14159265f;
        float off = bb * third;
        if (fast) {
            float theta = fast_acos(rr / sqrtf(q3));
            r[0] = -2.0f * sq * __cosf(theta * third) - off;
            r[1] = -2.0f * sq * __cosf((theta + two_pi) * third) - off;
            r[2] = -2.0f * sq * __cosf((theta - two_pi) * third) - off;
        } else {
            float theta = acosf(rr / sqrtf(q3));
            r[0] = -2.0f * sq * cosf(theta * third) - off;
            r[1] = -2.0f * sq * cosf((theta + two_pi) * third) - off;
            r[2] = -2.0f * sq * cosf((theta - two_pi) * third) - off;
        }
        return 3;
    }
    float ar = fast ? cardano_a_fast(rr, r2, q3) : cardano_a(rr, r2, q3);
    float br = (absf_(ar) > eps) ? q / ar : 0.f;
    r[0] = (ar + br) - bb * third;
    return 1;
}

// curves.rs:97-123
DVG_DEV void winding_line(float px, float py, float x0, float y0, float x1, float y1, uint32_t rule, float &w,
                          float &cr) {
    bool a = y0 <= py, b = y1 <= py;
    if (a != b) {
        float t = (py - y0) / (y1 - y0);
        float xi = x0 + t * (x1 - x0);
        if (xi > px) {
            if (rule == 1u) cr = 1.f - cr;
            else w += (y1 > y0) ? 1.f : -1.f;
        }
    }
}

// The reference CPU's line winding (distance/winding.rs:120-133): t in
// [0, 1] closed, curve at or right of px, horizontal lines skipped.
DVG_DEV void winding_line_cpu(float px, float py, float x0, float y0, float x1, float y1, uint32_t rule, float &w,
                              float &cr) {
    float dy = y1 - y0;
    if (fabsf(dy) <= 1.0e-8f) return;
    float t = (py - y0) / dy;
    if (t >= 0.f && t <= 1.f) {
        float tp = x0 - px + t * (x1 - x0);
        if (tp >= 0.f) {
            if (rule == 1u) cr = 1.f - cr;
            else w += (dy > 0.f) ? 1.f : -1.f;
        }
    }
}

// Per-segment terms of the winding tests of curves.rs:127-214 (quadratic and
// cubic segments, with solve_cubic of curves.rs:897-946), computed once per
// segment (k_wind_coef) with exactly the operations and order the reference
// applies to them, so that the per-sample tests below give bit-identical
// results while doing only the point-dependent part.
// Record of WIND_STRIDE floats:
//   cubic: a, b, c, q = (bb^2 - 3 cc) / 9, P = 2 bb^3 - 9 bb cc, q3 = q^3,
//          sqrt(q3), sqrt(q), bb / 3, 3a, 2b   (bb = b / a, cc = c / a),
//          root mask (below)
//   quad:  ay, by, -, by^2, 4 ay, 2 ay
//   all:   [12..15] = min y, max y, max x, min x of the control points (the
//          ray test's box, as the walk would compute it)
//   cubic: [16..20] = the monotone pieces of y(t) - y0 = a t^3 + b t^2 + c t
//          on [0, 1]: its critical points t1 <= t2 in (0, 1) (1 when absent)
//          and its values Y(t1), Y(t2), Y(1), computed in f64 and rounded;
//          [21] = 1 / (54 sqrt(q3)), [22] = 2 sqrt(q) + |bb / 3| + 1 (error
//          bounds of the backward's checked solve)
//   cubic: [24..29] = min and max of X(t) = ax t^3 + bx t^2 + cx t over each
//          monotone piece [0, t1], [t1, t2], [t2, 1] (ax, bx, cx the x
//          coefficients as the reference CPU forms them in f32; the pieces'
//          ends and X's extrema in f64, rounded), [30] = an absolute bound on
//          the rounding of those values (the crossing classification below)
//
// Root mask: in the three-root case root k is -2 sqrt(q) cos(phi_k) - bb/3
// with phi_0 in [0, pi/3], phi_1 in [2pi/3, pi], phi_2 in [-2pi/3, -pi/3]
// (theta = acos(.) in [0, pi]), so cos(phi_k) lies in [0.5, 1], [-1, -0.5],
// [-0.5, 0.5].  Bit k is clear when that interval, widened by a rounding
// margin, puts root k outside [0, 1]: the root would be discarded, so its
// cosine need not be evaluated.  NaN terms keep every bit set.  Bit 3 marks
// a well-conditioned segment (fast transcendentals, below).

// Range of X(t) = A t^3 + B t^2 + C t over [lo, hi] (f64; X's critical points
// e[0..ne) are the candidates inside).
DVG_DEV void xrange_piece(double A, double B, double C, const double *e, int ne, double lo, double hi,
                          double &mn, double &mx) {
    double v = ((A * lo + B) * lo + C) * lo;
    mn = mx = v;
    v = ((A * hi + B) * hi + C) * hi;
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
    for (int j = 0; j < ne; ++j) {
        if (!(e[j] > lo && e[j] < hi)) continue;
        v = ((A * e[j] + B) * e[j] + C) * e[j];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
}

// The critical points of A t^3 + B t^2 + C t (roots of 3A t^2 + 2B t + C) in
// the cancellation-free form; returns how many.
DVG_DEV int cubic_crit_f64(double A, double B, double C, double e[2]) {
    if (A == 0.0) {
        if (B == 0.0) return 0;
        e[0] = -C / (2.0 * B);
        return 1;
    }
    double disc = 4.0 * B * B - 12.0 * A * C;
    if (!(disc >= 0.0)) return 0;
    double sd = sqrt(disc);
    double qq = -0.5 * (2.0 * B + (B < 0.0 ? -sd : sd));
    if (qq == 0.0) {
        e[0] = -B / (3.0 * A);
        return 1;
    }
    e[0] = qq / (3.0 * A);
    e[1] = C / qq;
    return 2;
}

// Extremes over t in [0, 1] of one coordinate of a Bezier segment with
// control values p[0..n) (n = 2 line, 3 quadratic, 4 cubic), in f64: the
// ends and the roots of the derivative inside (0, 1).
DVG_DEV void bezier_range_f64(const double *p, int n, double &lo, double &hi) {
    lo = fmin(p[0], p[n - 1]);
    hi = fmax(p[0], p[n - 1]);
    double A = 0.0, B = 0.0, C = 0.0, e[2];  // p(t) - p[0] = A t^3 + B t^2 + C t
    int ne = 0;
    if (n == 3) {
        B = p[0] - 2.0 * p[1] + p[2];
        C = 2.0 * (p[1] - p[0]);
        if (B != 0.0) {
            e[0] = -C / (2.0 * B);
            ne = 1;
        }
    } else if (n == 4) {
        A = -p[0] + 3.0 * p[1] - 3.0 * p[2] + p[3];
        B = 3.0 * p[0] - 6.0 * p[1] + 3.0 * p[2];
        C = -3.0 * p[0] + 3.0 * p[1];
        ne = cubic_crit_f64(A, B, C, e);
    }
    for (int j = 0; j < ne; ++j) {
        double t = e[j];
        if (!(t > 0.0 && t < 1.0)) continue;
        double v = p[0] + ((A * t + B) * t + C) * t;
        lo = fmin(lo, v);
        hi = fmax(hi, v);
    }
}

// xs: the segment's four x control values (record slots 24..30), or null.
DVG_DEV void wind_coef_cubic(float y0, float y1, float y2, float y3, float *o, const float *xs = nullptr) {
    float a = -y0 + 3.0f * y1 - 3.0f * y2 + y3;
    float b = 3.0f * y0 - 6.0f * y1 + 3.0f * y2;
    float c = -3.0f * y0 + 3.0f * y1;
    o[0] = a; o[1] = b; o[2] = c;
    float bb = b / a, cc = c / a;
    float q = (bb * bb - 3.0f * cc) / 9.0f;
    o[3] = q;
    o[4] = 2.0f * bb * bb * bb - 9.0f * bb * cc;
    float q3 = q * q * q;
    o[5] = q3;
    o[6] = sqrtf(q3);
    o[7] = sqrtf(q);
    float off = bb * (1.0f / 3.0f), sq = o[7];
    o[8] = off;
    o[9] = 3.0f * a;
    o[10] = 2.0f * b;
    // root k = -2 sq cos(phi_k) - off with cos(phi_k) in [clo_k, chi_k]
    const float clo[3] = {0.4999990f, -1.0f, -0.5000010f}, chi[3] = {1.0f, -0.4999990f, 0.5000010f};
    float m = 1.0e-5f * (2.0f * fabsf(sq) + fabsf(off)) + 1.0e-30f;
    uint32_t mask = 0;
    for (int k = 0; k < 3; ++k) {
        float lo = -2.0f * sq * chi[k] - off, hi = -2.0f * sq * clo[k] - off;
        bool out = (hi < -m) || (lo > 1.0f + m);
        if (!out) mask |= 1u << k;
    }
    if (fabsf(sq) < WIND_FAST_COND && fabsf(off) < WIND_FAST_COND) mask |= 8u;  // well conditioned
    // Bit 4: the f32 solve loses the roots' position (a small next to b and c
    // inflates sqrt|q| and b / 3a, whose difference is the root), or a is
    // below the reference's quadratic cut-off.  The backward's inside tests
    // then solve in f64 as the reference CPU does.
    float sqa = sqrtf(fabsf(q));
    if (!(sqa < WIND_PIECE_COND && fabsf(off) < WIND_PIECE_COND) || !(fabsf(a) >= 1.0e-6f)) mask |= 16u;
    o[11] = (float)mask;
    // monotone pieces: roots of Y'(t) = 3a t^2 + 2b t + c in (0, 1)
    double A = a, B = b, C = c;
    double r0 = 2.0, r1 = 2.0;
    if (A == 0.0) {
        if (B != 0.0) r0 = -C / (2.0 * B);
    } else {
        double disc = 4.0 * B * B - 12.0 * A * C;
        if (disc >= 0.0) {
            double sd = sqrt(disc);
            double qq = -0.5 * (2.0 * B + (B < 0.0 ? -sd : sd));
            if (qq != 0.0) {
                r0 = qq / (3.0 * A);
                r1 = C / qq;
            } else {
                r0 = -B / (3.0 * A);
            }
        }
    }
    if (!(r0 > 0.0 && r0 < 1.0)) r0 = 1.0;
    if (!(r1 > 0.0 && r1 < 1.0)) r1 = 1.0;
    double t1 = r0 < r1 ? r0 : r1, t2 = r0 < r1 ? r1 : r0;
    o[16] = (float)t1;
    o[17] = (float)t2;
    o[18] = (float)(((A * t1 + B) * t1 + C) * t1);
    o[19] = (float)(((A * t2 + B) * t2 + C) * t2);
    o[20] = (float)(A + B + C);
    // the checked solve's per-segment terms: 1 / (54 sqrt(q3)), 2 sqrt(q) + |b / 3a| + 1
    o[21] = 1.0f / (54.0f * o[6]);
    o[22] = 2.0f * fabsf(sq) + fabsf(off) + 1.0f;
    o[23] = 0.f;
    for (int k = 24; k < WIND_STRIDE; ++k) o[k] = 0.f;
    if (xs) {
        // X(t) = ax t^3 + bx t^2 + cx t, the coefficients in f32 exactly as
        // distance/winding.rs:176-178 forms them
        float ax = -xs[0] + 3.0f * xs[1] - 3.0f * xs[2] + xs[3];
        float bx = 3.0f * xs[0] - 6.0f * xs[1] + 3.0f * xs[2];
        float cx = -3.0f * xs[0] + 3.0f * xs[1];
        double e[2];
        int ne = cubic_crit_f64(ax, bx, cx, e);
        double tb[4] = {0.0, t1, t2, 1.0};
        double big = 0.0;
        for (int k = 0; k < 3; ++k) {
            double mn, mx;
            xrange_piece(ax, bx, cx, e, ne, tb[k], tb[k + 1], mn, mx);
            o[24 + 2 * k] = (float)mn;
            o[25 + 2 * k] = (float)mx;
            big = fmax(big, fmax(fabs(mn), fabs(mx)));
        }
        // f64 -> f32 rounding of the ranges (2^-24 relative) and of the
        // query's sum, with a wide factor, plus an absolute floor
        o[30] = (float)(4.0e-6 * big + 1.0e-5);
    }
}

DVG_DEV void wind_coef_quad(float y0, float y1, float y2, float *o) {
    float ay = y0 - 2.0f * y1 + y2;
    float by = -2.0f * y0 + 2.0f * y1;
    o[0] = ay; o[1] = by; o[2] = 0.f;
    o[3] = by * by;
    o[4] = 4.0f * ay;
    o[5] = 2.0f * ay;
    for (int k = 6; k < 12; ++k) o[k] = 0.f;
    for (int k = 16; k < WIND_STRIDE; ++k) o[k] = 0.f;
}

// The quadratic winding test of curves.rs:127-166 with the segment's terms
// from its record.  CHK: raise `uns` where the reference CPU's f64 solve
// (winding.rs:135-165) could decide differently.
template <bool CHK = false>
DVG_DEV void winding_quad_w(float px, float py, float x0, float y0, float x1, float x2, const float *wc,
                            uint32_t rule, float &w, float &cr, float *uns = nullptr) {
    float ay = wc[0], by = wc[1];
    float cy = y0 - py;
    // solve_quadratic(ay, by, cy) with b*b and 4*a taken from the record
    float disc = wc[3] - wc[4] * cy;
    if (CHK && !(absf_(disc) > 1.0e-5f * (wc[3] + absf_(wc[4] * cy)))) *uns = 1.f;
    if (CHK && !(absf_(ay) >= 1.0e-6f)) *uns = 1.f;
    if (disc < 0.f) return;
    float root = sqrtf(disc);
    float qq = (by < 0.f) ? -0.5f * (by - root) : -0.5f * (by + root);
    float t0 = qq / ay, t1 = cy / qq;
    float r[2];
    if (t0 > t1) { r[0] = t1; r[1] = t0; }
    else { r[0] = t0; r[1] = t1; }
    for (int i = 0; i < 2; ++i) {
        float t = r[i];
        if (CHK && (absf_(t) <= 1.0e-5f || absf_(t - 1.f) <= 1.0e-5f)) *uns = 1.f;
        if (t >= 0.f && t <= 1.f) {
            float tt = 1.f - t;
            float x = (tt * tt) * x0 + (2.0f * tt * t) * x1 + (t * t) * x2;
            if (CHK && absf_(x - px) <= 1.0e-3f) *uns = 1.f;
            if (x > px) {
                if (rule == 1u) cr = 1.f - cr;
                else {
                    float dy = wc[5] * t + by;
                    w += (dy > 0.f) ? 1.f : -1.f;
                }
            }
        }
    }
}

// The cubic winding test of curves.rs:170-214 with the segment's terms from
// its record.
// FAST: the fast transcendentals on well-conditioned segments (record bit 3).
// The forward keeps the f32 library's: the fast ones move crossings by up to
// 1e-3 px, which flips about 0.1 % of the pixels of the 2048^2 scene of
// 10000 paths; the backward's inside tests take them.
//
// CHK: also raise `uns` when the reference CPU's f64 solve could decide
// differently (a root within its f32 error bound of 0 or 1, a crossing
// within its error bound of px, two roots about to merge).  The bound is the
// root formula's error, KT (below) per unit of 2 sqrt(q) + |b / 3a| + 1, with
// the acos argument's error amplified by 1 / sqrt(1 - rho^2) near a double
// root.
constexpr float WIND_KT_FAST = 3.0e-7f, WIND_KT_LIBM = 1.0e-7f;

DVG_DEV float bezier_x(float t, float x0, float x1, float x2, float x3) {
    float tt = 1.f - t;
    return (tt * tt * tt) * x0 + (3.0f * tt * tt * t) * x1 + (3.0f * tt * t * t) * x2 + (t * t * t) * x3;
}

DVG_DEV float bezier_dx(float t, float x0, float x1, float x2, float x3) {
    float tt = 1.f - t;
    return 3.0f * ((tt * tt) * (x1 - x0) + (2.0f * tt * t) * (x2 - x1) + (t * t) * (x3 - x2));
}

// CHK also raises *pieces where Cardano's f32 roots cannot settle the count
// (two roots about to merge, a root at a segment end): the caller then takes
// the monotone-piece evaluation, whose sign tests at the piece ends are exact.
template <bool FAST, bool CHK = false>
DVG_DEV void winding_cubic_w(float px, float py, float x0, float y0, float x1, float x2, float x3, const float *wc,
                             uint32_t rule, float &w, float &cr, float *uns = nullptr, bool *pieces = nullptr) {
    float a = wc[0];
    float d = y0 - py;
    float r[3];
    int n;
    // CHK: the roots' error bound is et0 + amp (amp: the double-root term),
    // evaluated only for roots near 0 or 1 or crossings near px
    float et0 = 0.f, amp = 0.f;
    bool risky = false;  // roots close to merging: no cheap pre-filter
    const float eps = 1.0e-6f;
    if (absf_(a) < eps) {
        n = solve_quadratic(wc[1], wc[2], d, r[0], r[1]);
        if (CHK) *uns = 1.f;  // the backward takes these on the piece path
    } else {
        // solve_cubic (curves.rs:897-946) from the per-segment terms
        float dd = d / a;
        float q = wc[3];
        float rr = (wc[4] + 27.0f * dd) / 54.0f;
        float r2 = rr * rr, q3 = wc[5];
        float off = wc[8];
        if (r2 < q3) {
            // roots the segment's mask rules out keep -1 (discarded below)
            uint32_t mask = (uint32_t)wc[11];
            r[0] = r[1] = r[2] = -1.0f;
            if ((mask & 7u) != 0u) {
                float sq = wc[7];
                const float third = 1.0f / 3.0f;
                const float two_pi = 2.0f * 3.14159265f;
                if (CHK) {
                    float gap = q3 - r2;  // q3 (1 - rho^2)
                    float kt = (FAST && (mask & 8u)) ? WIND_KT_FAST : WIND_KT_LIBM;
                    et0 = kt * wc[22];
                    if (!(gap > 1.0e-5f * q3)) *pieces = true;  // two roots about to merge
                    // acos argument error, amplified by 1 / sqrt(1 - rho^2)
                    float erho = 2.4e-7f * ((absf_(wc[4]) + 27.0f * absf_(dd)) * wc[21] + 1.0f);
                    amp = 0.667f * sq * erho * wc[6] * __builtin_amdgcn_rsqf(maxf_(gap, 1.0e-30f));
                    risky = gap < 1.0e-3f * q3 || et0 + amp > 1.0e-4f;
                }
                if (FAST && (mask & 8u)) {
                    float theta = fast_acos(rr / wc[6]);
                    if (mask & 1u) r[0] = -2.0f * sq * __cosf(theta * third) - off;
                    if (mask & 2u) r[1] = -2.0f * sq * __cosf((theta + two_pi) * third) - off;
                    if (mask & 4u) r[2] = -2.0f * sq * __cosf((theta - two_pi) * third) - off;
                } else {
                    float theta = acosf(rr / wc[6]);
                    if (mask & 1u) r[0] = -2.0f * sq * cosf(theta * third) - off;
                    if (mask & 2u) r[1] = -2.0f * sq * cosf((theta + two_pi) * third) - off;
                    if (mask & 4u) r[2] = -2.0f * sq * cosf((theta - two_pi) * third) - off;
                }
            }
            n = 3;
        } else {
            float ar = (FAST && ((uint32_t)wc[11] & 8u)) ? cardano_a_fast(rr, r2, q3) : cardano_a(rr, r2, q3);
            float br = (absf_(ar) > eps) ? q / ar : 0.f;
            r[0] = (ar + br) - off;
            n = 1;
            if (CHK) {
                float gap = r2 - q3;
                float kt = (FAST && ((uint32_t)wc[11] & 8u)) ? WIND_KT_FAST : WIND_KT_LIBM;
                et0 = 4.0f * kt * (absf_(ar) + absf_(br) + absf_(off) + 1.0f);
                if (!(gap > 1.0e-5f * r2)) *pieces = true;  // a complex pair about to turn real
                // sqrt(r2 - q3)'s error near the three-root boundary
                amp = 1.2e-7f * (absf_(ar) + absf_(br)) * absf_(rr) * __builtin_amdgcn_rsqf(maxf_(gap, 1.0e-30f));
                risky = gap < 1.0e-3f * r2 || et0 + amp > 1.0e-4f;
                if (absf_(absf_(ar) - eps) <= 4.0f * kt * absf_(ar)) *uns = 1.f;
            }
        }
    }
    for (int i = 0; i < n; ++i) {
        float t = r[i];
        if (CHK && (risky || absf_(t) <= 1.0e-2f || absf_(t - 1.f) <= 1.0e-2f)) {
            float et = et0 + amp;
            if (absf_(t) <= et || absf_(t - 1.f) <= et) *pieces = true;  // a root at a segment end
        }
        if (t >= 0.f && t <= 1.f) {
            float tt = 1.f - t;
            float x = (tt * tt * tt) * x0 + (3.0f * tt * tt * t) * x1 + (3.0f * tt * t * t) * x2 + (t * t * t) * x3;
            // (pre-filter: with et <= 1e-4 and |x'| <= 3 x 2048 the bound stays below 1 px)
            if (CHK && (risky || absf_(x - px) <= 1.0f)) {
                float xm = maxf_(maxf_(absf_(x0), absf_(x1)), maxf_(absf_(x2), absf_(x3)));
                float ex = absf_(bezier_dx(t, x0, x1, x2, x3)) * (et0 + amp) + 1.2e-7f * (xm + absf_(px));
                if (absf_(x - px) <= ex) *uns = 1.f;
            }
            if (x > px) {
                if (rule == 1u) cr = 1.f - cr;
                else {
                    float dy = wc[9] * t * t + wc[10] * t + wc[2];
                    w += (dy > 0.f) ? 1.f : -1.f;
                }
            }
        }
    }
}

// ---- the backward's cubic inside test on ill-conditioned segments -------------
// The reference CPU (distance/winding.rs:166-188) counts the roots t in [0, 1]
// of P(t) = a t^3 + b t^2 + c t + (y0 - py), solved in f64
// (distance/utils.rs:82-112), whose curve point lies strictly right of px,
// signed by P'(t).  On a segment with record bit 4 the same roots are found
// in f32 without Cardano's cancellation: each monotone piece of P (record
// slots 16..20) holds a root exactly when P changes sign over it, and the
// root is bracketed there, so safeguarded Newton steps on the unnormalised
// polynomial converge to it; the piece's direction is the sign of P'.  A
// cubic with |a| < 1e-6 is the CPU's quadratic b t^2 + c t + d
// (utils.rs:44-79), whose double roots count twice.

DVG_DEV float piece_root(float a, float b, float c, float d, float lo, float hi, float vlo, float vhi) {
    // regula falsi start, then Newton kept inside the bracket [lo, hi]
    float t = lo + (hi - lo) * (vlo / (vlo - vhi));
    bool neg_lo = vlo < 0.f;
    for (int it = 0; it < 12; ++it) {
        float p = ((a * t + b) * t + c) * t + d;
        if (p == 0.f) break;
        float dp = (3.0f * a * t + 2.0f * b) * t + c;
        if ((p < 0.f) == neg_lo) lo = t;
        else hi = t;
        float tn = t - p / dp;
        if (!(tn >= lo && tn <= hi)) tn = 0.5f * (lo + hi);
        bool done = fabsf(tn - t) <= 1.0e-7f;
        t = tn;
        if (done) break;
    }
    return t;
}

template <bool CHK>
DVG_DEV int winding_cubic_pieces(float px, float py, float x0, float y0, float x1, float x2, float x3,
                                 const float *wc, float &uns) {
    float a = wc[0], b = wc[1], c = wc[2];
    float d = y0 - py;
    int w = 0;
    float xs = 1.2e-7f * (fabsf(x0) + fabsf(x1) + fabsf(x2) + fabsf(x3) + fabsf(px)) + 2.0e-4f;
    if (!(fabsf(a) >= 1.0e-6f)) {
        if (CHK) uns = 1.f;  // the CPU's quadratic branch: rare, always taken exactly
        float r0, r1;
        int n = 0;
        if (fabsf(b) < 1.0e-12f) {
            if (fabsf(c) >= 1.0e-12f) { r0 = r1 = -d / c; n = 2; }
        } else {
            float disc = c * c - 4.0f * b * d;
            if (disc >= 0.f) {
                float root = sqrtf(disc);
                n = 2;
                if (root < 1.0e-12f) {
                    r0 = r1 = -0.5f * c / b;
                } else {
                    float q = (c < 0.f) ? -0.5f * (c - root) : -0.5f * (c + root);
                    if (fabsf(q) < 1.0e-12f) {
                        float inv = 0.5f / b;
                        r0 = (-c - root) * inv;
                        r1 = (-c + root) * inv;
                    } else {
                        r0 = q / b;
                        r1 = d / q;
                    }
                }
            }
        }
        for (int i = 0; i < n; ++i) {
            float t = i == 0 ? r0 : r1;
            if (t >= 0.f && t <= 1.f && bezier_x(t, x0, x1, x2, x3) > px)
                w += ((3.0f * a * t + 2.0f * b) * t + c > 0.f) ? 1 : -1;
        }
        return w;
    }
    float tk[4] = {0.f, wc[16], wc[17], 1.f};
    float vk[4] = {d, wc[18] + d, wc[19] + d, wc[20] + d};
    float ys = 1.0e-6f * (fabsf(wc[18]) + fabsf(wc[19]) + fabsf(wc[20]) + fabsf(d)) + 1.0e-6f;
    if (CHK)
        for (int k = 0; k < 4; ++k)
            if (fabsf(vk[k]) <= ys) uns = 1.f;  // a root at a piece end: the CPU's rounding decides
    for (int k = 0; k < 3; ++k) {
        float va = vk[k], vb = vk[k + 1];
        if ((va < 0.f) == (vb < 0.f) || !(tk[k] < tk[k + 1])) continue;
        float t = piece_root(a, b, c, d, tk[k], tk[k + 1], va, vb);
        float x = bezier_x(t, x0, x1, x2, x3);
        if (CHK && fabsf(x - px) <= xs) uns = 1.f;
        if (x > px) w += (vb > va) ? 1 : -1;
    }
    return w;
}

// The backward's cubic inside test: the f32 record solve, or the monotone
// pieces on segments whose Cardano solve is ill-conditioned (record bit 4).
// CHK: `uns` is raised where the f64 solve could decide differently.
//
// Before any root is solved, the crossings are classified on the monotone
// pieces: a piece over which P changes sign holds exactly one root, whose
// curve point lies inside the piece's x-range (record slots 24..29); when
// that range lies wholly right of px (beyond the record's rounding bound
// plus the query's) the crossing counts, signed by the piece's direction,
// and when wholly left it does not, which is what the f64 solve decides
// with its root anywhere in the piece.  Only when every crossing is so
// decided and no piece end lies within rounding of py (a root at a piece
// end, two roots about to merge: the solve's own rounding decides) does the
// classification stand; otherwise the solve below runs as before.
DVG_DEV bool winding_cubic_classify(float px, float py, float x0, float y0, const float *wc, int &n) {
    if (!(fabsf(wc[0]) >= 1.0e-6f)) return false;  // the CPU's quadratic branch
    float d = y0 - py;
    float v0 = d, v1 = wc[18] + d, v2 = wc[19] + d, v3 = wc[20] + d;
    float ys = 1.0e-6f * (fabsf(wc[18]) + fabsf(wc[19]) + fabsf(wc[20]) + fabsf(d)) + 1.0e-6f;
    if (!(fminf(fminf(fabsf(v0), fabsf(v1)), fminf(fabsf(v2), fabsf(v3))) > ys)) return false;
    float dx = x0 - px;
    float mx = wc[30] + 4.0e-6f * fabsf(dx);
    float t1 = wc[16], t2 = wc[17];
    int w = 0;
    bool ok = true;
    // piece k: ends (ta, tb), values (va, vb), X range (lo, hi)
    auto piece = [&](bool live, float va, float vb, float lo, float hi) {
        if (!live || (va < 0.f) == (vb < 0.f)) return;
        if (lo + dx > mx) w += (vb > va) ? 1 : -1;
        else if (!(hi + dx < -mx)) ok = false;
    };
    piece(t1 > 0.f, v0, v1, wc[24], wc[25]);
    piece(t1 < t2, v1, v2, wc[26], wc[27]);
    piece(t2 < 1.f, v2, v3, wc[28], wc[29]);
    n = w;
    return ok;
}

template <bool CHK>
DVG_DEV void winding_cubic_bwd(float px, float py, float x0, float y0, float x1, float x2, float x3,
                               const float *wc, uint32_t rule, float &w, float &cr, float &uns) {
    int nc;
    if (winding_cubic_classify(px, py, x0, y0, wc, nc)) {
        if (rule == 1u) {
            if (nc & 1) cr = 1.f - cr;
        } else {
            w += (float)nc;
        }
        return;
    }
    if ((uint32_t)wc[11] & 16u) {
        // the even-odd rule toggles per crossing: the parity of the signed sum
        int n = winding_cubic_pieces<CHK>(px, py, x0, y0, x1, x2, x3, wc, uns);
        if (rule == 1u) {
            if (n & 1) cr = 1.f - cr;
        } else {
            w += (float)n;
        }
    } else if (!CHK) {
        winding_cubic_w<true>(px, py, x0, y0, x1, x2, x3, wc, rule, w, cr);
    } else {
        float tw = 0.f, tc = 0.f, tu = 0.f;
        bool pieces = false;
        winding_cubic_w<true, true>(px, py, x0, y0, x1, x2, x3, wc, rule, tw, tc, &tu, &pieces);
        if (pieces) {
            int n = winding_cubic_pieces<true>(px, py, x0, y0, x1, x2, x3, wc, uns);
            tw = (float)n;
            tc = (n & 1) ? 1.f : 0.f;
        } else if (tu > 0.f) {
            uns = 1.f;
        }
        if (rule == 1u) {
            if (tc > 0.f) cr = 1.f - cr;
        } else {
            w += tw;
        }
    }
}

// ---- the reference CPU's winding in f64 -----------------------------------------
// distance/winding.rs:118-192 over distance/utils.rs:44-123, restated in f64
// with the same steps (no contraction: the library is built with
// -ffp-contract=off).  The boundary samples whose f32 inside tests are
// uncertain are evaluated again with these, and the self-test checks the f32
// tests against them.
DVG_DEV double cbrt_f64(double x) {  // utils.rs:115-123
    if (x > 0.0) return pow(x, 1.0 / 3.0);
    if (x < 0.0) return -pow(-x, 1.0 / 3.0);
    return 0.0;
}

DVG_DEV int quadratic_roots_f64(double a, double b, double c, double t[2]) {  // utils.rs:44-79
    const double eps = 1.0e-12;
    if (fabs(a) < eps) {
        if (fabs(b) < eps) return 0;
        t[0] = t[1] = -c / b;
        return 2;
    }
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) return 0;
    double root = sqrt(disc);
    if (fabs(root) < eps) {
        t[0] = t[1] = -0.5 * b / a;
        return 2;
    }
    double q = (b < 0.0) ? -0.5 * (b - root) : -0.5 * (b + root);
    double lo, hi;
    if (fabs(q) < eps) {
        double inv = 0.5 / a;
        lo = (-b - root) * inv;
        hi = (-b + root) * inv;
    } else {
        lo = q / a;
        hi = c / q;
    }
    t[0] = lo < hi ? lo : hi;
    t[1] = lo < hi ? hi : lo;
    return 2;
}

DVG_DEV int cubic_roots_f64(double a, double b, double c, double d, double t[3]) {  // utils.rs:82-112
    if (fabs(a) < 1.0e-6) return quadratic_roots_f64(b, c, d, t);
    double bn = b / a, cn = c / a, dn = d / a;
    double q = (bn * bn - 3.0 * cn) / 9.0;
    double r = (2.0 * bn * bn * bn - 9.0 * bn * cn + 27.0 * dn) / 54.0;
    double q3 = q * q * q;
    if (r * r < q3) {
        const double pi = 3.14159265358979323846;
        double theta = acos(r / sqrt(q3));
        double s = -2.0 * sqrt(q);
        t[0] = s * cos(theta / 3.0) - bn / 3.0;
        t[1] = s * cos((theta + 2.0 * pi) / 3.0) - bn / 3.0;
        t[2] = s * cos((theta - 2.0 * pi) / 3.0) - bn / 3.0;
        return 3;
    }
    double disc = sqrt(r * r - q3);
    double ar = (r > 0.0) ? -cbrt_f64(r + disc) : cbrt_f64(-r + disc);
    double br = (fabs(ar) > 1.0e-6) ? q / ar : 0.0;
    t[0] = (ar + br) - bn / 3.0;
    return 1;
}

DVG_DEV int winding_cubic_f64(float px, float py, const float *x, const float *y) {  // winding.rs:166-188
    double a = (double)(-y[0] + 3.0f * y[1] - 3.0f * y[2] + y[3]);
    double b = (double)(3.0f * y[0] - 6.0f * y[1] + 3.0f * y[2]);
    double c = (double)(-3.0f * y[0] + 3.0f * y[1]);
    double d = (double)(y[0] - py);
    double t[3];
    int n = cubic_roots_f64(a, b, c, d, t);
    double ax = (double)(-x[0] + 3.0f * x[1] - 3.0f * x[2] + x[3]);
    double bx = (double)(3.0f * x[0] - 6.0f * x[1] + 3.0f * x[2]);
    double cx = (double)(-3.0f * x[0] + 3.0f * x[1]);
    double dx = (double)(x[0] - px);
    int w = 0;
    for (int i = 0; i < n; ++i) {
        double s = t[i];
        if (s >= 0.0 && s <= 1.0) {
            double tp = ax * s * s * s + bx * s * s + cx * s + dx;
            if (tp > 0.0) w += (3.0 * a * s * s + 2.0 * b * s + c > 0.0) ? 1 : -1;
        }
    }
    return w;
}

// winding.rs:135-165 (quadratic segment)
DVG_DEV int winding_quad_f64(float px, float py, const float *x, const float *y) {
    double a = (double)(y[0] - 2.0f * y[1] + y[2]);
    double b = (double)(-2.0f * y[0] + 2.0f * y[1]);
    double c = (double)(y[0] - py);
    double t[2] = {0.0, 0.0};
    int n = 0;
    if (fabs(a) < 1.0e-12) {
        if (fabs(b) > 1.0e-12) { t[0] = -c / b; n = 1; }
    } else {
        n = quadratic_roots_f64(a, b, c, t);
    }
    double ax = (double)(x[0] - 2.0f * x[1] + x[2]);
    double bx = (double)(-2.0f * x[0] + 2.0f * x[1]);
    double cx = (double)(x[0] - px);
    int w = 0;
    for (int i = 0; i < n; ++i) {
        double s = t[i];
        if (s >= 0.0 && s <= 1.0) {
            double tp = ax * s * s + bx * s + cx;
            if (tp > 0.0) w += (2.0 * a * s + b > 0.0) ? 1 : -1;
        }
    }
    return w;
}

DVG_DEV float det2(float ax, float ay, float bx, float by) { return ax * by - ay * bx; }

// curves.rs:792-857
DVG_DEV void quad_closest_approx(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                                 float &ox, float &oy, float &ot) {
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
        if (absf_(denom) > 1.0e-8f) t = clamp01((ap + bp) / denom);
        else t = clamp01((ap + bp) * 0.5f);
    }
    float tt = 1.f - t;
    ox = (tt * tt) * x0 + (2.0f * tt * t) * x1 + (t * t) * x2;
    oy = (tt * tt) * y0 + (2.0f * tt * t) * y1 + (t * t) * y2;
    ot = t;
}

// curves.rs:218-279 (closest_point_quadratic_with_t; distance_to_quadratic is
// the same arithmetic)
template <bool CR = false>
DVG_DEV void quad_closest(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                          bool approx, float &dist, float &tout) {
    float best_t = 0.f;
    float best = dist_pt(px, py, x0, y0);
    if (approx) {
        float ax, ay, at;
        quad_closest_approx(px, py, x0, y0, x1, y1, x2, y2, ax, ay, at);
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
        float r[3];
        int n = solve_cubic<CR>(a, b, c, d, r);
        for (int i = 0; i < n; ++i) {
            float t = r[i];
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

DVG_DEV float eval_quintic(float t, float b, float c, float d, float e, float f) {  // curves.rs:777
    return ((((t + b) * t + c) * t + d) * t + e) * t + f;
}
DVG_DEV float eval_quintic_deriv(float t, float b, float c, float d, float e) {  // curves.rs:783
    return (((5.0f * t + 4.0f * b) * t + 3.0f * c) * t + 2.0f * d) * t + e;
}

// curves.rs:507-556
DVG_DEV void cubic_closest_approx(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                                  float x3, float y3, float &dist, float &tout) {
    const float inv = 1.0f / 8.0f;
    float prev_x = x0, prev_y = y0, prev_t = 0.f;
    float best = dist_pt(px, py, x0, y0), best_t = 0.f;
    for (int i = 1; i <= 8; ++i) {
        float t = (float)i * inv;
        float tt = 1.0f - t, tt2 = tt * tt, t2 = t * t;
        float a = tt2 * tt, b = 3.0f * tt2 * t, c = 3.0f * tt * t2, d = t2 * t;
        float cx = a * x0 + b * x1 + c * x2 + d * x3;
        float cy = a * y0 + b * y1 + c * y2 + d * y3;
        float sd, st;
        seg_dist_t(px, py, prev_x, prev_y, cx, cy, sd, st);
        float lt = prev_t + st * (t - prev_t);
        if (sd < best) { best = sd; best_t = lt; }
        prev_x = cx; prev_y = cy; prev_t = t;
    }
    dist = best;
    tout = best_t;
}

// curves.rs:561-731 (closest_point_cubic_with_t; distance_to_cubic,
// curves.rs:341-505, is the same arithmetic)
template <bool CR = false>
DVG_DEV void cubic_closest(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2, float x3,
                           float y3, bool approx, float &dist, float &tout) {
    if (approx) {
        cubic_closest_approx(px, py, x0, y0, x1, y1, x2, y2, x3, y3, dist, tout);
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
    if (absf_(a) > 1.0e-8f) {
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
        float pr[3];
        int num_p = solve_cubic<CR>(p1a, p1b, p1c, p1d, pr);
        float iv[4];
        int ni = 0;
        if (q_root >= 0.f && q_root <= 1.f) iv[ni++] = q_root;
        for (int i = 0; i < num_p; ++i) iv[ni++] = pr[i];
        for (int j = 1; j < ni; ++j) {
            int k = j;
            while (k > 0) {
                float u = iv[k - 1], v = iv[k];
                if (u <= v) break;
                iv[k - 1] = v;
                iv[k] = u;
                --k;
            }
        }
        float lower = 0.f;
        int idx = 0;
        while (idx <= ni) {
            if (idx < ni && iv[idx] < 0.f) { ++idx; continue; }
            float upper = (idx < ni) ? minf_(iv[idx], 1.f) : 1.f;
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
            }
            float t = (lb + ub) * 0.5f;
            for (int it = 0; it < 20; ++it) {
                if (t < lb || t > ub) t = (lb + ub) * 0.5f;
                float value = eval_quintic(t, b, c, d, e, f);
                if (absf_(value) < 1.0e-5f || it == 19) break;
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

// Closest point on an axis-aligned ellipse, src/distance/shape.rs:117-164
// (the CPU query used by the backward pass).
DVG_DEV void ellipse_closest(float cx, float cy, float rx_, float ry_, float px, float py, float &ox, float &oy) {
    float rx = absf_(rx_), ry = absf_(ry_);
    const float eps = 1.0e-6f;
    if (rx < eps && ry < eps) { ox = cx; oy = cy; return; }
    if (rx < eps) { ox = cx; oy = cy + fminf(fmaxf(py - cy, -ry), ry); return; }
    if (ry < eps) { ox = cx + fminf(fmaxf(px - cx, -rx), rx); oy = cy; return; }
    float dx = px - cx, dy = py - cy;
    if (absf_(dx) < eps && absf_(dy) < eps) { ox = cx + rx; oy = cy; return; }
    float sx = dx < 0.f ? -1.f : 1.f, sy = dy < 0.f ? -1.f : 1.f;
    dx = absf_(dx);
    dy = absf_(dy);
    float t = atan2f(dy * rx, dx * ry);
    for (int i = 0; i < 20; ++i) {
        float s = sinf(t), co = cosf(t);
        float g = rx * dx * s - ry * dy * co + (ry * ry - rx * rx) * s * co;
        float gt = rx * dx * co + ry * dy * s + (ry * ry - rx * rx) * (co * co - s * s);
        if (absf_(gt) < 1.0e-12f) break;
        float next = t - g / gt;
        next = next < 0.f ? 0.f : (next > 1.57079632679489661923f ? 1.57079632679489661923f : next);
        if (absf_(next - t) < 1.0e-6f) { t = next; break; }
        t = next;
    }
    float s = sinf(t), co = cosf(t);
    ox = cx + sx * rx * co;
    oy = cy + sy * ry * s;
}

// ---- the reference CPU's closest point on a segment ----------------------------
// distance/curve.rs:10-264, the query of the backward's distance adjoint and of
// the signed-distance gradients.  Points, distances and t are f32 as on the CPU;
// the quadratic's cubic and the cubic's quintic are formed and solved in f64
// (curve.rs:48-61, :145-242) with cubic_roots_f64, the CPU's solve_cubic.  The
// GPU forward's f32 solves (quad_closest / cubic_closest above) place t up to
// ~1e-3 away from these on ill-conditioned segments far from the point.

// curve.rs:10-21
DVG_DEV void seg_closest_cpu(float px, float py, float ax, float ay, float bx, float by, float &ox, float &oy,
                             float &dist, float &t) {
    float abx = bx - ax, aby = by - ay;
    float denom = abx * abx + aby * aby;
    if (denom <= 1.0e-10f) {
        ox = ax; oy = ay; t = 0.f;
        dist = dist_pt(px, py, ax, ay);
        return;
    }
    t = clamp01(((px - ax) * abx + (py - ay) * aby) / denom);
    ox = ax + abx * t;
    oy = ay + aby * t;
    dist = dist_pt(px, py, ox, oy);
}

// curve.rs:266-296 (the CPU's own approximation: t stays 0 on a degenerate
// denominator, unlike the GPU's curves.rs:792-857)
DVG_DEV void quad_approx_cpu(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                             float &ox, float &oy, float &ot) {
    float b0x = x0 - px, b0y = y0 - py, b1x = x1 - px, b1y = y1 - py, b2x = x2 - px, b2y = y2 - py;
    float a = det2(b0x, b0y, b2x, b2y);
    float b = 2.0f * det2(b1x, b1y, b0x, b0y);
    float d = 2.0f * det2(b2x, b2y, b1x, b1y);
    float f = b * d - a * a;
    float d21x = b2x - b1x, d21y = b2y - b1y, d10x = b1x - b0x, d10y = b1y - b0y, d20x = b2x - b0x,
          d20y = b2y - b0y;
    float gx = ((d21x * b) + (d10x * d)) + (d20x * a), gy = ((d21y * b) + (d10y * d)) + (d20y * a);
    gx = gx * 2.0f;
    gy = gy * 2.0f;
    float gfx = gy, gfy = -gx;
    float t = 0.f;
    float denom = gfx * gfx + gfy * gfy;
    if (denom > 1.0e-8f) {
        float s = -f / denom;
        float ppx = gfx * s, ppy = gfy * s;
        float d0px = b0x - ppx, d0py = b0y - ppy;
        float ap = det2(d0px, d0py, d20x, d20y);
        float bp = 2.0f * det2(d10x, d10y, d0px, d0py);
        float denom2 = 2.0f * a + b + d;
        if (absf_(denom2) > 1.0e-8f) t = clamp01((ap + bp) / denom2);
    }
    float tt = 1.0f - t;
    ox = ((x0 * (tt * tt)) + (x1 * (2.0f * tt * t))) + (x2 * (t * t));
    oy = ((y0 * (tt * tt)) + (y1 * (2.0f * tt * t))) + (y2 * (t * t));
    ot = t;
}

// curve.rs:26-78
DVG_DEV void quad_closest_cpu(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                              bool approx, float &ox, float &oy, float &dist, float &tout) {
    if (approx) {
        quad_approx_cpu(px, py, x0, y0, x1, y1, x2, y2, ox, oy, tout);
        dist = dist_pt(px, py, ox, oy);
        return;
    }
    float bt = 0.f, bx = x0, by = y0;
    float best = dist_pt(px, py, x0, y0);
    float de = dist_pt(px, py, x2, y2);
    if (de < best) { best = de; bx = x2; by = y2; bt = 1.f; }
    double ax = (double)(x0 - 2.0f * x1 + x2), ay = (double)(y0 - 2.0f * y1 + y2);
    double bbx = (double)(-x0 + x1), bby = (double)(-y0 + y1);
    double cx = (double)(x0 - px), cy = (double)(y0 - py);
    double a = ax * ax + ay * ay;
    double b = 3.0 * (ax * bbx + ay * bby);
    double c = 2.0 * (bbx * bbx + bby * bby) + (ax * cx + ay * cy);
    double d = bbx * cx + bby * cy;
    double r[3];
    int n = cubic_roots_f64(a, b, c, d, r);
    for (int i = 0; i < n; ++i) {
        double t = r[i];
        if (t >= 0.0 && t <= 1.0) {
            float t32 = (float)t, tt = 1.0f - t32;
            float qx = ((x0 * (tt * tt)) + (x1 * (2.0f * tt * t32))) + (x2 * (t32 * t32));
            float qy = ((y0 * (tt * tt)) + (y1 * (2.0f * tt * t32))) + (y2 * (t32 * t32));
            float dd = dist_pt(px, py, qx, qy);
            if (dd < best) { best = dd; bx = qx; by = qy; bt = t32; }
        }
    }
    ox = bx; oy = by; dist = best; tout = bt;
}

// curve.rs:80-119: the closest point of eight chords; the point returned is
// the one on the chord
DVG_DEV void cubic_approx_cpu(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                              float x3, float y3, float &ox, float &oy, float &dist, float &tout) {
    const float inv = 1.0f / 8.0f;
    float prev_x = x0, prev_y = y0, prev_t = 0.f;
    float bx = x0, by = y0, bt = 0.f, best = dist_pt(px, py, x0, y0);
    for (int i = 1; i <= 8; ++i) {
        float t = (float)i * inv;
        float tt = 1.0f - t, tt2 = tt * tt, t2 = t * t;
        float a = tt2 * tt, b = 3.0f * tt2 * t, c = 3.0f * tt * t2, d = t2 * t;
        float cx = a * x0 + b * x1 + c * x2 + d * x3;
        float cy = a * y0 + b * y1 + c * y2 + d * y3;
        float sx, sy, sd, st;
        seg_closest_cpu(px, py, prev_x, prev_y, cx, cy, sx, sy, sd, st);
        float lt = prev_t + st * (t - prev_t);
        if (sd < best) { best = sd; bx = sx; by = sy; bt = lt; }
        prev_x = cx; prev_y = cy; prev_t = t;
    }
    ox = bx; oy = by; dist = best; tout = bt;
}

// curve.rs:124-264
DVG_DEV void cubic_closest_cpu(float px, float py, float x0, float y0, float x1, float y1, float x2, float y2,
                               float x3, float y3, bool approx, float &ox, float &oy, float &dist, float &tout) {
    if (approx) {
        cubic_approx_cpu(px, py, x0, y0, x1, y1, x2, y2, x3, y3, ox, oy, dist, tout);
        return;
    }
    float bt = 0.f, bx = x0, by = y0;
    float best = dist_pt(px, py, x0, y0);
    float de = dist_pt(px, py, x3, y3);
    if (de < best) { best = de; bx = x3; by = y3; bt = 1.f; }
    ox = bx; oy = by; dist = best; tout = bt;
    double ax = (double)(-x0 + 3.0f * x1 - 3.0f * x2 + x3);
    double ay = (double)(-y0 + 3.0f * y1 - 3.0f * y2 + y3);
    double bbx = (double)(3.0f * x0 - 6.0f * x1 + 3.0f * x2);
    double bby = (double)(3.0f * y0 - 6.0f * y1 + 3.0f * y2);
    double cx = (double)(-3.0f * x0 + 3.0f * x1);
    double cy = (double)(-3.0f * y0 + 3.0f * y1);
    double dx = (double)(x0 - px), dy = (double)(y0 - py);
    double a = 3.0 * (ax * ax + ay * ay);
    if (fabs(a) < 1.0e-10) return;
    double b = 5.0 * (ax * bbx + ay * bby);
    double c = 4.0 * (ax * cx + ay * cy) + 2.0 * (bbx * bbx + bby * bby);
    double d = 3.0 * ((bbx * cx + bby * cy) + (ax * dx + ay * dy));
    double e = (cx * cx + cy * cy) + 2.0 * (dx * bbx + dy * bby);
    double f = dx * cx + dy * cy;
    b = b / a; c = c / a; d = d / a; e = e / a; f = f / a;
    double p1a = (2.0 / 5.0) * c - (4.0 / 25.0) * b * b;
    double p1b = (3.0 / 5.0) * d - (3.0 / 25.0) * b * c;
    double p1c = (4.0 / 5.0) * e - (2.0 / 25.0) * b * d;
    double p1d = f - b * e / 25.0;
    double q_root = -b / 5.0;
    double pr[3];
    int num_p = cubic_roots_f64(p1a, p1b, p1c, p1d, pr);
    double iv[4];
    int ni = 0;
    if (q_root >= 0.0 && q_root <= 1.0) iv[ni++] = q_root;
    for (int i = 0; i < num_p; ++i) iv[ni++] = pr[i];
    for (int j = 1; j < ni; ++j) {
        int k = j;
        while (k > 0 && iv[k - 1] > iv[k]) {
            double u = iv[k - 1];
            iv[k - 1] = iv[k];
            iv[k] = u;
            --k;
        }
    }
    double lower = 0.0;
    for (int j = 0; j <= ni; ++j) {
        if (j < ni && iv[j] < 0.0) continue;
        double upper = (j < ni) ? fmin(iv[j], 1.0) : 1.0;
        double lb = lower, ub = upper;
        double lbe = ((((lb + b) * lb + c) * lb + d) * lb + e) * lb + f;
        double ube = ((((ub + b) * ub + c) * ub + d) * ub + e) * ub + f;
        if (lbe * ube > 0.0) {
            lower = upper;
            continue;
        }
        if (lbe > ube) {
            double u = lb;
            lb = ub;
            ub = u;
        }
        double t = 0.5 * (lb + ub);
        for (int it = 0; it < 20; ++it) {
            if (!(t >= lb && t <= ub)) t = 0.5 * (lb + ub);
            double value = ((((t + b) * t + c) * t + d) * t + e) * t + f;
            if (fabs(value) < 1.0e-5 || it == 19) break;
            if (value > 0.0) ub = t; else lb = t;
            double der = (((5.0 * t + 4.0 * b) * t + 3.0 * c) * t + 2.0 * d) * t + e;
            t -= value / der;
        }
        if (t >= 0.0 && t <= 1.0) {
            float t32 = (float)t, tt = 1.0f - t32;
            float qx = (((x0 * (tt * tt * tt)) + (x1 * (3.0f * tt * tt * t32))) + (x2 * (3.0f * tt * t32 * t32))) +
                       (x3 * (t32 * t32 * t32));
            float qy = (((y0 * (tt * tt * tt)) + (y1 * (3.0f * tt * tt * t32))) + (y2 * (3.0f * tt * t32 * t32))) +
                       (y3 * (t32 * t32 * t32));
            float dd = dist_pt(px, py, qx, qy);
            if (dd < best) { best = dd; bx = qx; by = qy; bt = t32; }
        }
        if (upper >= 1.0) break;
        lower = upper;
    }
    ox = bx; oy = by; dist = best; tout = bt;
}

// ---- PCG32, src/renderer/rng.rs:16-43 (native 64-bit integer arithmetic) ----
struct Pcg32 {
    uint64_t state, inc;
    DVG_DEV Pcg32(uint64_t idx, uint64_t seed) {
        state = 0;
        inc = ((idx + 1) << 1) | 1;
        next_u32();
        state = state + (0x853c49e6748fea9bULL + seed);
        next_u32();
    }
    DVG_DEV uint32_t next_u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + (inc | 1);
        uint32_t xs = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    DVG_DEV float next_f32() {
        return (float)(next_u32() >> 9) * (1.0f / 8388608.0f);
    }
};

// rng.rs:167-197: the stratified jitter of sample (x, y, sx, sy); the stream
// index is ((y*W + x)*nsy + sy)*nsx + sx (u32 arithmetic, as the kernels do).
DVG_DEV void jitter_xy(uint32_t x, uint32_t y, uint32_t sx, uint32_t sy, uint32_t W, uint32_t nsx, uint32_t nsy,
                       uint32_t seed, bool jitter, float &rx, float &ry) {
    rx = 0.5f;
    ry = 0.5f;
    if (jitter) {
        uint32_t idx = ((y * W + x) * nsy + sy) * nsx + sx;
        Pcg32 rng((uint64_t)idx, (uint64_t)seed);
        rx = rng.next_f32();
        ry = rng.next_f32();
    }
}

// sampling.rs:143-181
DVG_DEV float filter_weight(uint32_t type, float dx, float dy, float radius) {
    float adx = absf_(dx), ady = absf_(dy);
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
        float fx = 0.5f * (1.0f - cosf(2.0f * pi * ndx));
        float fy = 0.5f * (1.0f - cosf(2.0f * pi * ndy));
        float norm = 1.0f / (radius * radius);
        return fx * fy * norm;
    }
    return 1.f;
}

}  // namespace dvgk
