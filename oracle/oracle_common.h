// Oracle: CPU restatement of factal/diffvg-rs's render/backward path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and
// bench.py's cpu_baseline leg may load liboracle.so; it is the checker, never
// the thing measured or shipped.
//
// Two halves, each restating one side of the reference:
//   * forward (oracle_forward.cpp): the reference's GPU forward kernels, the
//     only forward it has (src/gpu/kernels/forward.rs, curves.rs, sampling.rs,
//     math.rs, rng.rs), in f32 with the reference's operation order.  BVH and
//     tile culling are conservative in the reference, so the restatement loops
//     over groups and shapes with the same per-shape bounds tests.
//   * backward (oracle_backward.cpp): the reference CPU backward
//     (src/backward/*.rs over src/distance/*.rs), f64 root solvers included.
//
// Numerical conventions:
//   * built with -ffp-contract=off (no FMA contraction), no fast-math;
//   * an f32 transcendental (acos, cos, sin, atan2, pow) is evaluated in f64
//     and rounded once to f32 (correctly rounded f32);
//   * f64 transcendentals are the C library's.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

#include "../include/dvg.h"

// Operation counting (the liboracle_count.so build, counted.h): COUNT_ON
// switches the tally on or off, COUNT_PHASE picks its bucket.  No-ops in the
// checking build.
#ifdef ORACLE_COUNT
#define COUNT_ON(b) (cnt::on = (b))
#define COUNT_IS_ON() (cnt::on)
#define COUNT_PHASE(p) (cnt::phase = (p))
#else
#define COUNT_ON(b) ((void)0)
#define COUNT_IS_ON() (false)
#define COUNT_PHASE(p) ((void)0)
#endif

namespace oracle {

static const float PI_F = 3.14159265358979323846f;  // core::f32::consts::PI

static inline float acos_f(float x) { return (float)std::acos((double)x); }
static inline float cos_f(float x) { return (float)std::cos((double)x); }
static inline float sin_f(float x) { return (float)std::sin((double)x); }
static inline float atan2_f(float y, float x) { return (float)std::atan2((double)y, (double)x); }
static inline float pow_f(float x, float y) { return (float)std::pow((double)x, (double)y); }
static inline float sqrt_f(float x) { return std::sqrt(x); }

// Rust's f32::min / f32::max ignore a NaN operand, like fmin/fmax.
static inline float rmin(float a, float b) { return std::fmin(a, b); }
static inline float rmax(float a, float b) { return std::fmax(a, b); }
// Rust's f32::clamp.
static inline float rclamp(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// ---------------------------------------------------------------------------
// src/math.rs Vec2 / Vec4 / Mat3
struct V2 {
    float x, y;
    V2() : x(0.f), y(0.f) {}
    V2(float a, float b) : x(a), y(b) {}
    V2 operator+(V2 o) const { return V2(x + o.x, y + o.y); }
    V2 operator-(V2 o) const { return V2(x - o.x, y - o.y); }
    V2 operator*(float s) const { return V2(x * s, y * s); }
    V2 operator/(float s) const { return V2(x / s, y / s); }
    V2 &operator+=(V2 o) { x += o.x; y += o.y; return *this; }
    V2 &operator-=(V2 o) { x -= o.x; y -= o.y; return *this; }
    bool operator==(V2 o) const { return x == o.x && y == o.y; }
    float dot(V2 o) const { return x * o.x + y * o.y; }
    float cross(V2 o) const { return x * o.y - y * o.x; }
    float length_squared() const { return dot(*this); }
    float length() const { return sqrt_f(length_squared()); }
    V2 vmin(V2 o) const { return V2(rmin(x, o.x), rmin(y, o.y)); }
    V2 vmax(V2 o) const { return V2(rmax(x, o.x), rmax(y, o.y)); }
    V2 lerp(V2 o, float t) const { return V2(x + (o.x - x) * t, y + (o.y - y) * t); }
};

struct V4 {
    float x, y, z, w;
    V4() : x(0.f), y(0.f), z(0.f), w(0.f) {}
    V4(float a, float b, float c, float d) : x(a), y(b), z(c), w(d) {}
};

struct M3 {
    float m[3][3];
    static M3 identity() {
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = (i == j) ? 1.f : 0.f;
        return r;
    }
    static M3 zero() {
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = 0.f;
        return r;
    }
    M3 mul(const M3 &o) const {  // math.rs:185-195
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                r.m[i][j] = m[i][0] * o.m[0][j] + m[i][1] * o.m[1][j] + m[i][2] * o.m[2][j];
        return r;
    }
    V2 transform_point(V2 v) const {
        return V2(m[0][0] * v.x + m[0][1] * v.y + m[0][2], m[1][0] * v.x + m[1][1] * v.y + m[1][2]);
    }
    bool is_identity() const {  // math.rs:213-223
        const float e = 1.0e-6f;
        return std::fabs(m[0][0] - 1.f) < e && std::fabs(m[0][1]) < e && std::fabs(m[0][2]) < e &&
               std::fabs(m[1][0]) < e && std::fabs(m[1][1] - 1.f) < e && std::fabs(m[1][2]) < e &&
               std::fabs(m[2][0]) < e && std::fabs(m[2][1]) < e && std::fabs(m[2][2] - 1.f) < e;
    }
    void scale_factors(float &sx, float &sy) const {  // math.rs:248-253
        sx = sqrt_f(m[0][0] * m[0][0] + m[1][0] * m[1][0]);
        sy = sqrt_f(m[0][1] * m[0][1] + m[1][1] * m[1][1]);
    }
    float average_scale() const { float sx, sy; scale_factors(sx, sy); return 0.5f * (sx + sy); }
    float max_scale() const { float sx, sy; scale_factors(sx, sy); return rmax(sx, sy); }
    bool inverse(M3 &out) const {  // math.rs:269-290
        float a = m[0][0], b = m[0][1], c = m[1][0], d = m[1][1], tx = m[0][2], ty = m[1][2];
        float det = a * d - b * c;
        if (std::fabs(det) < 1.0e-6f) return false;
        float inv_det = 1.0f / det;
        float m00 = d * inv_det, m01 = -b * inv_det, m10 = -c * inv_det, m11 = a * inv_det;
        float m02 = -(m00 * tx + m01 * ty);
        float m12 = -(m10 * tx + m11 * ty);
        out = identity();
        out.m[0][0] = m00; out.m[0][1] = m01; out.m[0][2] = m02;
        out.m[1][0] = m10; out.m[1][1] = m11; out.m[1][2] = m12;
        return true;
    }
    M3 inverse_or_identity() const {
        M3 r;
        if (!inverse(r)) r = identity();
        return r;
    }
    M3 transpose() const {
        M3 r;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] = m[j][i];
        return r;
    }
    M3 add(const M3 &o) const {
        M3 r = *this;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) r.m[i][j] += o.m[i][j];
        return r;
    }
};

// ---------------------------------------------------------------------------
// PCG32, src/renderer/rng.rs:16-43 (O'Neill 2014, XSH-RR).
struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t idx, uint64_t seed) {
        state = 0;
        inc = ((idx + 1) << 1) | 1;
        next_u32();
        state = state + (0x853c49e6748fea9bULL + seed);
        next_u32();
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + (inc | 1);
        uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
        uint32_t rot = (uint32_t)(old >> 59);
        return (xorshifted >> rot) | (xorshifted << ((32 - rot) & 31));
    }
    float next_f32() {
        uint32_t u = next_u32();
        uint32_t bits = (u >> 9) | 0x3f800000u;
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};

// ---------------------------------------------------------------------------
// Scene model (src/scene.rs:142-290, src/geometry.rs:78-89), read from the
// flat dvg_scene_desc.
struct Stop { float offset; V4 color; };
struct Paint {
    uint32_t kind = DVG_PAINT_NONE;
    V4 color;        // solid
    bool radial = false;
    V2 p0, p1;       // linear start/end | radial center/radius
    std::vector<Stop> stops;
    uint32_t grad_index = 0;
    uint32_t stop_offset = 0;
    bool present() const { return kind != DVG_PAINT_NONE; }
};
struct Path {
    std::vector<uint8_t> num_control_points;
    std::vector<V2> points;
    bool has_thickness = false;
    std::vector<float> thickness;
    bool is_closed = false;
    bool use_distance_approx = false;
    uint32_t point_offset = 0;
    uint32_t thickness_offset = 0;
};
struct Shape {
    uint32_t kind = 0;
    V2 center;
    float radius_c = 0.f;
    V2 radius_e;
    V2 rmin, rmax;
    Path path;
    M3 transform;
    float stroke_width = 0.f;
    uint32_t join = DVG_JOIN_ROUND, cap = DVG_CAP_ROUND;
    float miter_limit = 4.f;
};
struct Group {
    std::vector<uint32_t> shape_indices;
    Paint fill, stroke;
    uint32_t fill_rule = DVG_FILL_NONZERO;
    M3 shape_to_canvas, canvas_to_shape;
};
struct Scene {
    uint32_t width = 0, height = 0;
    V4 background;
    bool has_bg_image = false;
    std::vector<float> bg_image;
    uint32_t filter_type = DVG_FILTER_BOX;
    float filter_radius = 0.5f;
    std::vector<Shape> shapes;
    std::vector<Group> groups;
    uint32_t num_gradients = 0, num_stops = 0;
    uint64_t path_points_len = 0, path_thickness_len = 0;
};

Scene read_scene(const dvg_scene_desc *d);

// path_utils.rs:7-39
static inline bool path_point(const Path &p, size_t index, V2 &out) {
    size_t total = p.points.size();
    if (total == 0) return false;
    if (p.is_closed) { out = p.points[index % total]; return true; }
    if (index >= total) return false;
    out = p.points[index];
    return true;
}
static inline bool path_point_radius(const Path &p, size_t index, float default_radius, V2 &pt, float &r) {
    if (!path_point(p, index, pt)) return false;
    size_t total = p.points.size();
    size_t ti = p.is_closed ? index % total : index;
    if (p.has_thickness && ti < p.thickness.size()) r = p.thickness[ti] * 1.0f;
    else r = default_radius;
    return true;
}

// One path segment with control points and per-point radii
// (src/renderer/path.rs:100-173, src/distance/bvh.rs:262-380).
struct Seg {
    uint32_t kind;       // 0 line, 1 quadratic, 2 cubic
    uint32_t index;      // segment index in num_control_points
    uint32_t point_id;   // first point index
    V2 p0, p1, p2, p3;
    float r0, r1, r2, r3;
};
// path_to_curve_segments (path.rs:100-173): stops at the first segment whose
// points are missing.
std::vector<Seg> path_segments(const Path &path, float stroke_width);

// Max |radius| over Path::flatten_with_thickness segments (geometry.rs:218-305).
float flattened_max_radius(const Path &path, float tolerance, float default_radius);

}  // namespace oracle
