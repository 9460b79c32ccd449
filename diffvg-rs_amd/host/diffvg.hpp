// C++ host API over the C ABI of libdvg.so (include/dvg.h), mirroring the
// crate's scene and renderer interface for the render / backward path:
//   scene types      src/scene.rs:1-291, src/geometry.rs:78-89, src/color.rs
//   options          src/renderer/types.rs:10-39, src/backward/render.rs:20-32
//   Renderer         src/renderer/renderer.rs:27-848 (render, render_sdf,
//                    eval_positions, eval_positions_backward,
//                    render_backward with its d_sdf term)
//   gradient tree    src/grad.rs:9-419 (SceneGrad, zeros_from_scene,
//                    accumulate_from)
//   apply_gpu_grads  src/renderer/backward_gpu.rs:599-864
// Names, argument meaning and error kinds follow the crate; errors are
// thrown as RenderError with the crate's kind (InvalidScene or Launch) and
// the library's message.  There is no CPU fallback: the Renderer needs the
// GPU library and a device.
#pragma once

#include <cstdint>
#include <optional>
#include <stdexcept>
#include <string>
#include <variant>
#include <vector>

#include "../../include/dvg.h"

namespace diffvg {

struct Vec2 {
    float x = 0.f, y = 0.f;
};

struct Color {
    float r = 0.f, g = 0.f, b = 0.f, a = 0.f;
};

struct Mat3 {
    float m[3][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    static Mat3 identity();
    // Mat3::inverse (src/math.rs:269-290); nullopt when singular
    std::optional<Mat3> inverse() const;
};

enum class FilterType : uint32_t { Box = 0, Tent = 1, RadialParabolic = 2, Hann = 3 };
enum class FillRule : uint32_t { NonZero = 0, EvenOdd = 1 };
enum class StrokeJoin : uint32_t { Miter = 0, Bevel = 1, Round = 2 };
enum class StrokeCap : uint32_t { Butt = 0, Square = 1, Round = 2 };

struct Filter {
    FilterType filter_type = FilterType::Box;
    float radius = 0.5f;
};

struct GradientStop {
    float offset = 0.f;
    Color color;
};

struct LinearGradient {
    Vec2 start, end;
    std::vector<GradientStop> stops;
};

struct RadialGradient {
    Vec2 center, radius;
    std::vector<GradientStop> stops;
};

// Paint::Solid(Color) | LinearGradient | RadialGradient
using Paint = std::variant<Color, LinearGradient, RadialGradient>;

struct Circle {
    Vec2 center;
    float radius = 0.f;
};

struct Ellipse {
    Vec2 center, radius;
};

struct Rect {
    Vec2 min, max;
};

// diffvg path encoding: num_control_points per segment (0 line, 1 quadratic,
// 2 cubic), the points, optional per-point stroke thickness.
struct Path {
    std::vector<uint8_t> num_control_points;
    std::vector<Vec2> points;
    std::optional<std::vector<float>> thickness;
    bool is_closed = false;
    bool use_distance_approx = false;
};

using ShapeGeometry = std::variant<Circle, Ellipse, Rect, Path>;

struct Shape {
    ShapeGeometry geometry;
    Mat3 transform = Mat3::identity();
    float stroke_width = 0.f;
    StrokeJoin stroke_join = StrokeJoin::Round;
    StrokeCap stroke_cap = StrokeCap::Round;
    float stroke_miter_limit = 4.f;
};

struct ShapeGroup {
    std::vector<uint32_t> shape_indices;
    std::optional<Paint> fill;
    std::optional<Paint> stroke;
    float opacity = 1.f;
    FillRule fill_rule = FillRule::NonZero;
    Mat3 shape_to_canvas = Mat3::identity();
    Mat3 canvas_to_shape = Mat3::identity();
    // sets both transforms (the inverse, or the identity when singular)
    void set_shape_to_canvas(const Mat3 &m);
};

struct Scene {
    uint32_t width = 0, height = 0;
    std::vector<Shape> shapes;
    std::vector<ShapeGroup> groups;
    Color background;  // transparent, as Scene::new (src/scene.rs:280-290)
    std::optional<std::vector<float>> background_image;  // width * height * 4
    Filter filter;
};

// RenderOptions::default (src/renderer/types.rs:27-39)
struct RenderOptions {
    float aa = 1.f;  // carried, never read on this path (as in the crate)
    float path_tolerance = 0.5f;
    uint32_t samples_x = 2, samples_y = 2;
    uint32_t seed = 0;
    bool jitter = true;
    bool use_prefiltering = true;
};

struct BackwardOptions {
    bool compute_translation = false;
};

struct Image {
    uint32_t width = 0, height = 0;
    std::vector<float> pixels;  // RGBA, row-major
};

// ---- gradient tree (src/grad.rs) --------------------------------------------

struct DGradientStop {
    float offset = 0.f;
    Color color;
};

struct DLinearGradient {
    Vec2 start, end;
    std::vector<DGradientStop> stops;
};

struct DRadialGradient {
    Vec2 center, radius;
    std::vector<DGradientStop> stops;
};

using DPaint = std::variant<Color, DLinearGradient, DRadialGradient>;

struct DCircle {
    Vec2 center;
    float radius = 0.f;
};

struct DEllipse {
    Vec2 center, radius;
};

struct DRect {
    Vec2 min, max;
};

struct DPath {
    std::vector<Vec2> points;
    std::optional<std::vector<float>> thickness;
};

using DShapeGeometry = std::variant<DCircle, DEllipse, DRect, DPath>;

struct DShape {
    DShapeGeometry geometry;
    Mat3 transform;
    float stroke_width = 0.f;
};

struct DShapeGroup {
    Mat3 shape_to_canvas;
    std::optional<DPaint> fill;
    std::optional<DPaint> stroke;
};

struct DFilter {
    float radius = 0.f;
};

struct SceneGrad {
    std::vector<DShape> shapes;
    std::vector<DShapeGroup> shape_groups;
    DFilter filter;
    Color background;
    std::optional<std::vector<float>> background_image;
    std::optional<std::vector<float>> translation;

    static SceneGrad zeros_from_scene(const Scene &scene, bool include_background_image, bool include_translation);
    // element-wise add; variable-length parts up to the shorter length, a
    // missing optional part taken from `other`
    void accumulate_from(const SceneGrad &other);
};

class RenderError : public std::runtime_error {
public:
    enum class Kind { InvalidScene, Launch };
    RenderError(Kind kind, const std::string &message);
    Kind kind() const { return kind_; }

private:
    Kind kind_;
};

// The flat description of a Scene (dvg_scene_desc) with the arrays it points
// into.  Gradient paints are numbered in the order the packer meets them
// (each group's fill, then its stroke; src/renderer/prepare.rs:262-273).
class FlatScene {
public:
    explicit FlatScene(const Scene &scene);
    FlatScene(const FlatScene &) = delete;
    FlatScene &operator=(const FlatScene &) = delete;
    const dvg_scene_desc &desc() const { return desc_; }

    std::vector<uint32_t> shape_kind, stroke_join, stroke_cap, path_point_offset, path_point_count,
        path_ctrl_offset, path_ctrl_count, path_thickness_offset, path_thickness_count, path_is_closed,
        path_use_distance_approx, path_num_controls, group_shape_offset, group_shape_count, group_shape_indices,
        fill_kind, stroke_kind, fill_gradient, stroke_gradient, fill_rule, gradient_kind, gradient_stop_offset,
        gradient_stop_count;
    std::vector<float> shape_params, shape_transform, stroke_width, stroke_miter_limit, path_points, path_thickness,
        fill_color, stroke_color, shape_to_canvas, canvas_to_shape, opacity, gradient_params, stop_offsets,
        stop_colors, background_image;

private:
    dvg_scene_desc desc_{};
};

// The 14 flat gradient arrays of the C ABI (include/dvg.h, dvg_grads), sized
// for one scene.
struct FlatGrads {
    FlatGrads(const FlatScene &flat, bool compute_translation);
    dvg_grads view();
    std::vector<float> d_shape_params, d_shape_points, d_shape_thickness, d_shape_stroke_width, d_shape_transform,
        d_group_transform, d_group_data, d_gradient_data, d_stop_offsets, d_stop_colors, d_filter_radius,
        d_background, d_background_image, d_translation;
};

// apply_gpu_grads (src/renderer/backward_gpu.rs:599-864): the flat arrays
// written into `grads` (a SceneGrad sized for `scene`).
void apply_gpu_grads(const Scene &scene, const FlatScene &flat, const FlatGrads &g, SceneGrad &grads);

class Renderer {
public:
    explicit Renderer(int device_ordinal = 0);
    ~Renderer();
    Renderer(const Renderer &) = delete;
    Renderer &operator=(const Renderer &) = delete;

    // checked boundary pass (library default: on)
    void set_exact_boundary(bool on);

    Image render(const Scene &scene, const RenderOptions &options);
    std::vector<float> render_sdf(const Scene &scene, const RenderOptions &options);
    std::vector<float> eval_positions(const Scene &scene, const RenderOptions &options,
                                      const std::vector<Vec2> &positions);
    // Renderer::render_backward: colour adjoint + boundary sampling for
    // d_render (W*H*4), plus the d_sdf term when d_sdf (W*H) is given
    SceneGrad render_backward(const Scene &scene, const RenderOptions &options, const std::vector<float> &d_render,
                              const BackwardOptions &backward_options = {},
                              const std::vector<float> *d_sdf = nullptr);
    SceneGrad eval_positions_backward(const Scene &scene, const RenderOptions &options,
                                      const std::vector<Vec2> &positions, const std::vector<float> &d_sdf,
                                      const BackwardOptions &backward_options = {});
    // the flat arrays of one render_backward call (what apply_gpu_grads reads)
    void render_backward_flat(const Scene &scene, const RenderOptions &options, const std::vector<float> &d_render,
                              const BackwardOptions &backward_options, const FlatScene &flat, FlatGrads &out);

private:
    void check(int32_t status) const;
    dvg_ctx *ctx_ = nullptr;
};

}  // namespace diffvg
