/*
 * dvg.h -- C ABI of the MI355X (gfx950) render/backward library, libdvg.so.
 *
 * This library replaces the `crate::gpu` seam of factal/diffvg-rs
 * (src/gpu/mod.rs:1-6, src/gpu/kernels/entrypoints.rs:5-28): the CubeCL launch
 * blocks of src/renderer/renderer.rs:213-535 (forward) and
 * src/renderer/backward_gpu.rs:223-488 (backward), which exchange flat packed
 * f32/u32 buffers with fixed strides (src/gpu/constants.rs:4-68).
 *
 * Three layers are exported:
 *
 *  1. Packed seam (what a Rust binding of src/gpu calls).
 *     dvg_render_packed / dvg_render_backward_packed take exactly the buffers
 *     that the crate's host code already builds: PreparedScene
 *     (src/renderer/prepare.rs:20-67), PreparedBackwardScene
 *     (src/renderer/prepare_backward.rs:12-39) and BoundarySamplingData
 *     (src/backward/boundary.rs:20-43), and return the 14 flat gradient
 *     arrays that apply_gpu_grads consumes (src/renderer/backward_gpu.rs:599).
 *
 *  2. Scene API (Renderer::render / Renderer::render_backward).
 *     dvg_render / dvg_render_backward take a flat description of a Scene
 *     (src/scene.rs:142-290) plus RenderOptions (src/renderer/types.rs:10-39),
 *     run the library's packer (a restatement of prepare_scene and friends)
 *     and then the packed seam.  Background gradients come back finalized as
 *     in src/backward/background.rs:29-61.
 *
 *  3. Device-resident scenes (no counterpart in the reference).  A packed
 *     scene is uploaded once; render / backward then read and write device
 *     pointers on the context's stream.  This is what a training loop or the
 *     benchmark uses, so that host<->device copies stay out of the hot path.
 *
 * Status codes: DVG_OK; DVG_INVALID_SCENE carries a static message, readable
 * with dvg_last_error(ctx), with the strings of RenderError::InvalidScene
 * (src/renderer/types.rs:83-88); DVG_DEVICE_ERROR maps RenderError::Launch.
 * No call aborts.  Host-pointer calls are synchronous: outputs are written
 * before return.  Host pointers stay owned by the caller.  One context per host
 * thread; calls on one context are not thread safe.
 */
#ifndef DVG_H
#define DVG_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DVG_OK 0
#define DVG_INVALID_SCENE 1
#define DVG_DEVICE_ERROR 2

/* Enum values: src/gpu/constants.rs:27-56, src/scene.rs:18-107. */
#define DVG_SHAPE_CIRCLE 0u
#define DVG_SHAPE_RECT 1u
#define DVG_SHAPE_PATH 2u
#define DVG_SHAPE_ELLIPSE 3u

#define DVG_PAINT_NONE 0u
#define DVG_PAINT_SOLID 1u
#define DVG_PAINT_LINEAR 2u
#define DVG_PAINT_RADIAL 3u

#define DVG_FILTER_BOX 0u
#define DVG_FILTER_TENT 1u
#define DVG_FILTER_RADIAL_PARABOLIC 2u
#define DVG_FILTER_HANN 3u

#define DVG_FILL_NONZERO 0u
#define DVG_FILL_EVENODD 1u

#define DVG_JOIN_MITER 0u
#define DVG_JOIN_BEVEL 1u
#define DVG_JOIN_ROUND 2u

#define DVG_CAP_BUTT 0u
#define DVG_CAP_SQUARE 1u
#define DVG_CAP_ROUND 2u

/* Packed strides, src/gpu/constants.rs:4-22. */
#define DVG_SHAPE_STRIDE 15u
#define DVG_GROUP_STRIDE 16u
#define DVG_SEGMENT_STRIDE 12u
#define DVG_CURVE_STRIDE 13u
#define DVG_GRADIENT_STRIDE 8u
#define DVG_XFORM_STRIDE 6u
#define DVG_BOUNDS_STRIDE 4u
#define DVG_BVH_NODE_STRIDE 4u
#define DVG_BVH_META_STRIDE 4u
#define DVG_BVH_NONE 0xffffffffu

/* RenderOptions, src/renderer/types.rs:10-39 (`aa` is carried and, as in the
 * reference, never read on this path). */
typedef struct dvg_render_options {
    float aa;
    float path_tolerance;
    uint32_t samples_x;
    uint32_t samples_y;
    uint32_t seed;
    uint32_t jitter;           /* bool */
    uint32_t use_prefiltering; /* bool */
} dvg_render_options;

/* BackwardOptions, src/backward/render.rs:20-32. */
typedef struct dvg_backward_options {
    uint32_t compute_translation; /* bool */
} dvg_backward_options;

/*
 * Flat description of a Scene (src/scene.rs:142-290).
 *
 * Shapes (S, scene order):
 *   shape_kind[S]          DVG_SHAPE_*
 *   shape_params[4*S]      circle: cx cy r 0 | ellipse: cx cy rx ry |
 *                          rect: minx miny maxx maxy | path: unused
 *   shape_transform[9*S]   Shape.transform, row-major 3x3
 *   stroke_width[S], stroke_join[S], stroke_cap[S], stroke_miter_limit[S]
 *   path_point_offset[S], path_point_count[S]   points, into path_points/2
 *   path_ctrl_offset[S], path_ctrl_count[S]     segments, into path_num_controls
 *   path_thickness_offset[S], path_thickness_count[S]  count 0 = no thickness
 *   path_is_closed[S], path_use_distance_approx[S]
 * Path pools: path_points[2*P] (x, y), path_num_controls[C] (0/1/2),
 *             path_thickness[T].
 *
 * Groups (G, draw order):
 *   group_shape_offset[G], group_shape_count[G] into group_shape_indices[]
 *   fill_kind[G], stroke_kind[G]   DVG_PAINT_*
 *   fill_color[4*G], stroke_color[4*G]   RGBA of solid paints
 *   fill_gradient[G], stroke_gradient[G] index into the gradient table
 *   fill_rule[G]
 *   shape_to_canvas[9*G], canvas_to_shape[9*G]   row-major 3x3
 *   opacity[G]  ShapeGroup.opacity (carried; the render path never reads it)
 * Gradients (NG): gradient_kind[NG] (0 linear, 1 radial),
 *   gradient_params[4*NG] (start.xy end.xy | center.xy radius.xy),
 *   gradient_stop_offset[NG], gradient_stop_count[NG] into
 *   stop_offsets[NS], stop_colors[4*NS].
 */
typedef struct dvg_scene_desc {
    uint32_t width;
    uint32_t height;
    float background[4];           /* straight RGBA */
    const float *background_image; /* NULL, or background_image_len floats */
    uint64_t background_image_len;
    uint32_t filter_type;
    float filter_radius;

    uint32_t num_shapes;
    const uint32_t *shape_kind;
    const float *shape_params;
    const float *shape_transform;
    const float *stroke_width;
    const uint32_t *stroke_join;
    const uint32_t *stroke_cap;
    const float *stroke_miter_limit;
    const uint32_t *path_point_offset;
    const uint32_t *path_point_count;
    const uint32_t *path_ctrl_offset;
    const uint32_t *path_ctrl_count;
    const uint32_t *path_thickness_offset;
    const uint32_t *path_thickness_count;
    const uint32_t *path_is_closed;
    const uint32_t *path_use_distance_approx;
    const float *path_points;
    uint64_t path_points_len; /* floats */
    const uint32_t *path_num_controls;
    uint64_t path_num_controls_len;
    const float *path_thickness;
    uint64_t path_thickness_len;

    uint32_t num_groups;
    const uint32_t *group_shape_offset;
    const uint32_t *group_shape_count;
    const uint32_t *group_shape_indices;
    uint64_t group_shape_indices_len;
    const uint32_t *fill_kind;
    const uint32_t *stroke_kind;
    const float *fill_color;
    const float *stroke_color;
    const uint32_t *fill_gradient;
    const uint32_t *stroke_gradient;
    const uint32_t *fill_rule;
    const float *shape_to_canvas;
    const float *canvas_to_shape;
    const float *opacity;

    uint32_t num_gradients;
    const uint32_t *gradient_kind;
    const float *gradient_params;
    const uint32_t *gradient_stop_offset;
    const uint32_t *gradient_stop_count;
    uint32_t num_stops;
    const float *stop_offsets;
    const float *stop_colors;
} dvg_scene_desc;

/*
 * The 14 gradient arrays of src/renderer/backward_gpu.rs:187-200, laid out as
 * apply_gpu_grads reads them (backward_gpu.rs:599-801).  Lengths:
 *   d_shape_params       8*S   circle cx cy r | ellipse cx cy rx ry | rect min max
 *   d_shape_points       len(path_points)   same indexing as path_points
 *   d_shape_thickness    len(path_thickness)
 *   d_shape_stroke_width S
 *   d_shape_transform    9*S   row-major 3x3
 *   d_group_transform    9*G   shape_to_canvas, row-major 3x3
 *   d_group_data         16*G  fill RGBA at +8, stroke RGBA at +12
 *   d_gradient_data      8*NG  p0..p3 at +1..+4
 *   d_stop_offsets       NS
 *   d_stop_colors        4*NS
 *   d_filter_radius      1
 *   d_background         4
 *   d_background_image   W*H*4 (NULL when the scene has no background image)
 *   d_translation        W*H*2 (NULL unless compute_translation)
 * The library zeroes them and accumulates into them.
 */
typedef struct dvg_grads {
    float *d_shape_params;
    float *d_shape_points;
    float *d_shape_thickness;
    float *d_shape_stroke_width;
    float *d_shape_transform;
    float *d_group_transform;
    float *d_group_data;
    float *d_gradient_data;
    float *d_stop_offsets;
    float *d_stop_colors;
    float *d_filter_radius;
    float *d_background;
    float *d_background_image;
    float *d_translation;
} dvg_grads;

/* PreparedScene, src/renderer/prepare.rs:20-67.  Indices into shape_data are
 * prepared indices: one per entry of the groups' shape lists, in group order
 * (prepare.rs:186-205).  segment_data and the path BVHs are accepted as the
 * reference fills them; the render and adjoint kernels cull with their own
 * tile lists.  The group BVHs (renderer/bvh.rs layout) are checked at upload
 * (DVG_INVALID_SCENE "invalid group BVH node" etc.) and walked by the signed
 * distance; a null group_bvh_meta, or a group's zero meta, is no BVH. */
typedef struct dvg_scene_pack {
    uint32_t num_groups;
    uint32_t num_prepared_shapes;
    const float *shape_data;          /* 15 * num_prepared_shapes */
    const float *segment_data;        uint64_t segment_data_len;
    const float *curve_data;          uint64_t curve_data_len;     /* 13 * Nc */
    const float *shape_bounds;        /* 4 * num_prepared_shapes */
    const float *group_bounds;        /* 4 * G */
    const float *group_data;          /* 16 * G */
    const float *group_xform;         /* 6 * G  canvas_to_shape */
    const float *group_shape_xform;   /* 6 * G  shape_to_canvas */
    const float *group_inv_scale;     /* G */
    const float *group_shapes;        uint64_t group_shapes_len;
    const float *shape_xform;         /* 6 * num_prepared_shapes, inverse */
    const float *group_bvh_bounds;    uint64_t group_bvh_bounds_len;
    const uint32_t *group_bvh_nodes;  uint64_t group_bvh_nodes_len;
    const uint32_t *group_bvh_indices; uint64_t group_bvh_indices_len;
    const uint32_t *group_bvh_meta;   /* 4 * G */
    const float *path_bvh_bounds;     uint64_t path_bvh_bounds_len;
    const uint32_t *path_bvh_nodes;   uint64_t path_bvh_nodes_len;
    const uint32_t *path_bvh_indices; uint64_t path_bvh_indices_len;
    const uint32_t *path_bvh_meta;    /* 4 * num_prepared_shapes */
    const float *gradient_data;       uint64_t gradient_data_len;
    const float *stop_offsets;        uint64_t stop_offsets_len;
    const float *stop_colors;         /* 4 * stop_offsets_len */
} dvg_scene_pack;

/* PreparedBackwardScene extras (prepare_backward.rs:12-39; per-shape arrays
 * in scene order) and BoundarySamplingData (backward/boundary.rs:20-43).
 * Prepared shapes are one per (group, shape) instance in group order, the
 * same order as the shape CDF entries (prepare.rs:186-205,
 * backward/boundary.rs:312-324); prepared_to_scene maps them to scene shapes.
 * shape_cdf_len == 0 disables boundary sampling. */
typedef struct dvg_backward_pack {
    uint32_t num_shapes;                    /* S (scene shapes) */
    const uint32_t *prepared_to_scene;      /* num_prepared_shapes */
    const float *shape_transform;           /* 6 * S, Shape.transform rows 0-1 */
    const float *path_points;               uint64_t path_points_len;
    const uint32_t *path_num_control_points; uint64_t path_num_control_points_len;
    const float *path_thickness;            uint64_t path_thickness_len;
    const uint32_t *shape_path_offsets;
    const uint32_t *shape_path_point_counts;
    const uint32_t *shape_path_ctrl_offsets;
    const uint32_t *shape_path_ctrl_counts;
    const uint32_t *shape_path_thickness_offsets;
    const uint32_t *shape_path_thickness_counts;
    const uint32_t *shape_path_is_closed;
    /* boundary sampling tables */
    const float *shape_lengths;             /* S */
    const float *shape_cdf;                 uint64_t shape_cdf_len;
    const float *shape_pmf;
    const uint32_t *shape_ids;
    const uint32_t *group_ids;
    const float *path_cdf;                  uint64_t path_cdf_len;
    const float *path_pmf;
    const uint32_t *path_point_ids;         uint64_t path_point_ids_len;
    const uint32_t *path_cdf_offsets;       /* S */
    const uint32_t *path_cdf_counts;        /* S */
    const uint32_t *path_point_offsets;     /* S */
} dvg_backward_pack;

/* Launch parameters of renderer.rs:146-156 / backward_gpu.rs:66-88. */
typedef struct dvg_render_params {
    uint32_t width;
    uint32_t height;
    uint32_t samples_x;
    uint32_t samples_y;
    uint32_t seed;
    uint32_t jitter;           /* effective flag: 0 when prefiltering */
    uint32_t use_prefiltering;
    uint32_t filter_type;
    float filter_radius;
    float background[4];       /* straight RGBA */
    const float *background_image; /* NULL or W*H*4 */
    uint32_t compute_translation;
} dvg_render_params;

/* ---- context -------------------------------------------------------------- */

typedef struct dvg_ctx dvg_ctx;

/* Replaces WgpuDevice / the CubeCL client (renderer.rs:21-37).  Owns a HIP
 * stream and device work buffers.  NULL when no device is usable. */
dvg_ctx *dvg_ctx_create(int device_ordinal);
void dvg_ctx_destroy(dvg_ctx *ctx);
/* Static message of the last non-OK status on this context. */
const char *dvg_last_error(const dvg_ctx *ctx);
/* Run all later work of this context on a caller stream (hipStream_t passed
 * as void*; NULL is the device's null stream).  Until called, the context
 * uses a non-blocking stream of its own. */
int32_t dvg_ctx_set_stream(dvg_ctx *ctx, void *hip_stream);
/* Checked boundary pass (no counterpart in the reference; on by default).
 * The single-scene backward passes (dvg_render_backward, the packed and
 * device-scene forms; not the batch) flag every boundary sample whose f32
 * inside test the f64 winding of distance/winding.rs could disagree with,
 * and evaluate those samples again in f64 (on a second stream of the
 * context, beside the colour adjoint, for the single-scene forms).  The
 * boundary gradients then follow the reference CPU sample for sample.
 * dvg_ctx_set_exact_boundary(ctx, 0) selects the unchecked f32 pass: about
 * 15 % faster on 1024^2 with 1000 cubic paths, but a sample whose query point
 * lies within f32 rounding of an edge can take the other side, and each such
 * sample carries a 1 / pdf weight (whole-image gradients off by up to ~2e-3
 * relative on adjoints that weight every edge). */
int32_t dvg_ctx_set_exact_boundary(dvg_ctx *ctx, int enable);

/* ---- scene API ------------------------------------------------------------ */

/* Renderer::render (renderer.rs:39-545): width*height*4 straight RGBA. */
int32_t dvg_render(dvg_ctx *ctx, const dvg_scene_desc *scene,
                   const dvg_render_options *options, float *out_rgba);

/* Renderer::render_sdf (renderer.rs:586-727 -> render_sdf_kernel,
 * src/gpu/kernels/sdf.rs:119-200): width*height signed distances, each the
 * mean over the pixel's samples (the forward's sample positions) of the
 * distance to the nearest shape boundary, negative inside the fill of that
 * shape's group.  All zero for a scene without groups. */
int32_t dvg_render_sdf(dvg_ctx *ctx, const dvg_scene_desc *scene,
                       const dvg_render_options *options, float *out_sdf);

/* Renderer::eval_positions (renderer.rs:729-848 -> eval_positions_kernel,
 * sdf.rs:207-262): the signed distance at `count` canvas points
 * (positions: x, y pairs). */
int32_t dvg_eval_positions(dvg_ctx *ctx, const dvg_scene_desc *scene,
                           const dvg_render_options *options,
                           const float *positions, uint64_t count, float *out);

/* Renderer::eval_positions_backward (renderer.rs:567-583 ->
 * render_backward_positions, src/backward/render.rs:213-280, which the
 * reference runs on the CPU): the gradients of the signed distances at
 * `count` points for the adjoint d_sdf (count floats), into the flat
 * gradient layout of dvg_render_backward (zeroed first; d_translation, when
 * requested, at the pixel holding each point). */
int32_t dvg_eval_positions_backward(dvg_ctx *ctx, const dvg_scene_desc *scene,
                                    const dvg_render_options *options,
                                    const dvg_backward_options *backward_options,
                                    const float *positions, uint64_t count,
                                    const float *d_sdf, const dvg_grads *grads);

/* The d_sdf term of Renderer::render_backward (renderer.rs:548 ->
 * backward_gpu.rs:512-521, which adds the CPU's render_backward with d_sdf
 * only: backward/render.rs:171-186): for a W*H adjoint of the signed-distance
 * image, sample_distance at every sample of every pixel with that pixel's
 * d_sdf, into the flat gradient layout (zeroed first).  A caller with both
 * adjoints adds these gradients to dvg_render_backward's. */
int32_t dvg_render_backward_sdf(dvg_ctx *ctx, const dvg_scene_desc *scene,
                                const dvg_render_options *options,
                                const dvg_backward_options *backward_options,
                                const float *d_sdf, const dvg_grads *grads);

/* Renderer::render_backward (renderer.rs:548 -> backward_gpu.rs:21-523) with
 * d_render present and d_sdf absent. */
int32_t dvg_render_backward(dvg_ctx *ctx, const dvg_scene_desc *scene,
                            const dvg_render_options *options,
                            const dvg_backward_options *backward_options,
                            const float *d_render, const dvg_grads *grads);

/* ---- packed seam ---------------------------------------------------------- */

/* rasterize_weights_f32 -> binning -> rasterize_splat_f32 -> resolve_splat_f32
 * of renderer.rs:213-537, on packed host buffers. */
int32_t dvg_render_packed(dvg_ctx *ctx, const dvg_scene_pack *pack,
                          const dvg_render_params *params, float *out_rgba);

/* weights -> colour backward -> boundary sampling of backward_gpu.rs:223-488,
 * on packed host buffers; gradients are raw (before
 * finalize_background_gradients), indexed by scene shape/group. */
int32_t dvg_render_backward_packed(dvg_ctx *ctx, const dvg_scene_pack *pack,
                                   const dvg_backward_pack *bpack,
                                   const dvg_render_params *params,
                                   const float *d_render, const dvg_grads *grads);

/* ---- packer (host side of the seam) --------------------------------------- */

typedef struct dvg_prepared dvg_prepared;
/* prepare_scene + prepare_scene_backward + build_boundary_sampling_data
 * (prepare.rs:114, prepare_backward.rs:45, backward/boundary.rs:46).  On
 * DVG_INVALID_SCENE *error_message is the reference's message. */
int32_t dvg_prepare(const dvg_scene_desc *scene, const dvg_render_options *options,
                    dvg_prepared **out, const char **error_message);
void dvg_prepared_free(dvg_prepared *p);
/* Views into the prepared buffers, valid until dvg_prepared_free. */
void dvg_prepared_pack(const dvg_prepared *p, dvg_scene_pack *pack,
                       dvg_backward_pack *bpack, dvg_render_params *params);

/* ---- device-resident scenes (no counterpart in the reference) ------------- */

typedef struct dvg_device_scene dvg_device_scene;
/* Upload a packed scene to device memory owned by the returned object. */
int32_t dvg_device_scene_create(dvg_ctx *ctx, const dvg_scene_pack *pack,
                                const dvg_backward_pack *bpack,
                                const dvg_render_params *params,
                                dvg_device_scene **out);
void dvg_device_scene_destroy(dvg_device_scene *scene);
/* Forward into a device buffer of W*H*4 floats; asynchronous on the stream. */
int32_t dvg_device_render(dvg_ctx *ctx, dvg_device_scene *scene, float *out_rgba_dev);
/* dvg_render_sdf / dvg_eval_positions on a device scene (created with its
 * backward pack) into device buffers; asynchronous on the stream. */
int32_t dvg_device_render_sdf(dvg_ctx *ctx, dvg_device_scene *scene, float *out_sdf_dev);
int32_t dvg_device_eval_positions(dvg_ctx *ctx, dvg_device_scene *scene, const float *positions_dev,
                                  uint64_t count, float *out_dev);
/* dvg_eval_positions_backward on a device scene: device positions, d_sdf
 * and gradient arrays (zeroed, then accumulated); asynchronous. */
int32_t dvg_device_eval_positions_backward(dvg_ctx *ctx, dvg_device_scene *scene,
                                           const float *positions_dev, uint64_t count,
                                           const float *d_sdf_dev, const dvg_grads *grads_dev);
/* dvg_render_backward_sdf on a device scene (W*H device adjoints). */
int32_t dvg_device_render_backward_sdf(dvg_ctx *ctx, dvg_device_scene *scene, const float *d_sdf_dev,
                                       const dvg_grads *grads_dev);
/* Backward from a device d_render (W*H*4) into device gradient arrays (raw
 * accumulations, as dvg_render_backward_packed); asynchronous on the stream. */
int32_t dvg_device_render_backward(dvg_ctx *ctx, dvg_device_scene *scene,
                                   const float *d_render_dev, const dvg_grads *grads_dev);
/* Block until the context's stream is idle. */
int32_t dvg_sync(dvg_ctx *ctx);

/* dvg_device_render without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * kernels as dvg_device_render; its weight and accumulation images live in a
 * static device array of DVG_NOALLOC_MAX_PIXELS * 5 floats (20 MiB), held by
 * each device the library runs on until the process ends.  Scenes of more
 * than DVG_NOALLOC_MAX_PIXELS pixels are rejected with DVG_INVALID_SCENE.  All
 * calls on one device share that device's array: calls that overlap on one
 * device (several streams, contexts or threads) race.  Asynchronous on the
 * context's stream, like dvg_device_render. */
#define DVG_NOALLOC_MAX_PIXELS (1024u * 1024u)
int32_t dvg_device_render_noalloc(dvg_ctx *ctx, dvg_device_scene *scene, float *out_rgba_dev);

/* Device-side re-packing for optimisation loops (the per-call prepare_scene
 * of renderer.rs:39 and build_boundary_sampling_data of
 * backward/boundary.rs:46, restated on the GPU for the inputs that change):
 * path_points (device, scene layout of dvg_backward_pack.path_points, or
 * NULL) and group_colors (device, 8 floats per group: fill RGBA then stroke
 * RGBA, or NULL) replace the scene's; curves, bounds, tile lists and boundary
 * tables follow.  Asynchronous on the context's stream.  The scene must have
 * been created with its backward tables. */
int32_t dvg_device_scene_update(dvg_ctx *ctx, dvg_device_scene *scene, const float *path_points,
                                const float *group_colors);
/* Synchronizes and reports DVG_DEVICE_ERROR if an update made the tile
 * lists outgrow their buffer (lists were cut short: recreate the scene). */
int32_t dvg_device_scene_check(dvg_ctx *ctx, dvg_device_scene *scene);

/* Launch shapes of one device scene, as in dvg_batch_config (sample blocks
 * 1 << sblock_log2 pixels, boundary buckets 1 << bnd_log2 pixels, tile work
 * in units of 256 << unit_log2 samples); results do not depend on them
 * beyond the order of float additions.  Defaults: 2, 1, 0. */
int32_t dvg_device_scene_set_shapes(dvg_device_scene *scene, uint32_t sblock_log2, uint32_t bnd_log2,
                                    uint32_t unit_log2);

/* ---- kernel timing (no counterpart in the reference) ---------------------- */

/* Kernel ids of dvg_kernel_times. */
#define DVG_KERNEL_ZERO 0
#define DVG_KERNEL_WEIGHTS 1
#define DVG_KERNEL_FORWARD 2
#define DVG_KERNEL_RESOLVE 3
#define DVG_KERNEL_INTERIOR 4
#define DVG_KERNEL_BOUNDARY_SORT 5
#define DVG_KERNEL_BOUNDARY_EVAL 6
#define DVG_NUM_KERNELS 7

/* When enabled, every later launch of the context is bracketed by HIP events
 * on its stream (the tracing hook the reference lacks). */
int32_t dvg_ctx_set_timing(dvg_ctx *ctx, int enable);
/* Synchronizes the stream, then reports per kernel id the accumulated
 * milliseconds and launch counts (arrays of n); reset != 0 clears them. */
int32_t dvg_kernel_times(dvg_ctx *ctx, double *total_ms, uint32_t *calls, int n, int reset);

/* ---- batches of scenes --------------------------------------------------- */
/* Many scenes (uneven sizes and modes) stepped together: forward, L2 loss
 * gradient and backward with one persistent launch per phase over a device
 * work queue of (scene, tile | sample block) items, heaviest tiles first.
 * This is the batched step of the scope (dvg_batch_step); the reference runs
 * one scene per call (renderer.rs:39, backward_gpu.rs:21) and has no
 * counterpart.  Everything lives in HBM; the caller owns the target images. */
typedef struct dvg_batch dvg_batch;

#define DVG_BATCH_FORWARD 1u  /* weights, forward, resolve into the batch images */
#define DVG_BATCH_LOSS 2u     /* d_render = 2 (image - target) (needs FORWARD) */
#define DVG_BATCH_BACKWARD 4u /* gradients of every scene from its d_render */

/* Launch shapes of a batch (all scenes): sample blocks of 1 << sblock_log2
 * pixels a side inside a tile (0..4), boundary buckets of 1 << bnd_log2
 * pixels (1..4), persistent workgroups per compute unit (1..64), tile work
 * split into units of 256 << unit_log2 samples (0..8; a unit never spans two
 * tiles); ms is the step time the autotuner measured for them (0 when not
 * measured). */
typedef struct dvg_batch_config {
    uint32_t sblock_log2;
    uint32_t bnd_log2;
    uint32_t wg_per_cu;
    uint32_t unit_log2;
    float ms;
} dvg_batch_config;

/* Uploads n scenes (packs as for dvg_device_scene_create) and allocates their
 * work buffers.  targets: n device pointers to W*H*4 target images (or NULL
 * when the loss is never asked for); they must outlive the batch.  The batch
 * runs the unchecked f32 boundary pass (see dvg_ctx_set_exact_boundary),
 * whatever the context's mode. */
int32_t dvg_batch_create(dvg_ctx *ctx, uint32_t n, const dvg_scene_pack *packs, const dvg_backward_pack *bpacks,
                         const dvg_render_params *params, const float *const *targets, dvg_batch **out);
void dvg_batch_destroy(dvg_batch *batch);
/* Device pointers of scene i's image, loss gradient (d_render, writable: the
 * caller may supply its own before a BACKWARD-only step) and gradients. */
int32_t dvg_batch_buffers(dvg_batch *batch, uint32_t i, float **img, float **d_render, dvg_grads *grads);
/* Copies scene i's image (W*H*4, may be NULL) and gradient arrays (NULL
 * members skipped) to host or device memory, then synchronizes. */
int32_t dvg_batch_read(dvg_ctx *ctx, dvg_batch *batch, uint32_t i, float *img, const dvg_grads *grads);
/* One asynchronous step on the context's stream; flags = DVG_BATCH_*. */
int32_t dvg_batch_step(dvg_ctx *ctx, dvg_batch *batch, uint32_t flags);
/* enable != 0: the step is captured once per flag set in a HIP graph and
 * replayed (the context's stream must not be the null stream). */
int32_t dvg_batch_set_graph(dvg_batch *batch, int enable);
int32_t dvg_batch_get_config(const dvg_batch *batch, dvg_batch_config *cfg);
int32_t dvg_batch_set_config(dvg_ctx *ctx, dvg_batch *batch, const dvg_batch_config *cfg);
/* Measures candidate launch shapes on this batch (reps timed steps each,
 * coordinate search) and keeps the fastest; best may be NULL. */
int32_t dvg_batch_autotune(dvg_ctx *ctx, dvg_batch *batch, uint32_t flags, int reps, dvg_batch_config *best);

/* ---- diagnostics (no counterpart in the reference) ------------------------ */

/* Self-test of the cubic winding solve: n random segment / point pairs (PCG32
 * stream `seed`, points inside each segment's y-range) are evaluated through
 * the per-segment record with its fast transcendentals and through the
 * curves.rs formula with libm's acos / cos / pow; *mismatches receives the
 * number of pairs whose winding differs.  Synchronous. */
int32_t dvg_selftest_winding(dvg_ctx *ctx, uint32_t n, uint32_t seed, uint32_t *mismatches);

/* Self-test of the backward's cubic inside test, which follows the
 * reference CPU's f64 winding (distance/winding.rs:166-188): n random
 * segment / point pairs in four families (random; near-quadratic cubics;
 * points near a y-extremum; points within 0.01 px of the curve) against an
 * f64 restatement of that formula.  counts[f] (f < 4) receives the pairs of
 * family f that differ without being flagged uncertain, counts[4 + f] the
 * pairs flagged (the exact boundary mode evaluates those again in f64),
 * counts[8 + f] the pairs where the default boundary pass's test differs
 * (12 counts).  Synchronous. */
int32_t dvg_selftest_inside(dvg_ctx *ctx, uint32_t n, uint32_t seed, uint32_t *counts);

/* Version string of the library build. */
const char *dvg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* DVG_H */
