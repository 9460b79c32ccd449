// Device scene layout and kernel launch interface shared by the host code and
// dvg_kernels.hip.
#pragma once
#include <algorithm>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dvgk {

constexpr int TILE = 16;            // tile edge in pixels (renderer/constants.rs:21)
constexpr int BLOCK = 256;          // threads per workgroup (4 waves of 64)
constexpr int FRAG_WINDOW = 8;      // fragments held per lane in the adjoint walk
constexpr int LDS_GROUP_CAP = 256;  // tile groups whose paint grads are staged in LDS
constexpr int WIND_STRIDE = 32;     // floats of a segment's winding-term record

// Canvas pixels by which a group's box is padded when it is binned into tiles.
// The prefiltered adjoint reaches one pixel past a boundary (smoothstep
// band); the boundary pass evaluates both query points of a boundary sample,
// each at most 1e-4 of the canvas size from the sampled point, through the
// list of the sampled point's tile.  The pad exceeds both reaches.
inline __host__ __device__ float tile_pad(uint32_t W, uint32_t H) {
    return 1.0f + floorf(1.0e-4f * (float)(W > H ? W : H));
}

// Scene arrays are read through the constant address space on the device:
// they do not change while a render or adjoint kernel runs, so loads at
// wave-uniform addresses (tile lists, group records, curve data of the group
// in hand) become scalar loads whatever the kernel writes elsewhere.  Kernels
// that rebuild scene arrays get writable pointers of their own.  The host
// pass sees plain pointers of the same size (the struct layout is one).
#if defined(__HIP_DEVICE_COMPILE__)
using kfloat = const __attribute__((address_space(4))) float;
using kuint = const __attribute__((address_space(4))) uint32_t;
#else
using kfloat = const float;
using kuint = const uint32_t;
#endif

// Device-resident scene: the packed buffers of PreparedScene /
// PreparedBackwardScene / BoundarySamplingData plus the tile lists.
struct DevScene {
    uint32_t W, H, nsx, nsy, spp, seed, jitter, prefilter, filter_type;
    // launch shapes (log2 edges in pixels): samples of a tile are visited in
    // square blocks of 1 << sblock_log2 pixels; boundary samples are bucketed
    // by blocks of 1 << bnd_log2 pixels
    uint32_t sblock_log2, bnd_log2;
    // tile work is split into units of unit_samples consecutive samples of a
    // tile (a multiple of the workgroup size; >= TILE*TILE*spp = whole tiles)
    uint32_t unit_samples;
    float filter_radius;
    int ri;
    float bg[4];            // straight RGBA
    kfloat *bg_image;  // W*H*4 or null
    // forward pack
    uint32_t num_groups, num_prepared;
    // per-segment winding terms (WIND_STRIDE floats a segment), derived on
    // the device from curve_data
    kfloat *curve_wind;
    kfloat *shape_data, *curve_data, *shape_bounds, *group_data, *group_xform, *group_shape_xform,
        *group_inv_scale, *group_shapes, *shape_xform, *gradient_data, *stop_offsets, *stop_colors;
    // tiles
    uint32_t tiles_x, tiles_y;
    kuint *tile_offsets, *tile_entries;
    // the backward walks' lists: bwd_entries from tile_offsets[t] to
    // tile_end[t], each tile's list without the groups that add nothing to
    // the backward's inside tests there; the forward's: tile_entries from
    // tile_offsets[t] to fwd_end[t], without those of them whose segments
    // the forward's f32 solve settles exactly (k_tile_keep, k_tile_compact)
    kuint *bwd_entries, *tile_end, *fwd_end;
    // backward pack (scene-indexed)
    uint32_t num_shapes;
    kuint *prepared_to_scene;
    kfloat *shape_transform;  // 6*S
    kfloat *path_points;
    kuint *path_ctrl;
    kfloat *path_thickness;
    kuint *sp_off, *sp_cnt, *sc_off, *sc_cnt, *st_off, *st_cnt, *s_closed;
    uint32_t n_cdf;
    kfloat *shape_lengths, *shape_cdf, *shape_pmf;
    kuint *shape_ids, *group_ids;
    kfloat *path_cdf, *path_pmf;
    kuint *path_point_ids, *path_cdf_off, *path_cdf_cnt, *path_point_off;
    kuint *scene_to_prepared;  // S: first prepared instance of each scene shape
    // 4 per group, group-local: the root box of the reference's CPU group BVH
    // (union of distance/shape.rs shape_bounds); min > max when it has none
    kfloat *group_cpu_bounds;
    // 2 per group: the largest stroke radius of its shapes (|stroke width|
    // and |thickness| over their curves) and the largest stroke width, the
    // distance cut-offs of the prefiltered forward and adjoint
    kfloat *group_pf;
    // the pack's group BVHs (renderer/bvh.rs): node boxes (4 floats), nodes
    // [left, skip, start, count], leaf shape indices, meta [node offset, node
    // count, index offset, index count] per group (all zero: no BVH)
    kfloat *gbvh_bounds;
    kuint *gbvh_nodes, *gbvh_indices, *gbvh_meta;
    // no group's shape-to-canvas map shrinks (smallest singular value >= 1):
    // the reference's signed-distance walk then returns the smallest distance
    // over all shapes, which the tiled signed-distance kernel computes
    uint32_t sdf_tiled;
};

// Signed-distance work arrays (derived on the device at every call):
// per prepared shape a canvas box of its boundary (4 floats, padded for
// rounding; min > max when it has none), its similarity scale (0 when its
// shape-to-canvas map is no similarity or it is a distance-approximation
// path) and its group; per curve segment the box of its control points and
// SDF_PIECES stadiums bounding sub-curves of equal parameter length (start
// point and extent, 4 floats; radius and 1 / |extent|^2, 2 floats); per tile
// the candidate radius, the shape nearest the centre and up to SDF_LCAP
// candidate shapes with their count.
constexpr uint32_t SDF_PIECES = 8, SDF_LCAP = 512, SDF_NONE = 0xffffffffu;
struct SdfWork {
    float *cbox, *sim, *cvbox, *piece, *piece_h, *tile_u;
    uint32_t *pgroup, *tile_seed, *tile_count, *tile_list;
};
// The same arrays read through the constant address space.
struct SdfView {
    kfloat *cbox, *sim, *cvbox, *piece, *piece_h, *tile_u;
    kuint *pgroup, *tile_seed, *tile_count, *tile_list;
};
// Bytes of the work arrays for np prepared shapes, nc curves and nt tiles, and
// their carving from one buffer of that size (256-byte aligned pieces).
size_t sdf_work_bytes(uint32_t np, uint32_t nc, uint32_t nt);
SdfWork sdf_work_carve(void *buf, uint32_t np, uint32_t nc, uint32_t nt);

struct DevGrads {
    float *d_shape_params, *d_shape_points, *d_shape_thickness, *d_shape_stroke_width, *d_shape_transform,
        *d_group_transform, *d_group_data, *d_gradient_data, *d_stop_offsets, *d_stop_colors, *d_filter_radius,
        *d_background, *d_background_image, *d_translation;
};

// Kernel launchers (dvg_kernels.hip).  All asynchronous on `stream`.
// The backward's and the forward's tile lists from the full lists in
// entries: the backward's copied to out without the groups that add nothing
// to the backward's inside tests in the tile, ends in tile_end; the
// forward's compacted in place, ends in fwd_end (cull false: every entry
// kept).  cap: the capacity of entries and out (at least
// tile_offsets[ntiles]).
void launch_tile_cull(const DevScene &S, uint32_t *entries, uint32_t *out, uint32_t *tile_end, uint32_t *fwd_end,
                      uint32_t cap, bool cull, hipStream_t stream);
void launch_bin_count(const float *group_aabb, uint32_t G, uint32_t tiles_x, uint32_t tiles_y, uint32_t *counts,
                      hipStream_t stream);
void launch_scan(const uint32_t *counts, uint32_t n, uint32_t *offsets, hipStream_t stream);
void launch_bin_write(const float *group_aabb, uint32_t G, uint32_t tiles_x, uint32_t tiles_y,
                      const uint32_t *offsets, uint32_t *entries, hipStream_t stream);
// Filter weight sums; `cpu` positions samples as the CPU backward does.
void launch_weights(const DevScene &S, float *weight, bool cpu, hipStream_t stream);
void launch_forward(const DevScene &S, const float *weight, float *accum, hipStream_t stream);
void launch_resolve(const DevScene &S, const float *weight, const float *accum, float *out, hipStream_t stream);
// Signed distance (sdf.rs): W*H per-pixel means of the samples, or n points
// (x, y pairs); both need the scene's backward tables (shape_transform,
// prepared_to_scene).
// `work` holds sdf_work_bytes(num_prepared, ncurves, tiles) bytes.
void launch_sdf(const DevScene &S, uint32_t ncurves, void *work, float *out, hipStream_t stream);
void launch_sdf_points(const DevScene &S, uint32_t ncurves, void *work, const float *pos, uint32_t n, float *out,
                       hipStream_t stream);
// Gradients of the signed distances at n points for the adjoint d_sdf (the
// reference CPU's render_backward_positions), accumulated into G.
void launch_sdf_points_backward(const DevScene &S, const float *pos, uint32_t n, const float *d_sdf,
                                const DevGrads &G, hipStream_t stream);
// The d_sdf image term of render_backward (W*H adjoints, one per pixel, at
// every sample), accumulated into G.
void launch_sdf_image_backward(const DevScene &S, const float *d_sdf, const DevGrads &G, hipStream_t stream);
// Bit-set slots of the colour adjoint for tile lists longer than the LDS bit
// set (LDS_MASK_GROUPS entries): `slots` workgroup slots of
// nw = (2 gcap + 31) / 32 words x BLOCK lanes at gmask, and one queue counter.
// gmask null: no slots (long lists then use fragment windows).
constexpr uint32_t LDS_MASK_GROUPS = 128;
// Workgroups of the colour adjoint resident on one CU (7 waves / SIMD; its
// 18 KiB of LDS allow 8).
constexpr uint32_t INTERIOR_WG_PER_CU = 7;
// Longest list given a global bit-set slot (a slot is then 256 KiB x 4).
constexpr uint32_t GMASK_MAX_GROUPS = 16384;
struct InteriorWork {
    uint32_t *gmask, *counter;
    uint32_t gcap, slots;
};
// A slot: the bit set (nw words a lane) then the adjoint's checkpoint ring
// (40 words a lane), both word-major over the BLOCK lanes.
__host__ __device__ inline size_t interior_slot_words(uint32_t gcap) {
    return ((size_t)((2u * gcap + 31u) / 32u) + 40u) * BLOCK;
}
// List capacity of the global bit-set slots for a scene whose longest tile
// list has `list_max` entries (0: the LDS bit set holds every list).  The
// margin leaves room for lists that grow under device-side updates.
inline uint32_t interior_gcap(uint32_t list_max) {
    if (list_max <= LDS_MASK_GROUPS) return 0;
    uint64_t c = (uint64_t)list_max + list_max / 4 + 64;
    c = (c + 63) / 64 * 64;
    return (uint32_t)std::min<uint64_t>(c, GMASK_MAX_GROUPS);
}
void launch_backward_interior(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                              const InteriorWork &iw, hipStream_t stream);
// Prefiltered (analytic coverage) colour adjoint with its distance adjoint.
void launch_backward_prefiltered(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                                 hipStream_t stream);
// Boundary samples are sorted by the pixel block (1 << bnd_log2 pixels a
// side) of their pixel, blocks numbered tile-major, so a wave shares one tile
// group list and a compact footprint.
__host__ __device__ inline uint32_t units_per_tile(const DevScene &S) {
    uint32_t total = TILE * TILE * S.spp;
    return (total + S.unit_samples - 1) / S.unit_samples;
}
__host__ __device__ inline uint32_t boundary_buckets(const DevScene &S) {
    uint32_t per_row = (uint32_t)TILE >> S.bnd_log2;
    return S.tiles_x * S.tiles_y * per_row * per_row;
}
// Work buffers of the boundary pass: keys/ranks/order hold one u32 per
// sample, counts buckets+1, offsets buckets+2, chunk_sums
// scan_chunks(buckets+1).
struct BoundaryWork {
    uint32_t *keys, *ranks, *order, *counts, *offsets, *chunk_sums;
    // Samples whose f32 inside tests are uncertain (a query point within
    // rounding distance of an edge), left by the pass to a redo with f64
    // windings: per unit of BLOCK bucket-ordered samples, a count
    // (flag_count[unit]) and FLAG_UNIT_SLOTS sample indices.
    uint32_t *flag_count, *flag_slots;
};
constexpr uint32_t FLAG_UNIT_SLOTS = 32;
// The bucket offsets come from a scan over many workgroups: SCAN_CHUNK
// counts a workgroup (256 threads x 16).
constexpr uint32_t SCAN_PER_THREAD = 16, SCAN_CHUNK = 256 * SCAN_PER_THREAD;
__host__ __device__ inline uint32_t scan_chunks(uint32_t n) { return (n + SCAN_CHUNK - 1) / SCAN_CHUNK; }
// Boundary samples: tile keys, counting sort; then the tile-ordered evaluation.
void launch_boundary_sort(const DevScene &S, const BoundaryWork &bw, hipStream_t stream);
void launch_boundary_eval(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                          const BoundaryWork &bw, hipStream_t stream);
// The samples launch_boundary_eval left uncertain, with f64 windings (after
// it, in stream order).
void launch_boundary_redo(const DevScene &S, const float *weight, const float *d_render, const DevGrads &G,
                          const BoundaryWork &bw, hipStream_t stream);
// Zero `count` (<= 14) device float arrays of the given lengths.
void launch_zero(float *const *ptrs, const uint32_t *lens, int count, hipStream_t stream);

// ---- device-side scene update (optimisation loops) --------------------------
// Buffers of a device scene that a geometry update rewrites, plus the plan
// computed once at creation.
struct UpdatePlan {
    uint32_t ncurves, n_points, tile_cap;
    const uint32_t *curve_pt;  // 4 per curve: global index of its control points
    const float *inst_pad;     // per prepared instance: padding of its bounds
    float *cpu_inst_bounds;    // per prepared instance: 4 floats scratch
    int32_t *ranges;           // per group: tile range (int4)
    uint32_t *tile_counts;     // per tile
    uint32_t *overflow;        // set when the tile lists outgrow tile_cap
};
// Rewrites the packed geometry of S from path_points (scene layout; NULL =
// unchanged) and group colours (8 per group: fill then stroke RGBA; NULL =
// unchanged), rebuilds tile lists and boundary tables.  Asynchronous.
// Counts into *mismatches (device) the random cubic segment / point pairs whose
// winding from the per-segment record differs from the libm formula.
void launch_selftest_winding(uint32_t n, uint32_t seed, uint32_t *mismatches, hipStream_t stream);
// counts[0..12): per family of random pairs, the checked backward inside
// tests that differ from the reference CPU's f64 winding without a flag, the
// pairs flagged, and the unchecked tests that differ.
void launch_selftest_inside(uint32_t n, uint32_t seed, uint32_t *counts, hipStream_t stream);
// Per-segment winding terms of curve_data (13 floats a segment) into
// curve_wind; after every change of curve_data.
void launch_wind_coef(const float *curve_data, uint32_t ncurves, float *curve_wind, const float *shape_data,
                      uint32_t nprep, hipStream_t stream);
void launch_scene_update(const DevScene &S, const UpdatePlan &P, const float *path_points, const float *group_colors,
                         hipStream_t stream);

// ---- batches: many scenes per launch ---------------------------------------
// One scene of a batch: its device scene and its work buffers.
constexpr int BATCH_ZERO_ARRAYS = 16;  // 14 gradient arrays, accum, counts
struct BatchScene {
    DevScene S;
    float *weight, *accum, *img, *target, *d_render;
    DevGrads G;
    uint32_t *keys, *ranks, *order, *counts, *offsets;
    uint32_t nb;  // boundary buckets
    float *zp[BATCH_ZERO_ARRAYS];
    uint32_t zn[BATCH_ZERO_ARRAYS];
};
// A unit of work: unit `unit` of scene `scene` in some phase.  Zero-phase
// units encode (array << 20) | chunk of ZERO_CHUNK floats.
struct WorkItem {
    uint32_t scene, unit;
};
constexpr uint32_t ZERO_CHUNK = 4096;
enum BatchPhase : int {
    PH_ZERO = 0,
    PH_WEIGHTS_FWD,
    PH_FORWARD,
    PH_FORWARD_PF,  // the forward of prefiltered scenes
    PH_RESOLVE,
    PH_WEIGHTS_BWD,
    PH_INTERIOR,
    PH_PREFILTERED,
    PH_BND_KEY,
    PH_SCAN,
    PH_BND_SCATTER,
    PH_BND_EVAL,
    NUM_PHASES
};
// One persistent launch of `grid` workgroups draining the phase's work queue
// (items[0..n_items), next item = atomicAdd(counter, 1)).  `loss` makes the
// resolve phase also write d_render = 2 (img - target).
void launch_batch_phase(int phase, const BatchScene *scenes, const WorkItem *items, uint32_t n_items,
                        uint32_t *counter, uint32_t grid, bool loss, uint32_t *gmask, uint32_t gcap,
                        hipStream_t stream);
// n words of p set to v (a kernel, so that it can be captured in a graph).
void launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t stream);

}  // namespace dvgk
