// Host packer: the C++ side of the seam that the crate's Rust host code fills
// (src/renderer/prepare.rs, prepare_backward.rs, backward/boundary.rs:46-110).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/dvg.h"

struct dvg_prepared {
    // PreparedScene (prepare.rs:20-67)
    std::vector<float> shape_data, segment_data, curve_data, shape_bounds, group_bounds, group_data, group_xform,
        group_shape_xform, group_inv_scale, group_shapes, shape_xform;
    std::vector<float> group_bvh_bounds, path_bvh_bounds;
    std::vector<uint32_t> group_bvh_nodes, group_bvh_indices, group_bvh_meta;
    std::vector<uint32_t> path_bvh_nodes, path_bvh_indices, path_bvh_meta;
    std::vector<float> gradient_data, stop_offsets, stop_colors;
    uint32_t num_groups = 0;
    // PreparedBackwardScene extras (prepare_backward.rs:12-39)
    std::vector<uint32_t> prepared_to_scene;
    std::vector<float> shape_transform, path_points, path_thickness;
    std::vector<uint32_t> path_num_control_points, shape_path_offsets, shape_path_point_counts,
        shape_path_ctrl_offsets, shape_path_ctrl_counts, shape_path_thickness_offsets, shape_path_thickness_counts,
        shape_path_is_closed;
    // BoundarySamplingData (backward/boundary.rs:20-43); empty when absent
    std::vector<float> shape_lengths, shape_cdf, shape_pmf, path_cdf, path_pmf;
    std::vector<uint32_t> shape_ids, group_ids, path_point_ids, path_cdf_offsets, path_cdf_counts,
        path_point_offsets;
    // launch parameters
    dvg_render_params params;
    std::vector<float> background_image;
};

namespace dvgh {
// Returns DVG_OK or DVG_INVALID_SCENE with the reference's message in *err.
int prepare(const dvg_scene_desc *scene, const dvg_render_options *options, dvg_prepared &out, const char **err);
void pack_views(const dvg_prepared &p, dvg_scene_pack *pack, dvg_backward_pack *bpack, dvg_render_params *params);
}  // namespace dvgh
