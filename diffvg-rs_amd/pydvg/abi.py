"""ctypes mirror of include/dvg.h (the C ABI of libdvg.so).

Field order and types follow the header exactly; every struct here is a
plain C struct of pointers, sizes and scalars.
"""
import ctypes as C
import os

_u32p = C.POINTER(C.c_uint32)
_f32p = C.POINTER(C.c_float)

DVG_OK = 0
DVG_INVALID_SCENE = 1
DVG_DEVICE_ERROR = 2
DVG_NOALLOC_MAX_PIXELS = 1024 * 1024

SHAPE_CIRCLE, SHAPE_RECT, SHAPE_PATH, SHAPE_ELLIPSE = 0, 1, 2, 3
PAINT_NONE, PAINT_SOLID, PAINT_LINEAR, PAINT_RADIAL = 0, 1, 2, 3
FILTER_BOX, FILTER_TENT, FILTER_RADIAL_PARABOLIC, FILTER_HANN = 0, 1, 2, 3
FILL_NONZERO, FILL_EVENODD = 0, 1
JOIN_MITER, JOIN_BEVEL, JOIN_ROUND = 0, 1, 2
CAP_BUTT, CAP_SQUARE, CAP_ROUND = 0, 1, 2


class RenderOptions(C.Structure):
    _fields_ = [("aa", C.c_float), ("path_tolerance", C.c_float), ("samples_x", C.c_uint32),
                ("samples_y", C.c_uint32), ("seed", C.c_uint32), ("jitter", C.c_uint32),
                ("use_prefiltering", C.c_uint32)]


class BackwardOptions(C.Structure):
    _fields_ = [("compute_translation", C.c_uint32)]


class SceneDescC(C.Structure):
    _fields_ = [
        ("width", C.c_uint32), ("height", C.c_uint32), ("background", C.c_float * 4),
        ("background_image", _f32p), ("background_image_len", C.c_uint64),
        ("filter_type", C.c_uint32), ("filter_radius", C.c_float),
        ("num_shapes", C.c_uint32), ("shape_kind", _u32p), ("shape_params", _f32p),
        ("shape_transform", _f32p), ("stroke_width", _f32p), ("stroke_join", _u32p),
        ("stroke_cap", _u32p), ("stroke_miter_limit", _f32p),
        ("path_point_offset", _u32p), ("path_point_count", _u32p),
        ("path_ctrl_offset", _u32p), ("path_ctrl_count", _u32p),
        ("path_thickness_offset", _u32p), ("path_thickness_count", _u32p),
        ("path_is_closed", _u32p), ("path_use_distance_approx", _u32p),
        ("path_points", _f32p), ("path_points_len", C.c_uint64),
        ("path_num_controls", _u32p), ("path_num_controls_len", C.c_uint64),
        ("path_thickness", _f32p), ("path_thickness_len", C.c_uint64),
        ("num_groups", C.c_uint32), ("group_shape_offset", _u32p), ("group_shape_count", _u32p),
        ("group_shape_indices", _u32p), ("group_shape_indices_len", C.c_uint64),
        ("fill_kind", _u32p), ("stroke_kind", _u32p), ("fill_color", _f32p), ("stroke_color", _f32p),
        ("fill_gradient", _u32p), ("stroke_gradient", _u32p), ("fill_rule", _u32p),
        ("shape_to_canvas", _f32p), ("canvas_to_shape", _f32p), ("opacity", _f32p),
        ("num_gradients", C.c_uint32), ("gradient_kind", _u32p), ("gradient_params", _f32p),
        ("gradient_stop_offset", _u32p), ("gradient_stop_count", _u32p),
        ("num_stops", C.c_uint32), ("stop_offsets", _f32p), ("stop_colors", _f32p),
    ]


class GradsC(C.Structure):
    _fields_ = [(n, _f32p) for n in (
        "d_shape_params", "d_shape_points", "d_shape_thickness", "d_shape_stroke_width",
        "d_shape_transform", "d_group_transform", "d_group_data", "d_gradient_data",
        "d_stop_offsets", "d_stop_colors", "d_filter_radius", "d_background",
        "d_background_image", "d_translation")]


class ScenePackC(C.Structure):
    _fields_ = [
        ("num_groups", C.c_uint32), ("num_prepared_shapes", C.c_uint32),
        ("shape_data", _f32p), ("segment_data", _f32p), ("segment_data_len", C.c_uint64),
        ("curve_data", _f32p), ("curve_data_len", C.c_uint64),
        ("shape_bounds", _f32p), ("group_bounds", _f32p), ("group_data", _f32p),
        ("group_xform", _f32p), ("group_shape_xform", _f32p), ("group_inv_scale", _f32p),
        ("group_shapes", _f32p), ("group_shapes_len", C.c_uint64), ("shape_xform", _f32p),
        ("group_bvh_bounds", _f32p), ("group_bvh_bounds_len", C.c_uint64),
        ("group_bvh_nodes", _u32p), ("group_bvh_nodes_len", C.c_uint64),
        ("group_bvh_indices", _u32p), ("group_bvh_indices_len", C.c_uint64),
        ("group_bvh_meta", _u32p),
        ("path_bvh_bounds", _f32p), ("path_bvh_bounds_len", C.c_uint64),
        ("path_bvh_nodes", _u32p), ("path_bvh_nodes_len", C.c_uint64),
        ("path_bvh_indices", _u32p), ("path_bvh_indices_len", C.c_uint64),
        ("path_bvh_meta", _u32p),
        ("gradient_data", _f32p), ("gradient_data_len", C.c_uint64),
        ("stop_offsets", _f32p), ("stop_offsets_len", C.c_uint64), ("stop_colors", _f32p),
    ]


class BackwardPackC(C.Structure):
    _fields_ = [
        ("num_shapes", C.c_uint32), ("prepared_to_scene", _u32p), ("shape_transform", _f32p),
        ("path_points", _f32p), ("path_points_len", C.c_uint64),
        ("path_num_control_points", _u32p), ("path_num_control_points_len", C.c_uint64),
        ("path_thickness", _f32p), ("path_thickness_len", C.c_uint64),
        ("shape_path_offsets", _u32p), ("shape_path_point_counts", _u32p),
        ("shape_path_ctrl_offsets", _u32p), ("shape_path_ctrl_counts", _u32p),
        ("shape_path_thickness_offsets", _u32p), ("shape_path_thickness_counts", _u32p),
        ("shape_path_is_closed", _u32p),
        ("shape_lengths", _f32p), ("shape_cdf", _f32p), ("shape_cdf_len", C.c_uint64),
        ("shape_pmf", _f32p), ("shape_ids", _u32p), ("group_ids", _u32p),
        ("path_cdf", _f32p), ("path_cdf_len", C.c_uint64), ("path_pmf", _f32p),
        ("path_point_ids", _u32p), ("path_point_ids_len", C.c_uint64),
        ("path_cdf_offsets", _u32p), ("path_cdf_counts", _u32p), ("path_point_offsets", _u32p),
    ]


class RenderParamsC(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("samples_x", C.c_uint32),
                ("samples_y", C.c_uint32), ("seed", C.c_uint32), ("jitter", C.c_uint32),
                ("use_prefiltering", C.c_uint32), ("filter_type", C.c_uint32),
                ("filter_radius", C.c_float), ("background", C.c_float * 4),
                ("background_image", _f32p), ("compute_translation", C.c_uint32)]


KERNEL_NAMES = ["zero", "weights", "forward", "resolve", "backward_interior", "boundary_sort", "boundary_eval"]

BATCH_FORWARD, BATCH_LOSS, BATCH_BACKWARD = 1, 2, 4


class BatchConfigC(C.Structure):
    _fields_ = [("sblock_log2", C.c_uint32), ("bnd_log2", C.c_uint32), ("wg_per_cu", C.c_uint32),
                ("unit_log2", C.c_uint32), ("ms", C.c_float)]


# Every symbol include/dvg.h declares, with its ctypes signature.
_vp = C.c_void_p
SIGNATURES = {
    "dvg_ctx_create": (_vp, [C.c_int]),
    "dvg_ctx_destroy": (None, [_vp]),
    "dvg_last_error": (C.c_char_p, [_vp]),
    "dvg_ctx_set_stream": (C.c_int32, [_vp, _vp]),
    "dvg_render": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions), _f32p]),
    "dvg_render_sdf": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions), _f32p]),
    "dvg_eval_positions": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions), _f32p, C.c_uint64,
                                       _f32p]),
    "dvg_device_render_sdf": (C.c_int32, [_vp, _vp, _vp]),
    "dvg_eval_positions_backward": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions),
                                                C.POINTER(BackwardOptions), _f32p, C.c_uint64, _f32p,
                                                C.POINTER(GradsC)]),
    "dvg_device_eval_positions_backward": (C.c_int32, [_vp, _vp, _vp, C.c_uint64, _vp, C.POINTER(GradsC)]),
    "dvg_render_backward_sdf": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions),
                                            C.POINTER(BackwardOptions), _f32p, C.POINTER(GradsC)]),
    "dvg_device_render_backward_sdf": (C.c_int32, [_vp, _vp, _vp, C.POINTER(GradsC)]),
    "dvg_device_eval_positions": (C.c_int32, [_vp, _vp, _vp, C.c_uint64, _vp]),
    "dvg_render_backward": (C.c_int32, [_vp, C.POINTER(SceneDescC), C.POINTER(RenderOptions),
                                        C.POINTER(BackwardOptions), _f32p, C.POINTER(GradsC)]),
    "dvg_render_packed": (C.c_int32, [_vp, C.POINTER(ScenePackC), C.POINTER(RenderParamsC), _f32p]),
    "dvg_render_backward_packed": (C.c_int32, [_vp, C.POINTER(ScenePackC), C.POINTER(BackwardPackC),
                                               C.POINTER(RenderParamsC), _f32p, C.POINTER(GradsC)]),
    "dvg_prepare": (C.c_int32, [C.POINTER(SceneDescC), C.POINTER(RenderOptions), C.POINTER(_vp),
                                C.POINTER(C.c_char_p)]),
    "dvg_prepared_free": (None, [_vp]),
    "dvg_prepared_pack": (None, [_vp, C.POINTER(ScenePackC), C.POINTER(BackwardPackC),
                                 C.POINTER(RenderParamsC)]),
    "dvg_device_scene_create": (C.c_int32, [_vp, C.POINTER(ScenePackC), C.POINTER(BackwardPackC),
                                            C.POINTER(RenderParamsC), C.POINTER(_vp)]),
    "dvg_device_scene_destroy": (None, [_vp]),
    "dvg_device_scene_set_shapes": (C.c_int32, [_vp, C.c_uint32, C.c_uint32, C.c_uint32]),
    "dvg_device_scene_update": (C.c_int32, [_vp, _vp, _vp, _vp]),
    "dvg_device_scene_check": (C.c_int32, [_vp, _vp]),
    "dvg_device_render": (C.c_int32, [_vp, _vp, _vp]),
    "dvg_device_render_noalloc": (C.c_int32, [_vp, _vp, _vp]),
    "dvg_device_render_backward": (C.c_int32, [_vp, _vp, _vp, C.POINTER(GradsC)]),
    "dvg_sync": (C.c_int32, [_vp]),
    "dvg_ctx_set_timing": (C.c_int32, [_vp, C.c_int]),
    "dvg_ctx_set_exact_boundary": (C.c_int32, [_vp, C.c_int]),
    "dvg_kernel_times": (C.c_int32, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int, C.c_int]),
    "dvg_batch_create": (C.c_int32, [_vp, C.c_uint32, C.POINTER(ScenePackC), C.POINTER(BackwardPackC),
                                     C.POINTER(RenderParamsC), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "dvg_batch_destroy": (None, [_vp]),
    "dvg_batch_buffers": (C.c_int32, [_vp, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(GradsC)]),
    "dvg_batch_read": (C.c_int32, [_vp, _vp, C.c_uint32, _vp, C.POINTER(GradsC)]),
    "dvg_batch_step": (C.c_int32, [_vp, _vp, C.c_uint32]),
    "dvg_batch_set_graph": (C.c_int32, [_vp, C.c_int]),
    "dvg_batch_get_config": (C.c_int32, [_vp, C.POINTER(BatchConfigC)]),
    "dvg_batch_set_config": (C.c_int32, [_vp, _vp, C.POINTER(BatchConfigC)]),
    "dvg_batch_autotune": (C.c_int32, [_vp, _vp, C.c_uint32, C.c_int, C.POINTER(BatchConfigC)]),
    "dvg_version": (C.c_char_p, []),
    "dvg_selftest_winding": (C.c_int32, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "dvg_selftest_inside": (C.c_int32, [_vp, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
}

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libdvg.so")

_lib = None


def load_library(path=None):
    """Load libdvg.so (built in-tree by __graft_entry__.build()).

    Raises OSError when the library is missing: there is no fallback path.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError(f"libdvg.so not built: {p} (run __graft_entry__.build())")
    # One HIP runtime per process.  torch ships its own libamdhip64 with the
    # same soname as the system one libdvg.so links; loading torch first makes
    # the dynamic loader bind libdvg.so to torch's copy, so device pointers and
    # streams pass between them.  Loaded the other way round, torch would map
    # a second runtime and find no device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib
