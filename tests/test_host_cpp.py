"""The C++ host API (diffvg-rs_amd/host/diffvg.hpp) against the Python host.

tests/host_check.cpp builds one scene with the C++ types (every shape kind,
both gradient kinds, a rotated group, a background image, per-point
thickness, the tent filter); _py_scene() below builds the same scene with
pydvg.  CPU tests: the C++ FlatScene equals pydvg's SceneDesc array by array,
Mat3::inverse equals mat3_inverse bit for bit, and the gradient tree the C++
apply_gpu_grads / accumulate_from make from given flat arrays equals pydvg's.
GPU tests: the C++ Renderer's image, SDF, position queries and gradient trees
against the oracle, with the tolerances of the other GPU parity tests.
"""
import ctypes as C
import os

import numpy as np
import pytest

from parity_util import image_report, tree_leaves
from pydvg import abi
from pydvg.scene import (Circle, Ellipse, Filter, FlatGrads, GradientStop, LinearGradient, Path, RadialGradient,
                         Rect, RenderOptions, Scene, SceneDesc, Shape, ShapeGroup, Solid, mat3_inverse)

HERE = os.path.dirname(os.path.abspath(__file__))
CHECK_LIB = os.path.join(HERE, "libhost_check.so")
HOST_LIB = os.path.join(os.path.dirname(HERE), "diffvg-rs_amd", "libdiffvg_host.so")
IMG_TOL = 1e-5
FLIP_FRACTION = 1e-3
GRAD_REL = 1e-3
CAP = 1 << 16


def _py_scene():
    W, H = 24, 20
    sc = Scene(W, H, background=(0.9, 0.95, 1.0, 1.0))
    sc.background_image = ((np.arange(W * H * 4) % 13).astype(np.float32) / np.float32(13.0)).reshape(H, W, 4)
    p1 = np.array([[2, 14], [6, 8], [10, 18], [14, 12], [16, 8], [20, 16], [22, 10]], np.float32)
    p4 = np.array([[12, 11], [18, 9], [21, 14], [19, 18], [16, 19], [13, 17]], np.float32)
    sc.shapes = [
        Shape(Circle((6.0, 6.0), 4.0), stroke_width=1.0),
        Shape(Path([2, 2], p1, thickness=[1.0, 1.25, 1.5, 1.75, 2.0, 1.5, 1.0]), stroke_width=1.5),
        Shape(Ellipse((15.0, 6.0), (5.0, 3.0))),
        Shape(Rect((3.0, 10.0), (9.0, 17.0)), stroke_width=0.75, stroke_join=abi.JOIN_MITER, stroke_cap=abi.CAP_BUTT),
        Shape(Path([1, 0, 2], p4, is_closed=True), stroke_width=0.5, stroke_join=abi.JOIN_BEVEL,
              stroke_miter_limit=2.0),
    ]
    lin = LinearGradient((0.0, 0.0), (8.0, 8.0), [GradientStop(0.0, (1, 0, 0, 1)), GradientStop(1.0, (0, 0, 1, 1))])
    rad = RadialGradient((15.0, 6.0), (4.0, 3.0), [GradientStop(0.0, (1, 1, 0, 1)), GradientStop(0.5, (0, 1, 0, 0.8)),
                                                   GradientStop(1.0, (0, 1, 1, 1))])
    lin2 = LinearGradient((3.0, 10.0), (21.0, 18.0), [GradientStop(0.2, (0.5, 0, 0.5, 1)),
                                                      GradientStop(0.9, (0, 0.5, 0, 0.6))])
    g2 = ShapeGroup([2], fill=rad, opacity=0.7)
    g2.set_shape_to_canvas(np.array([[0.8, -0.6, 6.6], [0.6, 0.8, -7.8], [0, 0, 1]], np.float32))
    sc.groups = [
        ShapeGroup([0], fill=Solid((1, 0, 0, 1)), stroke=lin),
        ShapeGroup([1], stroke=Solid((0, 0, 1, 0.5))),
        g2,
        ShapeGroup([3, 4], fill=lin2, stroke=Solid((0.2, 0.3, 0.4, 1)), fill_rule=abi.FILL_EVENODD),
    ]
    sc.filter = Filter(abi.FILTER_TENT, 1.0)
    return sc


def _py_options():
    return RenderOptions(samples_x=2, samples_y=2, seed=7, use_prefiltering=False)


_lib = None


def _check_lib():
    global _lib
    if _lib is None:
        if not os.path.exists(CHECK_LIB):
            raise RuntimeError(f"{CHECK_LIB} is not built (run __graft_entry__.build())")
        lib = C.CDLL(CHECK_LIB)
        dp, fp, u64 = C.POINTER(C.c_double), C.POINTER(C.c_float), C.c_uint64
        fpp = C.POINTER(fp)
        lib.hc_last_error.restype = C.c_char_p
        lib.hc_flat_field.argtypes = [C.c_char_p, dp, C.c_int]
        lib.hc_grad_sizes.argtypes = [C.POINTER(u64)]
        lib.hc_tree_leaves.argtypes = [fpp, dp, C.c_int]
        lib.hc_accumulate_leaves.argtypes = [fpp, fpp, dp, C.c_int]
        lib.hc_accumulate_into_zeros.argtypes = [fpp, dp, C.c_int]
        lib.hc_inverse.argtypes = [fp, fp]
        lib.hc_render.argtypes = [fp]
        lib.hc_render_sdf.argtypes = [fp]
        lib.hc_render_backward.argtypes = [fp, u64, fp, u64, dp, C.c_int]
        lib.hc_eval_positions.argtypes = [fp, u64, fp]
        lib.hc_eval_positions_backward.argtypes = [fp, u64, fp, u64, dp, C.c_int]
        _lib = lib
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _field(name):
    out = np.zeros(CAP, np.float64)
    n = _check_lib().hc_flat_field(name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), CAP)
    assert 0 <= n <= CAP, f"{name}: {n}"
    return out[:n]


def _leaves_call(fn, *args):
    out = np.zeros(CAP, np.float64)
    n = fn(*args, out.ctypes.data_as(C.POINTER(C.c_double)), CAP)
    assert 0 <= n <= CAP, _check_lib().hc_last_error().decode()
    return out[:n]


def _pointer_array(arrays):
    fp = C.POINTER(C.c_float)
    return (fp * len(arrays))(*[_fp(a) for a in arrays])


def _filled(desc, seed):
    g = FlatGrads(desc, translation=True)
    rng = np.random.default_rng(seed)
    for a in g.arrays.values():
        a[:] = rng.uniform(-1, 1, size=a.size).astype(np.float32)
    return g


# ---- CPU ------------------------------------------------------------------------

def test_host_library_links_the_c_abi():
    assert os.path.exists(HOST_LIB)
    C.CDLL(HOST_LIB)
    _check_lib()


def test_flat_scene_equals_pydvg_scene_desc():
    d = SceneDesc(_py_scene())
    py = {
        "shape_kind": d.shape_kind, "shape_params": d.shape_params, "shape_transform": d.shape_transform,
        "stroke_width": d.stroke_width, "stroke_join": d.stroke_join, "stroke_cap": d.stroke_cap,
        "stroke_miter_limit": d.stroke_miter, "path_point_offset": d.p_off, "path_point_count": d.p_cnt,
        "path_ctrl_offset": d.c_off, "path_ctrl_count": d.c_cnt, "path_thickness_offset": d.t_off,
        "path_thickness_count": d.t_cnt, "path_is_closed": d.closed, "path_use_distance_approx": d.approx,
        "path_points": d.path_points, "path_num_controls": d.path_num_controls, "path_thickness": d.path_thickness,
        "group_shape_offset": d.group_shape_offset, "group_shape_count": d.group_shape_count,
        "group_shape_indices": d.group_shape_indices, "fill_kind": d.fill_kind, "stroke_kind": d.stroke_kind,
        "fill_color": d.fill_color, "stroke_color": d.stroke_color, "fill_gradient": d.fill_gradient,
        "stroke_gradient": d.stroke_gradient, "fill_rule": d.fill_rule, "shape_to_canvas": d.shape_to_canvas,
        "canvas_to_shape": d.canvas_to_shape, "opacity": d.opacity, "gradient_kind": d.gradient_kind,
        "gradient_params": d.gradient_params, "gradient_stop_offset": d.gradient_stop_offset,
        "gradient_stop_count": d.gradient_stop_count, "stop_offsets": d.stop_offsets,
        "stop_colors": d.stop_colors, "background_image": d.bg_image,
    }
    for name, want in py.items():
        got = _field(name)
        assert np.array_equal(got, np.asarray(want, np.float64).ravel()), name
    c = d.c
    header = [c.width, c.height, *c.background, c.filter_type, c.filter_radius, c.num_shapes, c.num_groups,
              c.num_gradients, c.num_stops, c.path_points_len, c.path_num_controls_len, c.path_thickness_len,
              c.group_shape_indices_len, c.background_image_len]
    assert np.array_equal(_field("header"), np.asarray(header, np.float64))
    assert _check_lib().hc_flat_field(b"no_such_field", None, 0) == -1


def test_mat3_inverse_matches_the_reference_order():
    rng = np.random.default_rng(5)
    mats = [np.array([[0.8, -0.6, 6.6], [0.6, 0.8, -7.8], [0, 0, 1]], np.float32),
            np.array([[2.0, 4.0, 1.0], [1.0, 2.0, 3.0], [0, 0, 1]], np.float32),        # singular
            np.array([[1e-3, 0.0, 0.0], [0.0, 1e-4, 0.0], [0, 0, 1]], np.float32)]      # det 1e-7 < 1e-6
    for _ in range(20):
        m = np.eye(3, dtype=np.float32)
        m[:2, :] = rng.uniform(-3, 3, size=(2, 3)).astype(np.float32)
        mats.append(m)
    lib = _check_lib()
    for m in mats:
        out = np.zeros(9, np.float32)
        ok = lib.hc_inverse(_fp(np.ascontiguousarray(m.ravel())), _fp(out))
        want = mat3_inverse(m)
        assert bool(ok) == (want is not None)
        if want is not None:
            assert np.array_equal(out.reshape(3, 3), want)


def test_apply_gpu_grads_matches_pydvg():
    desc = SceneDesc(_py_scene())
    sizes = (C.c_uint64 * 14)()
    assert _check_lib().hc_grad_sizes(sizes) == 14
    f = _filled(desc, 11)
    from pydvg.scene import GRAD_FIELDS
    assert [int(s) for s in sizes] == [f.arrays[k].size for k in GRAD_FIELDS]
    got = _leaves_call(_check_lib().hc_tree_leaves, _pointer_array([f.arrays[k] for k in GRAD_FIELDS]))
    assert np.array_equal(got, tree_leaves(f.scene_grad()))


def test_accumulate_from_matches_pydvg():
    from pydvg.grad import SceneGrad
    from pydvg.scene import GRAD_FIELDS
    sc = _py_scene()
    desc = SceneDesc(sc)
    a, b = _filled(desc, 12), _filled(desc, 13)
    pa, pb = _pointer_array([a.arrays[k] for k in GRAD_FIELDS]), _pointer_array([b.arrays[k] for k in GRAD_FIELDS])
    ta, tb = a.scene_grad(), b.scene_grad()
    ta.accumulate_from(tb)
    got = _leaves_call(_check_lib().hc_accumulate_leaves, pa, pb)
    np.testing.assert_allclose(got, tree_leaves(ta), rtol=0, atol=1e-6)
    z = SceneGrad.zeros_from_scene(sc, False, False)
    z.accumulate_from(tb)
    got = _leaves_call(_check_lib().hc_accumulate_into_zeros, pb)
    assert np.array_equal(got, tree_leaves(z))


# ---- GPU ------------------------------------------------------------------------

def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-12))


@pytest.mark.gpu
def test_cpp_renderer_forward_matches_oracle(oracle):
    sc, opts = _py_scene(), _py_options()
    lib = _check_lib()
    print("case C++ Renderer::render / render_sdf / eval_positions", flush=True)
    img = np.zeros(sc.width * sc.height * 4, np.float32)
    assert lib.hc_render(_fp(img)) == 0, lib.hc_last_error().decode()
    ref = oracle.render(SceneDesc(sc), opts)
    frac, mx = image_report(img, ref, IMG_TOL)
    print(f"image: pixels beyond {IMG_TOL}: {frac:.2e}, max |diff| {mx:.3g}")
    assert frac <= FLIP_FRACTION
    sdf = np.zeros(sc.width * sc.height, np.float32)
    assert lib.hc_render_sdf(_fp(sdf)) == 0, lib.hc_last_error().decode()
    sref = oracle.render_sdf(SceneDesc(sc), opts)
    assert np.abs(sdf - sref.reshape(-1)).max() <= 1e-4
    pos = np.random.default_rng(4).uniform(0, 24, size=(64, 2)).astype(np.float32)
    out = np.zeros(64, np.float32)
    assert lib.hc_eval_positions(_fp(pos), 64, _fp(out)) == 0, lib.hc_last_error().decode()
    assert np.abs(out - oracle.eval_positions(SceneDesc(sc), opts, pos)).max() <= 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("with_sdf", [False, True])
def test_cpp_renderer_backward_matches_oracle(oracle, with_sdf):
    """render_backward's tree (with translation and the background image;
    with_sdf adds the d_sdf term through accumulate_from) against the tree
    pydvg makes from the oracle's flat gradients."""
    sc, opts = _py_scene(), _py_options()
    desc = SceneDesc(sc)
    lib = _check_lib()
    print(f"case C++ Renderer::render_backward with_sdf={with_sdf}", flush=True)
    ref = oracle.render(desc, opts)
    rng = np.random.default_rng(21)
    d_render = rng.normal(size=ref.size).astype(np.float32)
    d_sdf = rng.normal(size=sc.width * sc.height).astype(np.float32) if with_sdf else None
    got = _leaves_call(lib.hc_render_backward, _fp(d_render), d_render.size,
                       _fp(d_sdf) if with_sdf else None, d_sdf.size if with_sdf else 0)
    t = oracle.render_backward(desc, opts, d_render, translation=True).scene_grad()
    if with_sdf:
        t.accumulate_from(oracle.render_backward_sdf(desc, opts, d_sdf, translation=True).scene_grad())
    want = tree_leaves(t)
    rel = _rel(got, want)
    print(f"tree leaves rel L2 {rel:.3g}")
    assert got.size == want.size and rel <= GRAD_REL


@pytest.mark.gpu
def test_cpp_renderer_positions_backward_and_errors(oracle):
    sc, opts = _py_scene(), _py_options()
    desc = SceneDesc(sc)
    lib = _check_lib()
    print("case C++ Renderer::eval_positions_backward and size errors", flush=True)
    rng = np.random.default_rng(8)
    pos = rng.uniform(0, 24, size=(50, 2)).astype(np.float32)
    d_sdf = rng.normal(size=50).astype(np.float32)
    got = _leaves_call(lib.hc_eval_positions_backward, _fp(pos), 50, _fp(d_sdf), 50)
    want = tree_leaves(oracle.eval_positions_backward(desc, pos, d_sdf).scene_grad())
    rel = _rel(got, want)
    print(f"positions tree rel L2 {rel:.3g}")
    assert got.size == want.size and rel <= GRAD_REL
    # RenderError::InvalidScene with the crate's message for a wrong adjoint size
    short = np.zeros(10, np.float32)
    out = np.zeros(8, np.float64)
    st = lib.hc_render_backward(_fp(short), short.size, None, 0, out.ctypes.data_as(C.POINTER(C.c_double)), 8)
    assert st == -1 and lib.hc_last_error().decode() == "InvalidScene: d_render_image size mismatch"
    st = lib.hc_eval_positions_backward(_fp(pos), 50, _fp(d_sdf), 49, out.ctypes.data_as(C.POINTER(C.c_double)), 8)
    assert st == -1 and lib.hc_last_error().decode() == "InvalidScene: d_sdf_image size mismatch for eval_positions"
