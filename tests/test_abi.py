"""CPU checks of the C ABI: the library loads, exports every symbol the
header declares, and the host packer reports the reference's errors."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from pydvg import abi, scenes
from pydvg.renderer import PreparedScene, RenderError
from pydvg.scene import Circle, Path, RenderOptions, Scene, SceneDesc, Shape, ShapeGroup, Solid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "dvg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dvg_[a-z_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _declared() == sorted(abi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = abi.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.dvg_version().startswith(b"dvg-mi355x")


def test_struct_sizes_match_header_layout():
    # pointer/u64 fields are 8 bytes; the structs hold no hidden padding
    # beyond natural alignment, so ctypes and the C compiler agree.
    assert C.sizeof(abi.GradsC) == 14 * 8
    assert C.sizeof(abi.RenderOptions) == 7 * 4
    assert C.sizeof(abi.BackwardOptions) == 4


def _prep_error(scene, opts=RenderOptions(use_prefiltering=False)):
    with pytest.raises(RenderError) as e:
        PreparedScene(scene, opts)
    return e.value


def test_invalid_shape_index_message():
    sc = Scene(16, 16)
    sc.shapes.append(Shape(Circle((8, 8), 3)))
    sc.groups.append(ShapeGroup([1], fill=Solid((1, 0, 0, 1))))
    e = _prep_error(sc)
    assert e.kind == "InvalidScene" and e.message == "group references invalid shape index"


def test_background_image_size_message():
    sc = Scene(8, 8, background_image=np.zeros(8 * 8 * 4 - 4, np.float32))
    sc.shapes.append(Shape(Circle((4, 4), 2)))
    sc.groups.append(ShapeGroup([0], fill=Solid((1, 0, 0, 1))))
    assert _prep_error(sc).message == "background image size mismatch"


def test_packer_layout_of_a_closed_cubic_path():
    sc, opts = scenes.small_cubic_scene(32, 32, 3, seed=4, spp=2)
    p = PreparedScene(sc, opts)
    pk, bp = p.pack, p.bpack
    assert pk.num_groups == 3 and pk.num_prepared_shapes == 3
    shape_data = np.ctypeslib.as_array(pk.shape_data, shape=(3 * 15,)).reshape(3, 15)
    curve = np.ctypeslib.as_array(pk.curve_data, shape=(pk.curve_data_len,)).reshape(-1, 13)
    # shape_data layout (prepare.rs:198-205): kind 2 = path, curve offset/count
    assert np.all(shape_data[:, 0] == 2)
    assert curve.shape[0] == int(shape_data[:, 13].sum())
    # closed path: the last segment ends on point 0 (path_utils.rs:7-16)
    g0 = sc.shapes[0].geometry
    n0 = int(shape_data[0, 13])
    last = curve[n0 - 1]
    assert last[0] == 2 and np.allclose(last[7:9], g0.points[0])
    # boundary CDF ends at 1 and has one entry per group-shape instance
    cdf = np.ctypeslib.as_array(bp.shape_cdf, shape=(bp.shape_cdf_len,))
    assert bp.shape_cdf_len == 3 and abs(cdf[-1] - 1.0) < 1e-6 and np.all(np.diff(cdf) >= 0)


def test_gradient_table_must_follow_group_order():
    from pydvg.scene import GradientStop, LinearGradient
    sc = Scene(16, 16)
    for i in range(2):
        sc.shapes.append(Shape(Circle((8, 8), 3)))
        sc.groups.append(ShapeGroup([i], fill=LinearGradient((0, 0), (16, 16), [GradientStop(0, (1, 0, 0, 1)),
                                                                                 GradientStop(1, (0, 0, 1, 1))])))
    d = SceneDesc(sc)
    # swap the two gradient indices: the packer must refuse the table
    d.fill_gradient[:] = d.fill_gradient[::-1].copy()
    lib = abi.load_library()
    h, err = C.c_void_p(), C.c_char_p()
    st = lib.dvg_prepare(d.ref(), C.byref(RenderOptions().to_c()), C.byref(h), C.byref(err))
    assert st == abi.DVG_INVALID_SCENE and err.value == b"gradient table not in group order"
