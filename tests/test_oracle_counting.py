"""CPU checks of the op-counting oracle build (liboracle_count.so: the oracle's
sources with `float` / `double` replaced by the counting scalars of
oracle/counted.h), which gives the algorithmic flop counts of the roofline.

  * It computes what the checking build computes: images bit for bit, the
    backward's gradients bit for bit.
  * Its tallies add over slices of samples, and keep the reference's tile
    culling: groups outside a sample's tile list add nothing.
"""
import ctypes as C

import numpy as np

import oracle_py
from pydvg import scenes
from pydvg.scene import GRAD_FIELDS, RenderOptions, SceneDesc


def _scene():
    return scenes.mixed_scene(48, 40, 16, 3), RenderOptions(use_prefiltering=False, samples_x=2, samples_y=2)


def test_counting_build_computes_what_the_checking_build_computes(oracle):
    sc, opts = _scene()
    d = SceneDesc(sc)
    L = oracle_py.count_lib()
    img = np.zeros(sc.width * sc.height * 4, np.float32)
    assert L.oracle_render(d.ref(), C.byref(opts.to_c()), img.ctypes.data_as(C.POINTER(C.c_float)), None) == 0
    assert np.array_equal(img, np.asarray(oracle.render(d, opts)).reshape(-1))


def test_counts_add_over_slices(oracle):
    sc, opts = _scene()
    d = SceneDesc(sc)
    dr = np.random.default_rng(0).normal(size=sc.width * sc.height * 4).astype(np.float32)
    a = oracle_py.count_slice(d, opts, dr, 100, 400)
    b = oracle_py.count_slice(d, opts, dr, 400, 700)
    ab = oracle_py.count_slice(d, opts, dr, 100, 700)
    for k in ("fwd", "interior", "boundary"):
        assert ab[k][0] == a[k][0] + b[k][0] and ab[k][1] == a[k][1] + b[k][1], k
        assert ab[k][0] > 0


def test_groups_outside_the_tile_list_are_not_counted(oracle):
    """One scene with a circle in the top-left tile, one with six: samples of
    a bottom-right tile (no group in its list) count the same in both, samples
    of the top-left tile count more with six circles.  The oracle evaluates
    every group everywhere; the tally keeps the reference's tile culling."""
    from pydvg.scene import Circle, Scene, Shape, ShapeGroup, Solid
    opts = RenderOptions(use_prefiltering=False, samples_x=1, samples_y=1, jitter=False)

    def scene(n):
        shapes = [Shape(Circle((6.0 + k * 0.5, 6.0), 3.0)) for k in range(n)]
        groups = [ShapeGroup([k], fill=Solid((1, 0, 0, 0.5))) for k in range(n)]
        return Scene(64, 64, shapes=shapes, groups=groups, background=(1, 1, 1, 1))

    dr = np.ones(64 * 64 * 4, np.float32)
    far, near = 63 * 64 + 40, 3 * 64 + 3
    one, six = SceneDesc(scene(1)), SceneDesc(scene(6))
    assert oracle_py.count_slice(one, opts, dr, far, far + 20)["fwd"] == \
        oracle_py.count_slice(six, opts, dr, far, far + 20)["fwd"]
    assert oracle_py.count_slice(one, opts, dr, far, far + 20)["interior"] == \
        oracle_py.count_slice(six, opts, dr, far, far + 20)["interior"]
    assert oracle_py.count_slice(six, opts, dr, near, near + 8)["fwd"][0] > \
        oracle_py.count_slice(one, opts, dr, near, near + 8)["fwd"][0]
