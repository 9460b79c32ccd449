"""GPU: the packed seam itself -- dvg_render_packed / dvg_render_backward_packed,
the calls a replacement of the reference's src/gpu binds -- on packs filled
the way prepare.rs:114-320 fills PreparedScene (flattened segment_data, the
shape_data segment fields, group and path BVHs), against the host-scene
entry points dvg_render / dvg_render_backward on the same scenes.

Three packs per scene: the packer's; the packer's with segment_data and the
BVH buffers replaced by oracle/oracle_pack.py's restatement (the buffers a
Rust caller computes itself); and one with no BVH at all (all-zero meta, no
segments), which the reference's kernels also accept.  The forward must
equal dvg_render's up to the order of the splat's float atomics (1e-6); the
backward's raw accumulations, finalized as finalize_background_gradients
does, must equal dvg_render_backward up to the order of float atomics (1e-5
relative per field).
"""
import ctypes as C

import numpy as np
import pytest

import oracle_pack as op
from parity_util import assert_grads
from pydvg import abi, scenes
from pydvg.renderer import PreparedScene
from pydvg.scene import (GRAD_FIELDS, FlatGrads, Path, RenderOptions, Scene, ShapeGroup, Solid,
                         finalize_background_gradients)

pytestmark = pytest.mark.gpu


def _scaled_groups():
    sc = scenes.mixed_scene(96, 96, 48, 3)
    base = sc.groups
    sc.groups = []
    for k, cut in enumerate(range(0, 48, 12)):
        g = ShapeGroup(list(range(cut, cut + 12)), fill=base[cut].fill, stroke=base[cut].stroke)
        s = (0.6, 1.0, 1.5, 0.8)[k]
        g.set_shape_to_canvas(np.array([[s, 0.1, 4.0 * k], [-0.1, s, 3.0], [0, 0, 1]], np.float32))
        sc.groups.append(g)
    return sc


CASES = {
    "mixed": lambda: (scenes.mixed_scene(80, 72, 40, 7), RenderOptions(use_prefiltering=False, samples_x=2,
                                                                       samples_y=2, seed=3)),
    "scaled_groups": lambda: (_scaled_groups(), RenderOptions(use_prefiltering=False, seed=1)),
    "prefiltered": lambda: (scenes.mixed_scene(64, 64, 24, 5), RenderOptions(use_prefiltering=True)),
    "cubic": lambda: scenes.small_cubic_scene(64, 64, 12, seed=2, spp=2),
}


def _as(arr, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype)) if arr.size else C.cast(None, C.POINTER(ctype))


def _restated(p, sc, opts):
    """The pack with segment_data and both BVHs rebuilt in Python; returns
    (pack, keep-alive arrays)."""
    pk = p.pack
    n = pk.num_prepared_shapes
    sd = np.ctypeslib.as_array(pk.shape_data, shape=(15 * n,)).reshape(-1, 15)
    sb = np.ctypeslib.as_array(pk.shape_bounds, shape=(4 * n,)).reshape(-1, 4)
    cd = (np.ctypeslib.as_array(pk.curve_data, shape=(pk.curve_data_len,)).reshape(-1, 13)
          if pk.curve_data_len else np.zeros((0, 13), np.float32))
    gout = {"bounds": [], "nodes": [], "indices": []}
    pout = {"bounds": [], "nodes": [], "indices": []}
    gmeta, pmeta, segs = [], [], []
    prep = 0
    for g in sc.groups:
        boxes, ids = [], []
        for si in g.shape_indices:
            sh = sc.shapes[si]
            if isinstance(sh.geometry, Path):
                segs.extend(op.segments_to_f32(op.flatten_with_thickness(
                    sh.geometry, op.shape_tolerance(opts, g, sh), sh.stroke_width)).tolist())
                co, cc = int(sd[prep, 12]), int(sd[prep, 13])
                pmeta.extend(op.pack_bvh([op.curve_box(cd[co + k]) for k in range(cc)],
                                         list(range(co, co + cc)), pout))
            else:
                pmeta.extend([0, 0, 0, 0])
            if sb[prep, 0] <= sb[prep, 2] and sb[prep, 1] <= sb[prep, 3]:
                boxes.append(tuple(sb[prep]))
                ids.append(prep)
            prep += 1
        gmeta.extend(op.pack_bvh(boxes, ids, gout))
    keep = {
        "seg": np.asarray(segs, np.float32), "gb": np.asarray(gout["bounds"], np.float32),
        "gn": np.asarray(gout["nodes"], np.uint32), "gi": np.asarray(gout["indices"], np.uint32),
        "gm": np.asarray(gmeta, np.uint32), "pb": np.asarray(pout["bounds"], np.float32),
        "pn": np.asarray(pout["nodes"], np.uint32), "pi": np.asarray(pout["indices"], np.uint32),
        "pm": np.asarray(pmeta, np.uint32),
    }
    q = abi.ScenePackC()
    C.pointer(q)[0] = pk
    q.segment_data, q.segment_data_len = _as(keep["seg"], C.c_float), keep["seg"].size
    q.group_bvh_bounds, q.group_bvh_bounds_len = _as(keep["gb"], C.c_float), keep["gb"].size
    q.group_bvh_nodes, q.group_bvh_nodes_len = _as(keep["gn"], C.c_uint32), keep["gn"].size
    q.group_bvh_indices, q.group_bvh_indices_len = _as(keep["gi"], C.c_uint32), keep["gi"].size
    q.group_bvh_meta = _as(keep["gm"], C.c_uint32)
    q.path_bvh_bounds, q.path_bvh_bounds_len = _as(keep["pb"], C.c_float), keep["pb"].size
    q.path_bvh_nodes, q.path_bvh_nodes_len = _as(keep["pn"], C.c_uint32), keep["pn"].size
    q.path_bvh_indices, q.path_bvh_indices_len = _as(keep["pi"], C.c_uint32), keep["pi"].size
    q.path_bvh_meta = _as(keep["pm"], C.c_uint32)
    return q, keep


def _no_bvh(p):
    pk = p.pack
    G, n = pk.num_groups, pk.num_prepared_shapes
    keep = {"gm": np.zeros(4 * G, np.uint32), "pm": np.zeros(4 * n, np.uint32),
            "sd": np.ctypeslib.as_array(pk.shape_data, shape=(15 * n,)).copy()}
    keep["sd"].reshape(-1, 15)[:, 1:3] = 0
    q = abi.ScenePackC()
    C.pointer(q)[0] = pk
    q.shape_data = _as(keep["sd"], C.c_float)
    q.segment_data, q.segment_data_len = C.cast(None, C.POINTER(C.c_float)), 0
    for f in ("group_bvh_bounds", "path_bvh_bounds"):
        setattr(q, f, C.cast(None, C.POINTER(C.c_float)))
        setattr(q, f + "_len", 0)
    for f in ("group_bvh_nodes", "group_bvh_indices", "path_bvh_nodes", "path_bvh_indices"):
        setattr(q, f, C.cast(None, C.POINTER(C.c_uint32)))
        setattr(q, f + "_len", 0)
    q.group_bvh_meta, q.path_bvh_meta = _as(keep["gm"], C.c_uint32), _as(keep["pm"], C.c_uint32)
    return q, keep


@pytest.mark.parametrize("case", sorted(CASES))
def test_packed_seam_equals_host_entry_points(renderer, case):
    sc, opts = CASES[case]()
    print(f"case packed seam {case}: {sc.width}x{sc.height}, {len(sc.shapes)} shapes", flush=True)
    lib, ctx = renderer.lib, renderer.ctx
    img_ref = renderer.render(sc, opts)
    rng = np.random.default_rng(1)
    d_render = rng.normal(size=sc.width * sc.height * 4).astype(np.float32)
    g_ref = renderer.render_backward(sc, opts, d_render)
    p = PreparedScene(sc, opts)
    assert p.pack.segment_data_len > 0 and p.pack.group_bvh_nodes_len > 0
    packs = {"packer": (p.pack, None), "restated": _restated(p, sc, opts), "no_bvh": _no_bvh(p)}
    for name, (pk, _keep) in packs.items():
        img = np.zeros(sc.width * sc.height * 4, np.float32)
        st = lib.dvg_render_packed(ctx, C.byref(pk), C.byref(p.params), _as(img, C.c_float))
        assert st == abi.DVG_OK, (name, lib.dvg_last_error(ctx))
        # the splat's float atomics may add in another order from run to run
        err = float(np.abs(img - np.asarray(img_ref).reshape(-1)).max())
        assert err <= 1e-6, (name, err)
        g = FlatGrads(p.desc)
        st = lib.dvg_render_backward_packed(ctx, C.byref(pk), C.byref(p.bpack), C.byref(p.params),
                                            _as(d_render, C.c_float), C.byref(g.c))
        assert st == abi.DVG_OK, (name, lib.dvg_last_error(ctx))
        got = {k: np.array(g[k]) for k in GRAD_FIELDS}
        finalize_background_gradients(sc, got)
        total, _ = assert_grads(got, {k: g_ref[k] for k in GRAD_FIELDS}, 1e-5, label=f"{case} {name}")
        assert total <= 1e-5
