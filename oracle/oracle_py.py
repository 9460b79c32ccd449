"""ctypes binding of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module.  It is the checker; nothing in the product path loads it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        from pydvg import abi
        L = C.CDLL(LIB)
        f32p, u32p = C.POINTER(C.c_float), C.POINTER(C.c_uint32)
        L.oracle_render.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions), f32p, f32p]
        L.oracle_render.restype = C.c_int
        L.oracle_render_backward.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions),
                                             C.POINTER(abi.BackwardOptions), f32p, C.POINTER(abi.GradsC)]
        L.oracle_render_backward.restype = C.c_int
        L.oracle_render_rows.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions), C.c_uint32,
                                         C.c_uint32, f32p]
        L.oracle_render_rows.restype = C.c_int
        L.oracle_render_sdf.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions), f32p]
        L.oracle_render_sdf.restype = C.c_int
        L.oracle_eval_positions_backward.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.BackwardOptions),
                                                     f32p, C.c_uint64, f32p, C.POINTER(abi.GradsC)]
        L.oracle_eval_positions_backward.restype = C.c_int
        L.oracle_render_backward_sdf.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions),
                                                 C.POINTER(abi.BackwardOptions), f32p, C.POINTER(abi.GradsC)]
        L.oracle_render_backward_sdf.restype = C.c_int
        L.oracle_eval_positions.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions), f32p,
                                            C.c_uint64, f32p]
        L.oracle_eval_positions.restype = C.c_int
        L.oracle_render_backward_rows.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions),
                                                  C.POINTER(abi.BackwardOptions), f32p, C.POINTER(abi.GradsC),
                                                  C.c_uint32, C.c_uint32]
        L.oracle_render_backward_rows.restype = C.c_int
        L.oracle_pcg32.argtypes = [C.c_uint64, C.c_uint64, C.c_int, u32p, f32p]
        L.oracle_solve_cubic_f64.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)]
        L.oracle_solve_cubic_f64.restype = C.c_int
        L.oracle_solve_cubic_f32.argtypes = [C.c_float] * 4 + [f32p]
        L.oracle_solve_cubic_f32.restype = C.c_int
        L.oracle_closest_f64.argtypes = [C.c_uint32, f32p, C.c_float, C.c_float, C.c_int, f32p]
        L.oracle_closest_f32.argtypes = [C.c_uint32, f32p, C.c_float, C.c_float, C.c_int, f32p]
        L.oracle_winding_f64.argtypes = [C.c_uint32, f32p, C.c_float, C.c_float]
        L.oracle_winding_f64.restype = C.c_int
        L.oracle_winding_f32.argtypes = [C.c_uint32, f32p, C.c_float, C.c_float]
        L.oracle_winding_f32.restype = C.c_float
        L.oracle_within_distance.argtypes = [C.POINTER(abi.SceneDescC), C.c_uint32, C.c_float, C.c_float]
        L.oracle_is_inside.argtypes = [C.POINTER(abi.SceneDescC), C.c_uint32, C.c_float, C.c_float]
        L.oracle_filter_weight.argtypes = [C.c_uint32, C.c_float, C.c_float, C.c_float]
        L.oracle_filter_weight.restype = C.c_float
        L.oracle_prefiltered_objective.argtypes = [C.POINTER(abi.SceneDescC), C.POINTER(abi.RenderOptions), f32p]
        L.oracle_prefiltered_objective.restype = C.c_double
        _lib = L
    return _lib


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def render(desc, opts, want_samples=False):
    """Forward image (W*H*4) with the reference GPU-forward semantics."""
    sc = desc.scene
    out = np.zeros(sc.width * sc.height * 4, np.float32)
    spp = opts.samples_x * opts.samples_y
    samples = np.zeros(sc.width * sc.height * spp * 4, np.float32) if want_samples else None
    st = lib().oracle_render(desc.ref(), C.byref(opts.to_c()), _fp(out),
                             _fp(samples) if want_samples else None)
    if st != 0:
        raise RuntimeError(f"oracle_render status {st}")
    return (out, samples) if want_samples else out


def render_sdf(desc, opts):
    """Renderer::render_sdf: W * H mean signed distances of each pixel's
    samples (negative inside a fill)."""
    sc = desc.scene
    out = np.zeros(sc.width * sc.height, np.float32)
    st = lib().oracle_render_sdf(desc.ref(), C.byref(opts.to_c()), _fp(out))
    if st != 0:
        raise RuntimeError(f"oracle_render_sdf status {st}")
    return out


def eval_positions(desc, opts, positions):
    """Renderer::eval_positions: the signed distance at each canvas point
    (positions: n x 2)."""
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    out = np.zeros(pos.size // 2, np.float32)
    st = lib().oracle_eval_positions(desc.ref(), C.byref(opts.to_c()), _fp(pos), pos.size // 2, _fp(out))
    if st != 0:
        raise RuntimeError(f"oracle_eval_positions status {st}")
    return out


def eval_positions_backward(desc, positions, d_sdf, translation=False):
    """Renderer::eval_positions_backward (render_backward_positions, the
    reference CPU's): FlatGrads of the signed distances at the points."""
    from pydvg.scene import BackwardOptions, FlatGrads
    g = FlatGrads(desc, translation)
    pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
    d_sdf = np.ascontiguousarray(d_sdf, np.float32).reshape(-1)
    st = lib().oracle_eval_positions_backward(desc.ref(), C.byref(BackwardOptions(translation).to_c()), _fp(pos),
                                              d_sdf.size, _fp(d_sdf), g.ref())
    if st != 0:
        raise RuntimeError(f"oracle_eval_positions_backward status {st}")
    return g


def render_backward_sdf(desc, opts, d_sdf, translation=False):
    """The d_sdf image term of render_backward (the reference CPU's; d_render
    absent) as FlatGrads."""
    from pydvg.scene import BackwardOptions, FlatGrads
    g = FlatGrads(desc, translation)
    d_sdf = np.ascontiguousarray(d_sdf, np.float32).reshape(-1)
    st = lib().oracle_render_backward_sdf(desc.ref(), C.byref(opts.to_c()),
                                          C.byref(BackwardOptions(translation).to_c()), _fp(d_sdf), g.ref())
    if st != 0:
        raise RuntimeError(f"oracle_render_backward_sdf status {st}")
    return g


def render_rows(desc, opts, y0, y1):
    """Pixel rows [y0, y1) of the forward image ((y1 - y0) * W * 4), equal to
    the same rows of render(): only the samples whose filter reaches them are
    evaluated."""
    sc = desc.scene
    y1 = min(y1, sc.height)
    out = np.zeros(max(y1 - y0, 0) * sc.width * 4, np.float32)
    st = lib().oracle_render_rows(desc.ref(), C.byref(opts.to_c()), y0, y1, _fp(out))
    if st != 0:
        raise RuntimeError(f"oracle_render_rows status {st}")
    return out


def render_backward_rows(desc, opts, d_render, y0, y1, translation=False):
    """render_backward for a d_render that is zero outside pixel rows [y0, y1):
    the samples that gather a zero adjoint are skipped (they add exact zeros)."""
    from pydvg.scene import BackwardOptions, FlatGrads
    g = FlatGrads(desc, translation)
    d_render = np.ascontiguousarray(d_render, np.float32)
    st = lib().oracle_render_backward_rows(desc.ref(), C.byref(opts.to_c()),
                                           C.byref(BackwardOptions(translation).to_c()), _fp(d_render), g.ref(),
                                           y0, y1)
    if st != 0:
        raise RuntimeError(f"oracle_render_backward_rows status {st}")
    return g


def prefiltered_objective(desc, opts, d_render):
    """sum_samples dot(CPU prefiltered colour, gather_d_color): the scalar whose
    gradient render_backward returns for a prefiltered scene."""
    d_render = np.ascontiguousarray(d_render, np.float32)
    return float(lib().oracle_prefiltered_objective(desc.ref(), C.byref(opts.to_c()), _fp(d_render)))


def render_backward(desc, opts, d_render, translation=False):
    """Reference CPU render_backward gradients as pydvg.scene.FlatGrads."""
    from pydvg.scene import BackwardOptions, FlatGrads
    g = FlatGrads(desc, translation)
    d_render = np.ascontiguousarray(d_render, np.float32)
    st = lib().oracle_render_backward(desc.ref(), C.byref(opts.to_c()),
                                      C.byref(BackwardOptions(translation).to_c()), _fp(d_render), g.ref())
    if st != 0:
        raise RuntimeError(f"oracle_render_backward status {st}")
    return g


def pcg32(idx, seed, n):
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    lib().oracle_pcg32(idx, seed, n, u.ctypes.data_as(C.POINTER(C.c_uint32)), _fp(f))
    return u, f


def solve_cubic_f64(a, b, c, d):
    r = (C.c_double * 3)()
    n = lib().oracle_solve_cubic_f64(a, b, c, d, r)
    return sorted(r[i] for i in range(n))


def solve_cubic_f32(a, b, c, d):
    r = np.zeros(3, np.float32)
    n = lib().oracle_solve_cubic_f32(a, b, c, d, _fp(r))
    return sorted(float(x) for x in r[:n])


def closest(kind, pts, p, approx=False, f32=False):
    pts = np.ascontiguousarray(np.pad(np.asarray(pts, np.float32).reshape(-1), (0, 8 - np.asarray(pts).size)),
                               np.float32)
    if f32:
        out = np.zeros(2, np.float32)
        lib().oracle_closest_f32(kind, _fp(pts), p[0], p[1], int(approx), _fp(out))
        return float(out[0]), float(out[1])
    out = np.zeros(4, np.float32)
    lib().oracle_closest_f64(kind, _fp(pts), p[0], p[1], int(approx), _fp(out))
    return (float(out[0]), float(out[1])), float(out[2]), float(out[3])


def winding(kind, pts, p, f32=False):
    pts = np.ascontiguousarray(np.pad(np.asarray(pts, np.float32).reshape(-1), (0, 8 - np.asarray(pts).size)), np.float32)
    if f32:
        return lib().oracle_winding_f32(kind, _fp(pts), p[0], p[1])
    return lib().oracle_winding_f64(kind, _fp(pts), p[0], p[1])


def within_distance(desc, group, p):
    return bool(lib().oracle_within_distance(desc.ref(), group, p[0], p[1]))


def is_inside(desc, group, p):
    return bool(lib().oracle_is_inside(desc.ref(), group, p[0], p[1]))


def filter_weight(ftype, radius, dx, dy):
    return float(lib().oracle_filter_weight(ftype, radius, dx, dy))


def time_slice(desc, opts, d_render, seconds=15.0):
    """Time the oracle's per-sample forward + backward + boundary work on a
    contiguous slice of sample indices of the full workload, growing the
    slice until it takes about `seconds`; returns pixel-samples/s."""
    import time as _t
    L = lib()
    L.oracle_time_slice.argtypes = [C.POINTER(__import__("pydvg").abi.SceneDescC),
                                    C.POINTER(__import__("pydvg").abi.RenderOptions), C.POINTER(C.c_float),
                                    C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]
    L.oracle_time_slice.restype = C.c_double
    sc = desc.scene
    total = sc.width * sc.height * opts.samples_x * opts.samples_y
    d_render = np.ascontiguousarray(d_render, np.float32)
    copts = opts.to_c()
    # start the slice at the image centre rows so it sees typical coverage
    n = 4096
    ws = C.c_double(0.0)
    while True:
        i0 = max(0, total // 2 - n // 2)
        t = L.oracle_time_slice(desc.ref(), C.byref(copts), _fp(d_render), i0, i0 + n, C.byref(ws))
        if t >= seconds * 0.5 or n >= total:
            break
        n = min(total, int(n * max(2.0, seconds / max(t, 1e-3))))
    per_sample = t / n
    return {"samples_per_s": 1.0 / per_sample, "seconds": t, "samples": n, "weight_image_seconds": ws.value,
            "sample": f"{n} consecutive pixel-samples (indices {i0}..{i0 + n}) of the full workload, "
                      f"forward + interior backward + boundary sampling, 1 thread; the weight image "
                      f"({ws.value:.2f} s for all samples) is excluded"}


COUNT_LIB = os.path.join(_HERE, "liboracle_count.so")
_count_lib = None


def count_lib():
    """liboracle_count.so: the oracle built with op-counting scalars
    (counted.h); built on first use."""
    global _count_lib
    if _count_lib is None:
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle_count.so"])
        from pydvg import abi as _abi
        L = C.CDLL(COUNT_LIB)
        L.oracle_count_slice.argtypes = [C.POINTER(_abi.SceneDescC), C.POINTER(_abi.RenderOptions),
                                         C.POINTER(C.c_float), C.c_uint64, C.c_uint64, C.POINTER(C.c_double)]
        L.oracle_render.argtypes = [C.POINTER(_abi.SceneDescC), C.POINTER(_abi.RenderOptions),
                                    C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.oracle_render.restype = C.c_int
        _count_lib = L
    return _count_lib


def count_slice(desc, opts, d_render, i0, i1):
    """Arithmetic the oracle evaluates for the pixel-samples [i0, i1): forward,
    interior adjoint and boundary sampling, each as (flops, sfu ops), with the
    reference's tile culling (oracle_count_slice)."""
    L = count_lib()
    out = (C.c_double * 6)()
    dr = np.ascontiguousarray(d_render, np.float32)
    L.oracle_count_slice(desc.ref(), C.byref(opts.to_c()), _fp(dr), i0, i1, out)
    return {"fwd": (out[0], out[1]), "interior": (out[2], out[3]), "boundary": (out[4], out[5]),
            "samples": i1 - i0}
