"""Renderer over libdvg.so, mirroring the crate's Renderer
(src/renderer/renderer.rs:27-848): `render` and `render_backward` take a
Scene and RenderOptions and return an RGBA image / flat gradients.

`PreparedScene` exposes the packer (dvg_prepare) and `DeviceScene` the
device-resident path that keeps inputs in HBM between calls (torch tensors for
device memory; the work runs on the context's HIP stream).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .scene import BackwardOptions, FlatGrads, GRAD_FIELDS, RenderOptions, Scene, SceneDesc


class RenderError(RuntimeError):
    """RenderError (src/renderer/types.rs:83-88): .kind is 'InvalidScene' or 'Launch'."""

    def __init__(self, kind, message):
        super().__init__(f"{kind}: {message}")
        self.kind = kind
        self.message = message


def _check(lib, ctx, status):
    if status == abi.DVG_OK:
        return
    msg = lib.dvg_last_error(ctx)
    msg = msg.decode() if msg else ""
    raise RenderError("InvalidScene" if status == abi.DVG_INVALID_SCENE else "Launch", msg)


class PreparedScene:
    """dvg_prepare output: the packed seam buffers for one Scene."""

    def __init__(self, scene: Scene, options: RenderOptions):
        self.lib = abi.load_library()
        self.desc = SceneDesc(scene)
        self.options = options
        h = C.c_void_p()
        err = C.c_char_p()
        st = self.lib.dvg_prepare(self.desc.ref(), C.byref(options.to_c()), C.byref(h), C.byref(err))
        if st != abi.DVG_OK:
            raise RenderError("InvalidScene", (err.value or b"").decode())
        self.handle = h
        self.pack = abi.ScenePackC()
        self.bpack = abi.BackwardPackC()
        self.params = abi.RenderParamsC()
        self.lib.dvg_prepared_pack(h, C.byref(self.pack), C.byref(self.bpack), C.byref(self.params))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.lib.dvg_prepared_free(h)
            self.handle = None


class Renderer:
    """Renderer::new / render / render_backward on one GPU."""

    def __init__(self, device: int = 0):
        self.lib = abi.load_library()
        self.ctx = self.lib.dvg_ctx_create(device)
        if not self.ctx:
            raise RenderError("Launch", f"no usable HIP device {device}")

    def __del__(self):
        ctx = getattr(self, "ctx", None)
        if ctx:
            self.lib.dvg_ctx_destroy(ctx)
            self.ctx = None

    def render(self, scene: Scene, options: RenderOptions) -> np.ndarray:
        desc = SceneDesc(scene)
        out = np.zeros(scene.width * scene.height * 4, np.float32)
        st = self.lib.dvg_render(self.ctx, desc.ref(), C.byref(options.to_c()),
                                 out.ctypes.data_as(C.POINTER(C.c_float)))
        _check(self.lib, self.ctx, st)
        return out

    def render_sdf(self, scene: Scene, options: RenderOptions) -> np.ndarray:
        """Renderer::render_sdf: W * H signed distances (negative inside a fill)."""
        desc = SceneDesc(scene)
        out = np.zeros(scene.width * scene.height, np.float32)
        st = self.lib.dvg_render_sdf(self.ctx, desc.ref(), C.byref(options.to_c()),
                                     out.ctypes.data_as(C.POINTER(C.c_float)))
        _check(self.lib, self.ctx, st)
        return out

    def eval_positions(self, scene: Scene, positions: np.ndarray, options: RenderOptions) -> np.ndarray:
        """Renderer::eval_positions: the signed distance at each canvas point (n x 2)."""
        desc = SceneDesc(scene)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        out = np.zeros(pos.size // 2, np.float32)
        st = self.lib.dvg_eval_positions(self.ctx, desc.ref(), C.byref(options.to_c()),
                                         pos.ctypes.data_as(C.POINTER(C.c_float)), pos.size // 2,
                                         out.ctypes.data_as(C.POINTER(C.c_float)))
        _check(self.lib, self.ctx, st)
        return out

    def eval_positions_backward(self, scene: Scene, positions: np.ndarray, options: RenderOptions,
                                d_sdf: np.ndarray,
                                backward_options: BackwardOptions = BackwardOptions()) -> FlatGrads:
        """Renderer::eval_positions_backward: gradients of the signed distances
        at the points (n x 2) for the adjoint d_sdf (n)."""
        desc = SceneDesc(scene)
        pos = np.ascontiguousarray(positions, np.float32).reshape(-1)
        d_sdf = np.ascontiguousarray(d_sdf, np.float32).reshape(-1)
        if d_sdf.size != pos.size // 2:
            raise RenderError("InvalidScene", "d_sdf_image size mismatch for eval_positions")
        g = FlatGrads(desc, backward_options.compute_translation)
        st = self.lib.dvg_eval_positions_backward(self.ctx, desc.ref(), C.byref(options.to_c()),
                                                  C.byref(backward_options.to_c()),
                                                  pos.ctypes.data_as(C.POINTER(C.c_float)), d_sdf.size,
                                                  d_sdf.ctypes.data_as(C.POINTER(C.c_float)), g.ref())
        _check(self.lib, self.ctx, st)
        return g

    def render_backward_sdf(self, scene: Scene, options: RenderOptions, d_sdf: np.ndarray,
                            backward_options: BackwardOptions = BackwardOptions()) -> FlatGrads:
        """The d_sdf term of render_backward (W * H adjoints of the SDF image)."""
        desc = SceneDesc(scene)
        d_sdf = np.ascontiguousarray(d_sdf, np.float32).reshape(-1)
        if d_sdf.size != scene.width * scene.height:
            raise RenderError("InvalidScene", "d_sdf_image size mismatch")
        g = FlatGrads(desc, backward_options.compute_translation)
        st = self.lib.dvg_render_backward_sdf(self.ctx, desc.ref(), C.byref(options.to_c()),
                                              C.byref(backward_options.to_c()),
                                              d_sdf.ctypes.data_as(C.POINTER(C.c_float)), g.ref())
        _check(self.lib, self.ctx, st)
        return g

    def render_backward(self, scene: Scene, options: RenderOptions, d_render: np.ndarray,
                        backward_options: BackwardOptions = BackwardOptions()) -> FlatGrads:
        desc = SceneDesc(scene)
        d_render = np.ascontiguousarray(d_render, np.float32).reshape(-1)
        if d_render.size != scene.width * scene.height * 4:
            raise RenderError("InvalidScene", "d_render_image size mismatch")
        g = FlatGrads(desc, backward_options.compute_translation)
        st = self.lib.dvg_render_backward(self.ctx, desc.ref(), C.byref(options.to_c()),
                                          C.byref(backward_options.to_c()),
                                          d_render.ctypes.data_as(C.POINTER(C.c_float)), g.ref())
        _check(self.lib, self.ctx, st)
        return g

    # ---- device-resident path --------------------------------------------------
    def set_stream(self, stream_ptr: int):
        self.lib.dvg_ctx_set_stream(self.ctx, C.c_void_p(stream_ptr))

    def sync(self):
        _check(self.lib, self.ctx, self.lib.dvg_sync(self.ctx))

    def set_exact_boundary(self, enable: bool):
        """The checked boundary pass (the default): boundary samples whose f32
        inside test is uncertain are evaluated again in f64.  False selects
        the unchecked f32 pass."""
        _check(self.lib, self.ctx, self.lib.dvg_ctx_set_exact_boundary(self.ctx, int(enable)))

    def set_timing(self, enable: bool):
        self.lib.dvg_ctx_set_timing(self.ctx, int(enable))

    def kernel_times(self, reset=True):
        """{kernel name: (total ms, launches)} since the last reset."""
        n = len(abi.KERNEL_NAMES)
        ms = (C.c_double * n)()
        calls = (C.c_uint32 * n)()
        _check(self.lib, self.ctx, self.lib.dvg_kernel_times(self.ctx, ms, calls, n, int(reset)))
        return {abi.KERNEL_NAMES[k]: (ms[k], calls[k]) for k in range(n)}


class DeviceScene:
    """A packed scene uploaded once to HBM (dvg_device_scene_create).

    render()/render_backward() read and write torch CUDA tensors; they are
    asynchronous on the renderer's stream.
    """

    def __init__(self, renderer: Renderer, prepared: PreparedScene):
        self.r = renderer
        self.prepared = prepared
        h = C.c_void_p()
        st = renderer.lib.dvg_device_scene_create(renderer.ctx, C.byref(prepared.pack), C.byref(prepared.bpack),
                                                  C.byref(prepared.params), C.byref(h))
        _check(renderer.lib, renderer.ctx, st)
        self.handle = h
        self.width = prepared.params.width
        self.height = prepared.params.height

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.r.lib.dvg_device_scene_destroy(h)
            self.handle = None

    def update(self, path_points=None, group_colors=None):
        """Device-side re-packing from torch CUDA tensors: path points (P x 2,
        scene layout) and group colours (G x 8: fill then stroke RGBA)."""
        pp = C.c_void_p(path_points.data_ptr()) if path_points is not None else None
        gc = C.c_void_p(group_colors.data_ptr()) if group_colors is not None else None
        st = self.r.lib.dvg_device_scene_update(self.r.ctx, self.handle, pp, gc)
        _check(self.r.lib, self.r.ctx, st)

    def check(self):
        _check(self.r.lib, self.r.ctx, self.r.lib.dvg_device_scene_check(self.r.ctx, self.handle))

    def set_shapes(self, sblock_log2=2, bnd_log2=1, unit_log2=0):
        """Launch shapes (see dvg_device_scene_set_shapes)."""
        st = self.r.lib.dvg_device_scene_set_shapes(self.handle, sblock_log2, bnd_log2, unit_log2)
        _check(self.r.lib, self.r.ctx, st)

    def render(self, out) -> None:
        st = self.r.lib.dvg_device_render(self.r.ctx, self.handle, C.c_void_p(out.data_ptr()))
        _check(self.r.lib, self.r.ctx, st)

    def render_noalloc(self, out) -> None:
        """render() through dvg_device_render_noalloc: no device allocation;
        the work images live in the library's static per-device scratch
        (scenes of at most DVG_NOALLOC_MAX_PIXELS pixels)."""
        st = self.r.lib.dvg_device_render_noalloc(self.r.ctx, self.handle, C.c_void_p(out.data_ptr()))
        _check(self.r.lib, self.r.ctx, st)

    def render_sdf(self, out) -> None:
        """W * H signed distances into a device tensor (asynchronous)."""
        st = self.r.lib.dvg_device_render_sdf(self.r.ctx, self.handle, C.c_void_p(out.data_ptr()))
        _check(self.r.lib, self.r.ctx, st)

    def eval_positions(self, positions, out) -> None:
        """Signed distances at the points of a device tensor (n x 2) into out (n)."""
        n = positions.numel() // 2
        st = self.r.lib.dvg_device_eval_positions(self.r.ctx, self.handle, C.c_void_p(positions.data_ptr()), n,
                                                  C.c_void_p(out.data_ptr()))
        _check(self.r.lib, self.r.ctx, st)

    def grad_buffers(self, torch, device="cuda"):
        """Zeroed device gradient arrays with the lengths of the C ABI."""
        desc = self.prepared.desc
        host = FlatGrads(desc)
        bufs = {k: torch.zeros(max(host[k].size, 1), dtype=torch.float32, device=device) for k in GRAD_FIELDS}
        return bufs, {k: host[k].size for k in GRAD_FIELDS}

    def render_backward(self, d_render, bufs, sizes) -> None:
        g = abi.GradsC()
        for k in GRAD_FIELDS:
            ptr = bufs[k].data_ptr() if sizes[k] > 0 else None
            if k in ("d_background_image", "d_translation") and sizes[k] == 0:
                ptr = None
            setattr(g, k, C.cast(C.c_void_p(ptr), C.POINTER(C.c_float)) if ptr else
                    C.cast(None, C.POINTER(C.c_float)))
        st = self.r.lib.dvg_device_render_backward(self.r.ctx, self.handle, C.c_void_p(d_render.data_ptr()),
                                                   C.byref(g))
        _check(self.r.lib, self.r.ctx, st)
