"""Python mirror of the crate's scene / shape / grad API.

Types follow src/scene.rs:1-291, src/geometry.rs:78-89, src/grad.rs:9-126 and
src/renderer/types.rs:10-39 (names and meaning kept).  `SceneDesc` flattens a
Scene into the `dvg_scene_desc` struct of include/dvg.h; `FlatGrads` holds the
14 flat gradient arrays of the C ABI (the layout apply_gpu_grads reads,
src/renderer/backward_gpu.rs:599-801); `FlatGrads.scene_grad()` unpacks them
into the `SceneGrad` tree of pydvg.grad.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import abi

Vec2 = Tuple[float, float]
Color = Tuple[float, float, float, float]


def identity3() -> np.ndarray:
    return np.eye(3, dtype=np.float32)


def mat3_inverse(m: np.ndarray) -> Optional[np.ndarray]:
    """Mat3::inverse (src/math.rs:269-290), in f32 with the reference's order."""
    m = np.asarray(m, dtype=np.float32)
    a, b, c, d = m[0, 0], m[0, 1], m[1, 0], m[1, 1]
    tx, ty = m[0, 2], m[1, 2]
    det = np.float32(a * d - b * c)
    if abs(det) < np.float32(1e-6):
        return None
    inv_det = np.float32(1.0) / det
    m00, m01, m10, m11 = d * inv_det, -b * inv_det, -c * inv_det, a * inv_det
    m02 = -(m00 * tx + m01 * ty)
    m12 = -(m10 * tx + m11 * ty)
    return np.array([[m00, m01, m02], [m10, m11, m12], [0, 0, 1]], dtype=np.float32)


# ---- scene.rs ---------------------------------------------------------------
@dataclass
class Filter:
    filter_type: int = abi.FILTER_BOX
    radius: float = 0.5


@dataclass
class GradientStop:
    offset: float
    color: Color


@dataclass
class LinearGradient:
    start: Vec2
    end: Vec2
    stops: List[GradientStop]


@dataclass
class RadialGradient:
    center: Vec2
    radius: Vec2
    stops: List[GradientStop]


@dataclass
class Solid:
    color: Color


Paint = Union[Solid, LinearGradient, RadialGradient]


@dataclass
class Circle:
    center: Vec2
    radius: float


@dataclass
class Ellipse:
    center: Vec2
    radius: Vec2


@dataclass
class Rect:
    min: Vec2
    max: Vec2


@dataclass
class Path:
    """diffvg path encoding (src/geometry.rs:78-89)."""
    num_control_points: np.ndarray  # uint8 per segment: 0 line, 1 quad, 2 cubic
    points: np.ndarray              # (P, 2) float32
    thickness: Optional[np.ndarray] = None
    is_closed: bool = False
    use_distance_approx: bool = False

    def __post_init__(self):
        self.num_control_points = np.asarray(self.num_control_points, dtype=np.uint8).reshape(-1)
        self.points = np.asarray(self.points, dtype=np.float32).reshape(-1, 2)
        if self.thickness is not None:
            self.thickness = np.asarray(self.thickness, dtype=np.float32).reshape(-1)


Geometry = Union[Circle, Ellipse, Rect, Path]


@dataclass
class Shape:
    geometry: Geometry
    transform: np.ndarray = field(default_factory=identity3)
    stroke_width: float = 0.0
    stroke_join: int = abi.JOIN_ROUND
    stroke_cap: int = abi.CAP_ROUND
    stroke_miter_limit: float = 4.0


@dataclass
class ShapeGroup:
    shape_indices: List[int]
    fill: Optional[Paint] = None
    stroke: Optional[Paint] = None
    opacity: float = 1.0
    fill_rule: int = abi.FILL_NONZERO
    shape_to_canvas: np.ndarray = field(default_factory=identity3)
    canvas_to_shape: np.ndarray = field(default_factory=identity3)

    def set_shape_to_canvas(self, m):
        self.shape_to_canvas = np.asarray(m, dtype=np.float32)
        inv = mat3_inverse(self.shape_to_canvas)
        self.canvas_to_shape = inv if inv is not None else identity3()


@dataclass
class Scene:
    width: int
    height: int
    shapes: List[Shape] = field(default_factory=list)
    groups: List[ShapeGroup] = field(default_factory=list)
    background: Color = (0.0, 0.0, 0.0, 0.0)
    background_image: Optional[np.ndarray] = None
    filter: Filter = field(default_factory=Filter)


@dataclass
class RenderOptions:
    """src/renderer/types.rs:10-39 (defaults included)."""
    aa: float = 1.0
    path_tolerance: float = 0.5
    samples_x: int = 2
    samples_y: int = 2
    seed: int = 0
    jitter: bool = True
    use_prefiltering: bool = True

    def to_c(self) -> abi.RenderOptions:
        return abi.RenderOptions(self.aa, self.path_tolerance, self.samples_x, self.samples_y,
                                 self.seed, int(self.jitter), int(self.use_prefiltering))


@dataclass
class BackwardOptions:
    compute_translation: bool = False

    def to_c(self) -> abi.BackwardOptions:
        return abi.BackwardOptions(int(self.compute_translation))


# ---- flattening into dvg_scene_desc ---------------------------------------------
def _ptr(arr: np.ndarray, ctype):
    if arr is None or arr.size == 0:
        return C.cast(None, C.POINTER(ctype))
    return arr.ctypes.data_as(C.POINTER(ctype))


class SceneDesc:
    """Flat arrays of a Scene plus the ctypes struct pointing into them.

    Gradient paints are listed in the order the packer meets them (each group's
    fill, then its stroke; src/renderer/prepare.rs:262-273), so gradient and
    stop indices here equal the packed indices.
    """

    def __init__(self, scene: Scene):
        S, G = len(scene.shapes), len(scene.groups)
        self.scene = scene
        self.num_shapes, self.num_groups = S, G
        kind = np.zeros(S, np.uint32)
        params = np.zeros((S, 4), np.float32)
        xf = np.zeros((S, 9), np.float32)
        sw = np.zeros(S, np.float32)
        join = np.zeros(S, np.uint32)
        cap = np.zeros(S, np.uint32)
        miter = np.zeros(S, np.float32)
        p_off = np.zeros(S, np.uint32)
        p_cnt = np.zeros(S, np.uint32)
        c_off = np.zeros(S, np.uint32)
        c_cnt = np.zeros(S, np.uint32)
        t_off = np.zeros(S, np.uint32)
        t_cnt = np.zeros(S, np.uint32)
        closed = np.zeros(S, np.uint32)
        approx = np.zeros(S, np.uint32)
        pts, ctrls, thick = [], [], []
        npts = nctrl = nthick = 0
        for i, sh in enumerate(scene.shapes):
            g = sh.geometry
            xf[i] = np.asarray(sh.transform, np.float32).reshape(9)
            sw[i] = sh.stroke_width
            join[i] = sh.stroke_join
            cap[i] = sh.stroke_cap
            miter[i] = sh.stroke_miter_limit
            if isinstance(g, Circle):
                kind[i] = abi.SHAPE_CIRCLE
                params[i] = (g.center[0], g.center[1], g.radius, 0.0)
            elif isinstance(g, Ellipse):
                kind[i] = abi.SHAPE_ELLIPSE
                params[i] = (g.center[0], g.center[1], g.radius[0], g.radius[1])
            elif isinstance(g, Rect):
                kind[i] = abi.SHAPE_RECT
                params[i] = (g.min[0], g.min[1], g.max[0], g.max[1])
            else:
                kind[i] = abi.SHAPE_PATH
                p_off[i], p_cnt[i] = npts, len(g.points)
                c_off[i], c_cnt[i] = nctrl, len(g.num_control_points)
                pts.append(g.points)
                ctrls.append(g.num_control_points.astype(np.uint32))
                npts += len(g.points)
                nctrl += len(g.num_control_points)
                if g.thickness is not None:
                    t_off[i], t_cnt[i] = nthick, len(g.thickness)
                    thick.append(g.thickness)
                    nthick += len(g.thickness)
                else:
                    t_off[i] = nthick
                closed[i] = int(g.is_closed)
                approx[i] = int(g.use_distance_approx)
        self.path_points = (np.concatenate(pts).astype(np.float32).reshape(-1) if pts
                            else np.zeros(0, np.float32))
        self.path_num_controls = (np.concatenate(ctrls).astype(np.uint32) if ctrls
                                  else np.zeros(0, np.uint32))
        self.path_thickness = (np.concatenate(thick).astype(np.float32) if thick
                               else np.zeros(0, np.float32))
        self.shape_kind, self.shape_params, self.shape_transform = kind, params.reshape(-1), xf.reshape(-1)
        self.stroke_width, self.stroke_join, self.stroke_cap, self.stroke_miter = sw, join, cap, miter
        self.p_off, self.p_cnt, self.c_off, self.c_cnt = p_off, p_cnt, c_off, c_cnt
        self.t_off, self.t_cnt, self.closed, self.approx = t_off, t_cnt, closed, approx

        g_off = np.zeros(G, np.uint32)
        g_cnt = np.zeros(G, np.uint32)
        idx = []
        fk = np.zeros(G, np.uint32)
        sk = np.zeros(G, np.uint32)
        fc = np.zeros((G, 4), np.float32)
        scol = np.zeros((G, 4), np.float32)
        fg = np.zeros(G, np.uint32)
        sg = np.zeros(G, np.uint32)
        rule = np.zeros(G, np.uint32)
        s2c = np.zeros((G, 9), np.float32)
        c2s = np.zeros((G, 9), np.float32)
        opac = np.zeros(G, np.float32)
        gk, gp, gso, gsc, so, scl = [], [], [], [], [], []
        self.paint_gradient_index = []  # (group, is_stroke) -> gradient index

        def add_paint(p, k_arr, c_arr, g_arr, gi):
            if p is None:
                k_arr[gi] = abi.PAINT_NONE
                return
            if isinstance(p, Solid):
                k_arr[gi] = abi.PAINT_SOLID
                c_arr[gi] = p.color
                return
            k_arr[gi] = abi.PAINT_LINEAR if isinstance(p, LinearGradient) else abi.PAINT_RADIAL
            g_arr[gi] = len(gk)
            gk.append(0 if isinstance(p, LinearGradient) else 1)
            if isinstance(p, LinearGradient):
                gp.append((p.start[0], p.start[1], p.end[0], p.end[1]))
            else:
                gp.append((p.center[0], p.center[1], p.radius[0], p.radius[1]))
            gso.append(len(so))
            gsc.append(len(p.stops))
            for st in p.stops:
                so.append(st.offset)
                scl.append(st.color)

        for gi, grp in enumerate(scene.groups):
            g_off[gi] = len(idx)
            g_cnt[gi] = len(grp.shape_indices)
            idx.extend(grp.shape_indices)
            add_paint(grp.fill, fk, fc, fg, gi)
            add_paint(grp.stroke, sk, scol, sg, gi)
            rule[gi] = grp.fill_rule
            s2c[gi] = np.asarray(grp.shape_to_canvas, np.float32).reshape(9)
            c2s[gi] = np.asarray(grp.canvas_to_shape, np.float32).reshape(9)
            opac[gi] = grp.opacity
        self.group_shape_offset, self.group_shape_count = g_off, g_cnt
        self.group_shape_indices = np.asarray(idx, np.uint32)
        self.fill_kind, self.stroke_kind = fk, sk
        self.fill_color, self.stroke_color = fc.reshape(-1), scol.reshape(-1)
        self.fill_gradient, self.stroke_gradient, self.fill_rule = fg, sg, rule
        self.shape_to_canvas, self.canvas_to_shape, self.opacity = s2c.reshape(-1), c2s.reshape(-1), opac
        self.num_gradients = len(gk)
        self.gradient_kind = np.asarray(gk, np.uint32)
        self.gradient_params = np.asarray(gp, np.float32).reshape(-1)
        self.gradient_stop_offset = np.asarray(gso, np.uint32)
        self.gradient_stop_count = np.asarray(gsc, np.uint32)
        self.num_stops = len(so)
        self.stop_offsets = np.asarray(so, np.float32)
        self.stop_colors = np.asarray(scl, np.float32).reshape(-1)
        self.bg_image = (None if scene.background_image is None
                         else np.ascontiguousarray(scene.background_image, np.float32).reshape(-1))

        f32, u32 = C.c_float, C.c_uint32
        d = abi.SceneDescC()
        d.width, d.height = scene.width, scene.height
        d.background[:] = [float(v) for v in scene.background]
        d.background_image = _ptr(self.bg_image, f32)
        d.background_image_len = 0 if self.bg_image is None else self.bg_image.size
        d.filter_type, d.filter_radius = scene.filter.filter_type, scene.filter.radius
        d.num_shapes = S
        d.shape_kind, d.shape_params, d.shape_transform = _ptr(kind, u32), _ptr(self.shape_params, f32), _ptr(self.shape_transform, f32)
        d.stroke_width, d.stroke_join, d.stroke_cap = _ptr(sw, f32), _ptr(join, u32), _ptr(cap, u32)
        d.stroke_miter_limit = _ptr(miter, f32)
        d.path_point_offset, d.path_point_count = _ptr(p_off, u32), _ptr(p_cnt, u32)
        d.path_ctrl_offset, d.path_ctrl_count = _ptr(c_off, u32), _ptr(c_cnt, u32)
        d.path_thickness_offset, d.path_thickness_count = _ptr(t_off, u32), _ptr(t_cnt, u32)
        d.path_is_closed, d.path_use_distance_approx = _ptr(closed, u32), _ptr(approx, u32)
        d.path_points, d.path_points_len = _ptr(self.path_points, f32), self.path_points.size
        d.path_num_controls, d.path_num_controls_len = _ptr(self.path_num_controls, u32), self.path_num_controls.size
        d.path_thickness, d.path_thickness_len = _ptr(self.path_thickness, f32), self.path_thickness.size
        d.num_groups = G
        d.group_shape_offset, d.group_shape_count = _ptr(g_off, u32), _ptr(g_cnt, u32)
        d.group_shape_indices, d.group_shape_indices_len = _ptr(self.group_shape_indices, u32), self.group_shape_indices.size
        d.fill_kind, d.stroke_kind = _ptr(fk, u32), _ptr(sk, u32)
        d.fill_color, d.stroke_color = _ptr(self.fill_color, f32), _ptr(self.stroke_color, f32)
        d.fill_gradient, d.stroke_gradient, d.fill_rule = _ptr(fg, u32), _ptr(sg, u32), _ptr(rule, u32)
        d.shape_to_canvas, d.canvas_to_shape = _ptr(self.shape_to_canvas, f32), _ptr(self.canvas_to_shape, f32)
        d.opacity = _ptr(opac, f32)
        d.num_gradients = self.num_gradients
        d.gradient_kind, d.gradient_params = _ptr(self.gradient_kind, u32), _ptr(self.gradient_params, f32)
        d.gradient_stop_offset, d.gradient_stop_count = _ptr(self.gradient_stop_offset, u32), _ptr(self.gradient_stop_count, u32)
        d.num_stops = self.num_stops
        d.stop_offsets, d.stop_colors = _ptr(self.stop_offsets, f32), _ptr(self.stop_colors, f32)
        self.c = d

    def ref(self):
        return C.byref(self.c)


# ---- gradients ------------------------------------------------------------------
GRAD_FIELDS = ("d_shape_params", "d_shape_points", "d_shape_thickness", "d_shape_stroke_width",
               "d_shape_transform", "d_group_transform", "d_group_data", "d_gradient_data",
               "d_stop_offsets", "d_stop_colors", "d_filter_radius", "d_background",
               "d_background_image", "d_translation")


class FlatGrads:
    """The 14 flat gradient arrays of the C ABI, sized for one scene."""

    def __init__(self, desc: SceneDesc, translation: bool = False):
        sc = desc.scene
        S, G = desc.num_shapes, desc.num_groups
        self.arrays = {
            "d_shape_params": np.zeros(8 * S, np.float32),
            "d_shape_points": np.zeros(desc.path_points.size, np.float32),
            "d_shape_thickness": np.zeros(desc.path_thickness.size, np.float32),
            "d_shape_stroke_width": np.zeros(S, np.float32),
            "d_shape_transform": np.zeros(9 * S, np.float32),
            "d_group_transform": np.zeros(9 * G, np.float32),
            "d_group_data": np.zeros(16 * G, np.float32),
            "d_gradient_data": np.zeros(8 * desc.num_gradients, np.float32),
            "d_stop_offsets": np.zeros(desc.num_stops, np.float32),
            "d_stop_colors": np.zeros(4 * desc.num_stops, np.float32),
            "d_filter_radius": np.zeros(1, np.float32),
            "d_background": np.zeros(4, np.float32),
            "d_background_image": (np.zeros(sc.width * sc.height * 4, np.float32)
                                   if sc.background_image is not None else np.zeros(0, np.float32)),
            "d_translation": (np.zeros(sc.width * sc.height * 2, np.float32)
                              if translation else np.zeros(0, np.float32)),
        }
        self.desc = desc
        self.c = abi.GradsC()
        for k in GRAD_FIELDS:
            setattr(self.c, k, _ptr(self.arrays[k], C.c_float))

    def scene_grad(self):
        """The arrays unpacked into a SceneGrad of the scene (apply_gpu_grads,
        src/renderer/backward_gpu.rs:599-801), with the background-image and
        translation parts when these arrays hold them."""
        from .grad import SceneGrad, apply_gpu_grads
        sc = self.desc.scene
        g = SceneGrad.zeros_from_scene(sc, self.arrays["d_background_image"].size > 0,
                                       self.arrays["d_translation"].size > 0)
        return apply_gpu_grads(sc, self.desc, self, g)

    def __getitem__(self, k):
        return self.arrays[k]

    def ref(self):
        return C.byref(self.c)

    def vector(self, fields: Sequence[str] = GRAD_FIELDS) -> np.ndarray:
        return np.concatenate([self.arrays[k] for k in fields])


def finalize_background_gradients(scene: Scene, grads) -> None:
    """finalize_background_gradients (src/backward/background.rs:29-61), in
    place on flat gradients: the premultiplied background adjoint that the
    packed seam, the device scenes and the batch return becomes the adjoint
    of the straight colour (rgb times alpha; alpha from every channel).  The
    host entry points (dvg_render_backward) apply it themselves."""
    if scene.background_image is not None:
        d = np.asarray(grads["d_background_image"]).reshape(-1, 4)
        if d.size == 0:
            return
        im = np.asarray(scene.background_image, np.float32).reshape(-1, 4)
        r, g, b, a = im[:, 0], im[:, 1], im[:, 2], im[:, 3]
        da = d[:, 0] * r + d[:, 1] * g + d[:, 2] * b + d[:, 3]
        d[:, 0], d[:, 1], d[:, 2] = d[:, 0] * a, d[:, 1] * a, d[:, 2] * a
        d[:, 3] = da
    else:
        d = grads["d_background"]
        r, g, b, a = (np.float32(v) for v in scene.background)
        d0, d1, d2, d3 = (np.float32(v) for v in d[:4])
        d[0], d[1], d[2], d[3] = d0 * a, d1 * a, d2 * a, d0 * r + d1 * g + d2 * b + d3
