"""The crate's gradient tree and the unpacking of the flat GPU gradients.

Types mirror src/grad.rs:9-126 (`SceneGrad` and its parts, names and meaning
kept), `SceneGrad::zeros_from_scene` / `accumulate_from` (src/grad.rs:130-419)
and `to_diffvg_layout` (src/grad.rs:474-560).  `apply_gpu_grads` is the
crate's unpacking of the 14 flat arrays a GPU backward returns into that tree
(src/renderer/backward_gpu.rs:599-864), reading the packed layout this
library's C ABI returns: shape parameters 8 a shape, path points and
thicknesses at the scene's point / thickness offsets, transforms 9 a shape /
group, paint colours at group_data[16 g + 8] (fill) and [+ 12] (stroke),
gradient geometry at d_gradient_data[8 k + 1 .. + 4] and stops at the
gradient's stop offset.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field
from typing import List, Optional, Union

import numpy as np

from . import abi
from .scene import (Circle, Ellipse, LinearGradient, Path, RadialGradient, Rect, Scene, SceneDesc, Solid)

Color = tuple


def _zero_color():
    return (0.0, 0.0, 0.0, 0.0)


def _add_color(a, b):
    return tuple(float(x) + float(y) for x, y in zip(a, b))


@dataclass
class DGradientStop:
    offset: float
    color: Color


@dataclass
class DLinearGradient:
    start: np.ndarray
    end: np.ndarray
    stops: List[DGradientStop]


@dataclass
class DRadialGradient:
    center: np.ndarray
    radius: np.ndarray
    stops: List[DGradientStop]


@dataclass
class DSolid:
    """DPaint::Solid."""
    color: Color


DPaint = Union[DSolid, DLinearGradient, DRadialGradient]


@dataclass
class DCircle:
    center: np.ndarray
    radius: float


@dataclass
class DEllipse:
    center: np.ndarray
    radius: np.ndarray


@dataclass
class DRect:
    min: np.ndarray
    max: np.ndarray


@dataclass
class DPath:
    points: np.ndarray                     # (P, 2)
    thickness: Optional[np.ndarray] = None


DShapeGeometry = Union[DCircle, DEllipse, DRect, DPath]


@dataclass
class DShape:
    geometry: DShapeGeometry
    transform: np.ndarray = field(default_factory=lambda: np.zeros((3, 3), np.float32))
    stroke_width: float = 0.0


@dataclass
class DShapeGroup:
    shape_to_canvas: np.ndarray = field(default_factory=lambda: np.zeros((3, 3), np.float32))
    fill: Optional[DPaint] = None
    stroke: Optional[DPaint] = None


@dataclass
class DFilter:
    radius: float = 0.0


def _v2(x=0.0, y=0.0):
    return np.array([x, y], np.float32)


def _zero_geometry(shape) -> DShapeGeometry:
    g = shape.geometry
    if isinstance(g, Circle):
        return DCircle(_v2(), 0.0)
    if isinstance(g, Ellipse):
        return DEllipse(_v2(), _v2())
    if isinstance(g, Rect):
        return DRect(_v2(), _v2())
    return DPath(np.zeros((len(g.points), 2), np.float32),
                 None if g.thickness is None else np.zeros(len(g.thickness), np.float32))


def _zero_paint(p) -> DPaint:
    if isinstance(p, Solid):
        return DSolid(_zero_color())
    stops = [DGradientStop(0.0, _zero_color()) for _ in p.stops]
    if isinstance(p, LinearGradient):
        return DLinearGradient(_v2(), _v2(), stops)
    return DRadialGradient(_v2(), _v2(), stops)


def _add_geometry(dst: DShapeGeometry, src: DShapeGeometry) -> None:
    if isinstance(dst, DCircle) and isinstance(src, DCircle):
        dst.center = dst.center + src.center
        dst.radius += src.radius
    elif isinstance(dst, DEllipse) and isinstance(src, DEllipse):
        dst.center = dst.center + src.center
        dst.radius = dst.radius + src.radius
    elif isinstance(dst, DRect) and isinstance(src, DRect):
        dst.min = dst.min + src.min
        dst.max = dst.max + src.max
    elif isinstance(dst, DPath) and isinstance(src, DPath):
        n = min(len(dst.points), len(src.points))
        dst.points[:n] += src.points[:n]
        if dst.thickness is not None:
            if src.thickness is not None:
                m = min(len(dst.thickness), len(src.thickness))
                dst.thickness[:m] += src.thickness[:m]
        elif src.thickness is not None:
            dst.thickness = src.thickness.copy()


def _add_paint(dst: DPaint, src: DPaint) -> None:
    if isinstance(dst, DSolid) and isinstance(src, DSolid):
        dst.color = _add_color(dst.color, src.color)
    elif ((isinstance(dst, DLinearGradient) and isinstance(src, DLinearGradient))
          or (isinstance(dst, DRadialGradient) and isinstance(src, DRadialGradient))):
        if isinstance(dst, DLinearGradient):
            dst.start, dst.end = dst.start + src.start, dst.end + src.end
        else:
            dst.center, dst.radius = dst.center + src.center, dst.radius + src.radius
        for i in range(min(len(dst.stops), len(src.stops))):
            dst.stops[i].offset += src.stops[i].offset
            dst.stops[i].color = _add_color(dst.stops[i].color, src.stops[i].color)


def _add_paint_option(dst: Optional[DPaint], src: Optional[DPaint]) -> Optional[DPaint]:
    if dst is not None:
        if src is not None:
            _add_paint(dst, src)
        return dst
    return copy.deepcopy(src)


def _add_slice_option(dst, src):
    if dst is not None:
        if src is not None:
            n = min(len(dst), len(src))
            dst[:n] += src[:n]
        return dst
    return None if src is None else np.array(src, np.float32, copy=True)


@dataclass
class DiffvgPathGrad:
    num_control_points: np.ndarray
    points: np.ndarray
    thickness: Optional[np.ndarray]
    is_closed: bool
    use_distance_approx: bool


@dataclass
class DiffvgShapeGrad:
    geometry: Union[DCircle, DEllipse, DRect, DiffvgPathGrad]
    transform: np.ndarray
    stroke_width: float


@dataclass
class DiffvgSceneGrad:
    shapes: List[DiffvgShapeGrad]
    shape_groups: List[DShapeGroup]
    filter: DFilter
    background: Color
    background_image: Optional[np.ndarray]
    translation: Optional[np.ndarray]


@dataclass
class SceneGrad:
    shapes: List[DShape]
    shape_groups: List[DShapeGroup]
    filter: DFilter = field(default_factory=DFilter)
    background: Color = (0.0, 0.0, 0.0, 0.0)
    background_image: Optional[np.ndarray] = None
    translation: Optional[np.ndarray] = None

    @staticmethod
    def zeros_from_scene(scene: Scene, include_background_image: bool = False,
                         include_translation: bool = False) -> "SceneGrad":
        """SceneGrad::zeros_from_scene (src/grad.rs:130-183)."""
        shapes = [DShape(_zero_geometry(s)) for s in scene.shapes]
        groups = [DShapeGroup(fill=None if g.fill is None else _zero_paint(g.fill),
                              stroke=None if g.stroke is None else _zero_paint(g.stroke)) for g in scene.groups]
        bg = None
        if include_background_image and scene.background_image is not None:
            bg = np.zeros(np.asarray(scene.background_image).size, np.float32)
        tr = np.zeros(scene.width * scene.height * 2, np.float32) if include_translation else None
        return SceneGrad(shapes, groups, DFilter(0.0), (0.0, 0.0, 0.0, 0.0), bg, tr)

    def accumulate_from(self, other: "SceneGrad") -> None:
        """SceneGrad::accumulate_from (src/grad.rs:190-213): element-wise add,
        variable-length parts up to the shorter length, missing optional parts
        taken from `other`."""
        for d, s in zip(self.shapes, other.shapes):
            d.stroke_width += s.stroke_width
            d.transform = d.transform + s.transform
            _add_geometry(d.geometry, s.geometry)
        for d, s in zip(self.shape_groups, other.shape_groups):
            d.shape_to_canvas = d.shape_to_canvas + s.shape_to_canvas
            d.fill = _add_paint_option(d.fill, s.fill)
            d.stroke = _add_paint_option(d.stroke, s.stroke)
        self.filter.radius += other.filter.radius
        self.background = _add_color(self.background, other.background)
        self.background_image = _add_slice_option(self.background_image, other.background_image)
        self.translation = _add_slice_option(self.translation, other.translation)

    def to_diffvg_layout(self, scene: Scene) -> DiffvgSceneGrad:
        """SceneGrad::to_diffvg_layout (src/grad.rs:474-560): path metadata
        copied from the scene, per-point buffers sized to the scene's path."""
        shapes = []
        for i, d in enumerate(self.shapes):
            sg = scene.shapes[i].geometry if i < len(scene.shapes) else None
            g = d.geometry
            if isinstance(g, DPath):
                if isinstance(sg, Path):
                    pts = np.zeros((len(sg.points), 2), np.float32)
                    n = min(len(pts), len(g.points))
                    pts[:n] = g.points[:n]
                    if g.thickness is not None and sg.thickness is not None:
                        th = np.zeros(len(sg.thickness), np.float32)
                        m = min(len(th), len(g.thickness))
                        th[:m] = g.thickness[:m]
                    elif g.thickness is not None:
                        th = g.thickness.copy()
                    elif sg.thickness is not None:
                        th = np.zeros(len(sg.thickness), np.float32)
                    else:
                        th = None
                    geom = DiffvgPathGrad(sg.num_control_points.copy(), pts, th, bool(sg.is_closed),
                                          bool(sg.use_distance_approx))
                else:
                    geom = DiffvgPathGrad(np.zeros(0, np.uint8), g.points.copy(),
                                          None if g.thickness is None else g.thickness.copy(), False, False)
            else:
                geom = copy.deepcopy(g)
            shapes.append(DiffvgShapeGrad(geom, d.transform.copy(), d.stroke_width))
        return DiffvgSceneGrad(shapes, copy.deepcopy(self.shape_groups), DFilter(self.filter.radius),
                               self.background,
                               None if self.background_image is None else self.background_image.copy(),
                               None if self.translation is None else self.translation.copy())


def _mat3(a, base):
    return np.asarray(a[base:base + 9], np.float32).reshape(3, 3).copy()


def _paint_grad(paint, kind: int, grad_index: int, color_base: int, desc: SceneDesc, flat) -> DPaint:
    """build_paint_grad (src/renderer/backward_gpu.rs:803-864)."""
    d_group_data = flat["d_group_data"]
    if isinstance(paint, Solid):
        c = d_group_data[color_base:color_base + 4]
        return DSolid(tuple(float(x) for x in c) if len(c) == 4 else _zero_color())
    if kind == abi.PAINT_NONE:
        return DSolid(_zero_color())
    dg = flat["d_gradient_data"]
    base = grad_index * 8
    p = [float(dg[base + k]) if base + k < len(dg) else 0.0 for k in range(1, 5)]
    so = int(desc.gradient_stop_offset[grad_index])
    sc = int(desc.gradient_stop_count[grad_index])
    d_off, d_col = flat["d_stop_offsets"], flat["d_stop_colors"]
    stops = []
    for i in range(sc):
        o = so + i
        off = float(d_off[o]) if o < len(d_off) else 0.0
        col = tuple(float(d_col[4 * o + c]) if 4 * o + c < len(d_col) else 0.0 for c in range(4))
        stops.append(DGradientStop(off, col))
    if int(desc.gradient_kind[grad_index]) == 0:
        return DLinearGradient(_v2(p[0], p[1]), _v2(p[2], p[3]), stops)
    return DRadialGradient(_v2(p[0], p[1]), _v2(p[2], p[3]), stops)


def apply_gpu_grads(scene: Scene, desc: SceneDesc, flat, grads: SceneGrad) -> SceneGrad:
    """apply_gpu_grads (src/renderer/backward_gpu.rs:599-801): the flat arrays
    `flat` (a FlatGrads or a mapping of the 14 field names) written into
    `grads` (a SceneGrad sized for `scene`), which is returned."""
    d_params, d_points = flat["d_shape_params"], flat["d_shape_points"]
    d_thick, d_sw = flat["d_shape_thickness"], flat["d_shape_stroke_width"]
    d_st, d_gt = flat["d_shape_transform"], flat["d_group_transform"]
    for i, shape in enumerate(scene.shapes):
        d = grads.shapes[i]
        if i < len(d_sw):
            d.stroke_width = float(d_sw[i])
        if 9 * i + 8 < len(d_st):
            d.transform = _mat3(d_st, 9 * i)
        g, base = shape.geometry, 8 * i
        if isinstance(g, Circle) and base + 2 < len(d_params) and isinstance(d.geometry, DCircle):
            d.geometry.center = _v2(d_params[base], d_params[base + 1])
            d.geometry.radius = float(d_params[base + 2])
        elif isinstance(g, Ellipse) and base + 3 < len(d_params) and isinstance(d.geometry, DEllipse):
            d.geometry.center = _v2(d_params[base], d_params[base + 1])
            d.geometry.radius = _v2(d_params[base + 2], d_params[base + 3])
        elif isinstance(g, Rect) and base + 3 < len(d_params) and isinstance(d.geometry, DRect):
            d.geometry.min = _v2(d_params[base], d_params[base + 1])
            d.geometry.max = _v2(d_params[base + 2], d_params[base + 3])
        elif isinstance(g, Path) and isinstance(d.geometry, DPath):
            p_off, p_cnt = int(desc.p_off[i]), int(desc.p_cnt[i])
            for k in range(min(p_cnt, len(d.geometry.points))):
                b = (p_off + k) * 2
                if b + 1 < len(d_points):
                    d.geometry.points[k] = (d_points[b], d_points[b + 1])
            if d.geometry.thickness is not None:
                t_off, t_cnt = int(desc.t_off[i]), int(desc.t_cnt[i])
                for k in range(min(t_cnt, len(d.geometry.thickness))):
                    if t_off + k < len(d_thick):
                        d.geometry.thickness[k] = d_thick[t_off + k]
    for gi, group in enumerate(scene.groups):
        d = grads.shape_groups[gi]
        if 9 * gi + 8 < len(d_gt):
            d.shape_to_canvas = _mat3(d_gt, 9 * gi)
        gb = 16 * gi
        if group.fill is not None:
            d.fill = _paint_grad(group.fill, int(desc.fill_kind[gi]), int(desc.fill_gradient[gi]), gb + 8, desc, flat)
        if group.stroke is not None:
            d.stroke = _paint_grad(group.stroke, int(desc.stroke_kind[gi]), int(desc.stroke_gradient[gi]), gb + 12,
                                   desc, flat)
    if len(flat["d_filter_radius"]) > 0:
        grads.filter.radius = float(flat["d_filter_radius"][0])
    bgd = flat["d_background"]
    if len(bgd) >= 4:
        grads.background = tuple(float(x) for x in bgd[:4])
    if grads.background_image is not None and len(flat["d_background_image"]) >= len(grads.background_image):
        grads.background_image[:] = flat["d_background_image"][:len(grads.background_image)]
    if grads.translation is not None and len(flat["d_translation"]) >= len(grads.translation):
        grads.translation[:] = flat["d_translation"][:len(grads.translation)]
    return grads
