"""Seeded synthetic scenes for the benchmark workloads and parity tests.

Every generator draws from numpy's PCG64 stream seeded with `seed` (a stream
separate from the renderer's per-sample PCG32 streams) and builds one shape per
group, shapes listed in group order, so that prepared and scene shape indices
coincide.

  headline(): 1024x1024, 1000 closed cubic paths of 3-8 segments; centre
      uniform in the image, control points = centre + U(-64, 64)^2 px; fill
      non-zero, RGB ~ U[0,1), alpha ~ U[0.3,1.0); white background; 2x2
      jittered samples; box filter r = 0.5; no prefiltering.
  config1(): 256x256, 64 closed cubic paths of 3-8 segments, control points
      uniform in the image, 4x4 samples, box r = 0.5.
  config2(): 512x512, 512 mixed shapes (circles, ellipses, rects with radii
      5-60 px, quadratic and cubic open/closed paths), 30 % stroked with width
      1-8 px, fill rules 50/50, paints 50 % solid / 25 % linear / 25 % radial
      with 2-4 stops, 2x2 samples, tent filter r = 1.
  config3(): 1024x1024 prefiltered, 1 sample per pixel; 2000 stroked open cubic
      paths (4-16 segments, width 0.5-4, random walk N(0, 20 px)) and 500
      filled polygons of 3-12 vertices.
  config4(): 64 scenes of config2 shapes, path counts log-uniform in 16-4096,
      resolutions from {128, 256, 512, 1024}^2.
  config5(): 2048x2048, 10000 closed cubic paths, 2x2 samples.
  config5_reduced(): the config5 generator at 256x256 with the path density
      kept (156 paths): the size at which the CPU oracle runs its Adam loop.
  gaussian_target(): the procedural optimisation target of config5, a sum of
      256 random anisotropic Gaussians in RGB.
  jittered(scene, sigma, seed): the same scene with every path point moved by
      N(0, sigma) -- the optimisation target of configs 1 and 5.
"""
from __future__ import annotations

import copy

import numpy as np

from . import abi
from .scene import (Circle, Ellipse, Filter, GradientStop, LinearGradient, Path, RadialGradient, Rect,
                    RenderOptions, Scene, Shape, ShapeGroup, Solid)

WHITE = (1.0, 1.0, 1.0, 1.0)


def _closed_cubic(rng, n_seg, center, spread, clip=None):
    pts = center + rng.uniform(-spread, spread, size=(3 * n_seg, 2))
    if clip is not None:
        pts = np.clip(pts, 0.0, clip)
    return Path(np.full(n_seg, 2, np.uint8), pts.astype(np.float32), is_closed=True)


def _rgba(rng):
    rgb = rng.uniform(0.0, 1.0, size=3)
    a = rng.uniform(0.3, 1.0)
    return (float(rgb[0]), float(rgb[1]), float(rgb[2]), float(a))


def cubic_paths_scene(width, height, n_paths, seed, spread=None, seg_range=(3, 8)):
    """Closed cubic paths, one solid-filled group each (headline / configs 1, 5)."""
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, background=WHITE, filter=Filter(abi.FILTER_BOX, 0.5))
    for i in range(n_paths):
        n_seg = int(rng.integers(seg_range[0], seg_range[1] + 1))
        if spread is None:
            path = Path(np.full(n_seg, 2, np.uint8),
                        (rng.uniform(0, 1, size=(3 * n_seg, 2)) * (width, height)).astype(np.float32),
                        is_closed=True)
        else:
            center = rng.uniform(0, 1, size=2) * (width, height)
            path = _closed_cubic(rng, n_seg, center, spread)
        sc.shapes.append(Shape(path))
        sc.groups.append(ShapeGroup([i], fill=Solid(_rgba(rng))))
    return sc


def headline(seed=0):
    sc = cubic_paths_scene(1024, 1024, 1000, seed, spread=64.0)
    opts = RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts


def config1(seed=0):
    sc = cubic_paths_scene(256, 256, 64, seed, spread=None)
    opts = RenderOptions(samples_x=4, samples_y=4, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts


def config5(seed=0):
    sc = cubic_paths_scene(2048, 2048, 10000, seed, spread=64.0)
    opts = RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts


def config5_reduced(seed=0):
    sc = cubic_paths_scene(256, 256, 156, seed, spread=64.0)
    opts = RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts


def gaussian_target(width, height, n=256, seed=0):
    """RGBA float32 image (H*W*4, alpha 1): RGB = clip(sum_k a_k g_k(x), 0, 1)
    with g_k(x) = exp(-0.5 (x - c_k)^T R_k diag(s_k)^-2 R_k^T (x - c_k)),
    centre c_k uniform in the image, standard deviations s_k uniform in
    [20, 300] px x width / 2048, rotation angle uniform in [0, pi), amplitude
    a_k uniform in [-0.15, 0.25]^3 (mean image value about 0.5), evaluated in
    float64 at pixel centres within
    4.5 standard deviations of the centre (the rest is below 4e-5)."""
    rng = np.random.default_rng(seed)
    img = np.zeros((height, width, 3), np.float64)
    for _ in range(n):
        cx, cy = rng.uniform(0, width), rng.uniform(0, height)
        sx, sy = rng.uniform(20.0, 300.0, size=2) * (width / 2048.0)
        th = rng.uniform(0.0, np.pi)
        amp = rng.uniform(-0.15, 0.25, size=3)
        r = 4.5 * max(sx, sy)
        x0, x1 = max(int(cx - r), 0), min(int(cx + r) + 1, width)
        y0, y1 = max(int(cy - r), 0), min(int(cy + r) + 1, height)
        if x0 >= x1 or y0 >= y1:
            continue
        ys, xs = np.mgrid[y0:y1, x0:x1]
        dx, dy = xs + 0.5 - cx, ys + 0.5 - cy
        c, s = np.cos(th), np.sin(th)
        u, v = (c * dx + s * dy) / sx, (-s * dx + c * dy) / sy
        img[y0:y1, x0:x1] += np.exp(-0.5 * (u * u + v * v))[..., None] * amp
    out = np.ones((height, width, 4), np.float32)
    out[..., :3] = np.clip(img, 0.0, 1.0)
    return out.reshape(-1)


def _stops(rng):
    n = int(rng.integers(2, 5))
    offs = np.sort(rng.uniform(0, 1, size=n))
    offs[0], offs[-1] = 0.0, 1.0
    return [GradientStop(float(o), _rgba(rng)) for o in offs]


def _paint(rng, cx, cy, r):
    u = rng.uniform()
    if u < 0.5:
        return Solid(_rgba(rng))
    if u < 0.75:
        ang = rng.uniform(0, 2 * np.pi)
        d = np.array([np.cos(ang), np.sin(ang)]) * r
        return LinearGradient((float(cx - d[0]), float(cy - d[1])), (float(cx + d[0]), float(cy + d[1])),
                              _stops(rng))
    return RadialGradient((float(cx), float(cy)), (float(r), float(r * rng.uniform(0.5, 1.5))), _stops(rng))


def mixed_shape(rng, width, height):
    """One shape of configs 2 and 4, with its paint choices."""
    cx, cy = rng.uniform(0, width), rng.uniform(0, height)
    kind = int(rng.integers(0, 6))
    r = float(rng.uniform(5, 60))
    if kind == 0:
        geom = Circle((float(cx), float(cy)), r)
    elif kind == 1:
        geom = Ellipse((float(cx), float(cy)), (r, float(rng.uniform(5, 60))))
    elif kind == 2:
        h = float(rng.uniform(5, 60))
        geom = Rect((float(cx - r), float(cy - h)), (float(cx + r), float(cy + h)))
    else:
        n_seg = int(rng.integers(2, 7))
        nc = 1 if kind in (3, 4) else 2
        closed = kind in (3, 5)
        # a closed path wraps its last segment to point 0; an open one ends on
        # an extra point
        n_pts = (nc + 1) * n_seg if closed else (nc + 1) * n_seg + 1
        pts = np.array([cx, cy]) + rng.uniform(-r, r, size=(n_pts, 2))
        geom = Path(np.full(n_seg, nc, np.uint8), pts.astype(np.float32), is_closed=closed)
    stroked = rng.uniform() < 0.3
    shape = Shape(geom, stroke_width=float(rng.uniform(1, 8)) if stroked else 0.0)
    fill = _paint(rng, cx, cy, r)
    stroke = _paint(rng, cx, cy, r) if stroked else None
    fill_rule = abi.FILL_EVENODD if rng.uniform() < 0.5 else abi.FILL_NONZERO
    return shape, fill, stroke, fill_rule


def mixed_scene(width, height, n_shapes, seed):
    rng = np.random.default_rng(seed)
    sc = Scene(width, height, background=WHITE, filter=Filter(abi.FILTER_TENT, 1.0))
    for i in range(n_shapes):
        shape, fill, stroke, rule = mixed_shape(rng, width, height)
        sc.shapes.append(shape)
        sc.groups.append(ShapeGroup([i], fill=fill, stroke=stroke, fill_rule=rule))
    return sc


def config2(seed=0):
    sc = mixed_scene(512, 512, 512, seed)
    opts = RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts


def prefiltered_scene(width, height, n_strokes, n_polygons, seed, step=20.0, radius=(5.0, 80.0),
                      seg_range=(4, 16)):
    """Stroked open cubic paths (random walk control points, width U[0.5, 4])
    and filled polygons of 3-12 vertices on a circle, drawn in that order."""
    rng = np.random.default_rng(seed)
    W, H = width, height
    sc = Scene(W, H, background=WHITE, filter=Filter(abi.FILTER_BOX, 0.5))
    for i in range(n_strokes):
        n_seg = int(rng.integers(seg_range[0], seg_range[1] + 1))
        start = rng.uniform(0, 1, size=2) * (W, H)
        steps = rng.normal(0.0, step, size=(3 * n_seg, 2))
        pts = np.vstack([start, start + np.cumsum(steps, axis=0)]).astype(np.float32)
        sc.shapes.append(Shape(Path(np.full(n_seg, 2, np.uint8), pts, is_closed=False),
                               stroke_width=float(rng.uniform(0.5, 4.0))))
        sc.groups.append(ShapeGroup([i], fill=None, stroke=Solid(_rgba(rng))))
    for j in range(n_polygons):
        n = int(rng.integers(3, 13))
        c = rng.uniform(0, 1, size=2) * (W, H)
        rad = rng.uniform(*radius)
        ang = np.sort(rng.uniform(0, 2 * np.pi, size=n))
        pts = (c + rad * np.stack([np.cos(ang), np.sin(ang)], axis=1)).astype(np.float32)
        idx = n_strokes + j
        sc.shapes.append(Shape(Path(np.zeros(n, np.uint8), pts, is_closed=True)))
        sc.groups.append(ShapeGroup([idx], fill=Solid(_rgba(rng))))
    opts = RenderOptions(samples_x=1, samples_y=1, seed=0, jitter=False, use_prefiltering=True)
    return sc, opts


def config3(seed=0):
    return prefiltered_scene(1024, 1024, 2000, 500, seed)


def config4(seed=0, n_scenes=64):
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        n = int(np.floor(np.exp(rng.uniform(np.log(16), np.log(4096)))))
        res = int(rng.choice([128, 256, 512, 1024]))
        sc = mixed_scene(res, res, n, seed=s)
        out.append((sc, RenderOptions(samples_x=2, samples_y=2, seed=0, jitter=True, use_prefiltering=False)))
    return out


def jittered(scene: Scene, sigma=2.0, seed=1) -> Scene:
    """Copy of `scene` with every path point moved by N(0, sigma) px."""
    rng = np.random.default_rng(seed)
    out = copy.deepcopy(scene)
    for sh in out.shapes:
        if isinstance(sh.geometry, Path):
            sh.geometry.points = (sh.geometry.points
                                  + rng.normal(0.0, sigma, size=sh.geometry.points.shape)).astype(np.float32)
    return out


def small_cubic_scene(width=32, height=32, n_paths=4, seed=0, spp=2):
    """A tiny scene for oracle-speed parity tests."""
    sc = cubic_paths_scene(width, height, n_paths, seed, spread=width * 0.35)
    opts = RenderOptions(samples_x=spp, samples_y=spp, seed=0, jitter=True, use_prefiltering=False)
    return sc, opts
