"""Adam optimisation loop over path points and fill colours.

The reference has no optimiser; this is the harness of the optimisation
workloads: Adam (Kingma & Ba 2015; beta1 0.9, beta2 0.999, eps 1e-8) on the
path control points and the solid fill colours, L2 loss
L = sum((image - target)^2), so d_render = 2 (image - target).

Every step repacks the scene on the host (dvg_prepare), uploads it, renders,
runs the backward pass on the device and updates the parameters.
"""
from __future__ import annotations

import numpy as np

from .renderer import DeviceScene, PreparedScene
from .scene import FlatGrads, Path, Solid


class Adam:
    def __init__(self, shape, lr, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(shape, np.float32)
        self.v = np.zeros(shape, np.float32)
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.t = 0

    def update(self, x, g):
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        return (x - self.lr * mh / (np.sqrt(vh) + self.eps)).astype(np.float32)


class AdamLoop:
    """Optimise a scene of solid-filled paths towards a device target image."""

    def __init__(self, renderer, scene, options, target, lr_points=0.5, lr_colors=1e-2):
        import torch
        self.torch = torch
        self.r = renderer
        self.scene = scene
        self.options = options
        self.target = target
        self.paths = [(i, s.geometry) for i, s in enumerate(scene.shapes) if isinstance(s.geometry, Path)]
        self.solid = [gi for gi, g in enumerate(scene.groups) if isinstance(g.fill, Solid)]
        pts = np.concatenate([p.points.reshape(-1) for _, p in self.paths]) if self.paths else np.zeros(0, np.float32)
        cols = np.array([scene.groups[gi].fill.color for gi in self.solid], np.float32).reshape(-1)
        self.opt_p = Adam(pts.shape, lr_points)
        self.opt_c = Adam(cols.shape, lr_colors)
        self.img = torch.empty(scene.width * scene.height * 4, dtype=torch.float32, device="cuda")
        self.d_render = torch.empty_like(self.img)

    def _params(self):
        pts = np.concatenate([p.points.reshape(-1) for _, p in self.paths])
        cols = np.array([self.scene.groups[gi].fill.color for gi in self.solid], np.float32).reshape(-1)
        return pts, cols

    def _set(self, pts, cols):
        o = 0
        for _, p in self.paths:
            n = p.points.size
            p.points = pts[o:o + n].reshape(-1, 2)
            o += n
        cols = np.clip(cols, 0.0, 1.0).reshape(-1, 4)
        for k, gi in enumerate(self.solid):
            self.scene.groups[gi].fill = Solid(tuple(float(c) for c in cols[k]))

    def step(self) -> float:
        torch = self.torch
        prep = PreparedScene(self.scene, self.options)
        ds = DeviceScene(self.r, prep)
        ds.render(self.img)
        torch.sub(self.img, self.target, out=self.d_render)
        loss = float((self.d_render * self.d_render).sum())
        self.d_render.mul_(2.0)
        bufs, sizes = ds.grad_buffers(torch)
        ds.render_backward(self.d_render, bufs, sizes)
        d_pts = bufs["d_shape_points"][: sizes["d_shape_points"]].cpu().numpy()
        d_grp = bufs["d_group_data"][: sizes["d_group_data"]].cpu().numpy().reshape(-1, 16)
        pts, cols = self._params()
        # d_shape_points is laid out like the flat path_points (scene order),
        # which is the concatenation order of self.paths.
        g_cols = np.concatenate([d_grp[gi, 8:12] for gi in self.solid]) if self.solid else np.zeros(0, np.float32)
        self._set(self.opt_p.update(pts, d_pts), self.opt_c.update(cols, g_cols))
        return loss
