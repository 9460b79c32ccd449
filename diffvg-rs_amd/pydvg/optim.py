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


class DeviceAdamLoop:
    """The optimisation of AdamLoop with the parameters, the Adam state and
    the scene packing resident on the device: a step is device re-packing
    (dvg_device_scene_update) -> forward -> loss gradient -> backward -> Adam,
    with no host round trip except the optional loss readback.

    graph=True: the first step runs eagerly (it sizes every work buffer of
    the library), the second captures the whole step once in a HIP graph
    (torch.cuda.CUDAGraph; the library's kernels join the capture on the
    capture stream) and every step replays it.  The Adam bias corrections
    b^t are device scalars, so the graph carries no step count.  The loss,
    when asked for, is reduced after the replay (a reduction stays out of
    the graph).

    The renderer's stream is set to torch's current stream, so that the
    library's kernels and the torch arithmetic are ordered on one stream."""

    def __init__(self, renderer, scene, options, target, lr_points=0.5, lr_colors=1e-2, graph=False):
        import torch
        self.torch = torch
        self.r = renderer
        renderer.set_stream(torch.cuda.current_stream().cuda_stream)
        self.prepared = PreparedScene(scene, options)
        self.ds = DeviceScene(renderer, self.prepared)
        desc = self.prepared.desc
        G = desc.num_groups
        self.G = G
        self.points = torch.tensor(np.asarray(desc.path_points, np.float32), device="cuda")
        cols = np.concatenate([np.asarray(desc.fill_color, np.float32).reshape(G, 4),
                               np.asarray(desc.stroke_color, np.float32).reshape(G, 4)], axis=1)
        self.colors = torch.tensor(cols.reshape(-1), device="cuda")
        mask = np.zeros((G, 8), np.float32)
        for gi, g in enumerate(scene.groups):
            if isinstance(g.fill, Solid):
                mask[gi, :4] = 1.0
        self.mask = torch.tensor(mask.reshape(-1), device="cuda")
        self.lr_p, self.lr_c = lr_points, lr_colors
        self.b1, self.b2, self.eps = 0.9, 0.999, 1e-8
        self.t = 0
        # b1^t and b2^t, advanced on the device every step
        self.pw1 = torch.ones((), dtype=torch.float64, device="cuda")
        self.pw2 = torch.ones((), dtype=torch.float64, device="cuda")
        self.m_p = torch.zeros_like(self.points)
        self.v_p = torch.zeros_like(self.points)
        self.m_c = torch.zeros_like(self.colors)
        self.v_c = torch.zeros_like(self.colors)
        self.target = target
        self.img = torch.empty(scene.width * scene.height * 4, dtype=torch.float32, device="cuda")
        self.d_render = torch.empty_like(self.img)
        self.bufs, self.sizes = self.ds.grad_buffers(torch)
        self.use_graph = graph
        self.graph = None

    def _adam(self, x, g, m, v, lr):
        m.mul_(self.b1).add_(g, alpha=1 - self.b1)
        v.mul_(self.b2).addcmul_(g, g, value=1 - self.b2)
        mh = m / (1.0 - self.pw1).float()
        vh = v / (1.0 - self.pw2).float()
        x.sub_(lr * mh / (vh.sqrt() + self.eps))

    def _body(self):
        """One step without the loss: update, forward, 2 (img - target),
        backward, Adam."""
        torch = self.torch
        self.ds.update(self.points, self.colors)
        self.ds.render(self.img)
        torch.sub(self.img, self.target, out=self.d_render)
        self.d_render.mul_(2.0)
        self.ds.render_backward(self.d_render, self.bufs, self.sizes)
        self.pw1.mul_(self.b1)
        self.pw2.mul_(self.b2)
        n = self.points.numel()
        if n:
            self._adam(self.points, self.bufs["d_shape_points"][:n], self.m_p, self.v_p, self.lr_p)
        g_c = self.bufs["d_group_data"][: 16 * self.G].view(self.G, 16)[:, 8:16].reshape(-1) * self.mask
        old = self.colors.clone()
        self._adam(self.colors, g_c, self.m_c, self.v_c, self.lr_c)
        # only solid fill colours move, kept in [0, 1]
        self.colors.copy_(torch.where(self.mask > 0, self.colors.clamp(0.0, 1.0), old))

    def _capture(self):
        torch = self.torch
        cur = torch.cuda.current_stream()
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(cur)
        try:
            with torch.cuda.graph(g, stream=side):
                self.r.set_stream(side.cuda_stream)
                self._body()
        finally:
            self.r.set_stream(cur.cuda_stream)
        cur.wait_stream(side)
        self.graph = g

    def step(self, want_loss=True):
        """One step; the loss L = sum((img - target)^2) of the image the step
        rendered (before its update), as a float or (want_loss=False) a
        device scalar."""
        torch = self.torch
        self.t += 1
        if self.use_graph and self.t >= 2:
            if self.graph is None:
                # the capture records without running: this step's work is
                # its first replay
                self._capture()
            self.graph.replay()
        else:
            self._body()
        loss = (self.d_render * self.d_render).sum() * 0.25
        return float(loss) if want_loss else loss
