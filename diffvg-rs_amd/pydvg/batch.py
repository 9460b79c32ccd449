"""Batches of scenes over libdvg's dvg_batch_* entry points: many scenes of
uneven size stepped together (forward, L2 loss gradient, backward), one
persistent launch per phase, optionally replayed from a HIP graph, with
autotuned launch shapes."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .renderer import PreparedScene, Renderer, _check
from .scene import FlatGrads, GRAD_FIELDS

FORWARD, LOSS, BACKWARD = abi.BATCH_FORWARD, abi.BATCH_LOSS, abi.BATCH_BACKWARD
STEP = FORWARD | LOSS | BACKWARD


def _cfg(c):
    return {"sblock_log2": c.sblock_log2, "bnd_log2": c.bnd_log2, "wg_per_cu": c.wg_per_cu,
            "unit_log2": c.unit_log2, "ms": c.ms}


class Batch:
    """`items` is a list of (Scene, RenderOptions); `targets` (optional) a list
    of torch CUDA float32 tensors of W*H*4 target pixels, kept alive here."""

    def __init__(self, renderer: Renderer, items, targets=None):
        self.r = renderer
        self.prepared = [PreparedScene(sc, opts) for sc, opts in items]
        n = len(items)
        self.n = n
        self._targets = list(targets) if targets is not None else None
        packs = (abi.ScenePackC * n)(*[p.pack for p in self.prepared])
        bpacks = (abi.BackwardPackC * n)(*[p.bpack for p in self.prepared])
        params = (abi.RenderParamsC * n)(*[p.params for p in self.prepared])
        tg = (C.c_void_p * n)(*[t.data_ptr() for t in self._targets]) if self._targets is not None else None
        h = C.c_void_p()
        st = renderer.lib.dvg_batch_create(renderer.ctx, n, packs, bpacks, params, tg, C.byref(h))
        _check(renderer.lib, renderer.ctx, st)
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self.r.lib.dvg_batch_destroy(h)
            self.handle = None

    def step(self, flags=STEP):
        _check(self.r.lib, self.r.ctx, self.r.lib.dvg_batch_step(self.r.ctx, self.handle, flags))

    def set_graph(self, enable: bool):
        self.r.lib.dvg_batch_set_graph(self.handle, int(enable))

    def config(self):
        c = abi.BatchConfigC()
        self.r.lib.dvg_batch_get_config(self.handle, C.byref(c))
        return _cfg(c)

    def set_config(self, sblock_log2=2, bnd_log2=2, wg_per_cu=4, unit_log2=8):
        c = abi.BatchConfigC(sblock_log2, bnd_log2, wg_per_cu, unit_log2, 0.0)
        _check(self.r.lib, self.r.ctx, self.r.lib.dvg_batch_set_config(self.r.ctx, self.handle, C.byref(c)))

    def autotune(self, flags=STEP, reps=2):
        c = abi.BatchConfigC()
        st = self.r.lib.dvg_batch_autotune(self.r.ctx, self.handle, flags, reps, C.byref(c))
        _check(self.r.lib, self.r.ctx, st)
        return _cfg(c)

    def read(self, i):
        """Scene i's image (numpy W*H*4) and gradients (FlatGrads), synchronously."""
        p = self.prepared[i]
        img = np.zeros(p.params.width * p.params.height * 4, np.float32)
        g = FlatGrads(p.desc)
        st = self.r.lib.dvg_batch_read(self.r.ctx, self.handle, i, img.ctypes.data_as(C.c_void_p), g.ref())
        _check(self.r.lib, self.r.ctx, st)
        return img, g
