"""GPU: the whole-image gradient of the 1024^2 / 1000-path workload against
the oracle's, every interior and boundary sample of the image (the fixture
tests/golden/headline_grad.npz, written by tests/golden/make_headline_grad.py;
the oracle needs about four minutes for it on one core).

The adjoint is d_render = 2 (0.5 - T), T the procedural 256-Gaussian target:
like the optimisation loss's 2 (image - T) it weights every edge of the
canvas.  Bound: relative L2 <= 1e-3 for the whole vector and every field, in
the default boundary pass (which checks inside tests near edges and
re-evaluates the uncertain samples in f64) and, reported beside it, in the
unchecked f32 pass, which is not bounded here.
"""
import importlib.util
import os

import numpy as np
import pytest

from parity_util import assert_grads, grad_report
from pydvg import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _fixture():
    spec = importlib.util.spec_from_file_location("make_headline_grad",
                                                  os.path.join(GOLDEN, "make_headline_grad.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    ref = np.load(os.path.join(GOLDEN, "headline_grad.npz"))
    dr = mod.headline_adjoint()
    # the adjoint is recomputed here: it must be the one the fixture used
    assert abs(float(np.linalg.norm(dr.astype(np.float64))) - float(ref["d_render_norm"])) <= 1e-9 * float(
        ref["d_render_norm"])
    return dr, {k: ref[k] for k in ref.files if k != "d_render_norm"}


def test_whole_image_gradient_matches_oracle(renderer):
    sc, opts = scenes.headline()
    dr, ref = _fixture()
    print("case whole-image headline gradient (Gaussian-target adjoint)", flush=True)
    fields = tuple(ref)
    renderer.set_exact_boundary(False)
    try:
        g_fast = renderer.render_backward(sc, opts, dr)
    finally:
        renderer.set_exact_boundary(True)
    fast, per_fast = grad_report({k: g_fast[k] for k in fields}, ref, fields)
    print(f"unchecked f32 boundary pass: rel L2 {fast:.3g}; per field "
          + ", ".join(f"{k} {v:.3g}" for k, v in per_fast.items()), flush=True)
    g = renderer.render_backward(sc, opts, dr)
    assert_grads({k: g[k] for k in fields}, ref, 1e-3, fields, label="default boundary pass")
