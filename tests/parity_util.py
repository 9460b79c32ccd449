"""Comparison helpers for the GPU-vs-oracle parity tests.

Gradient parity is judged per field: every non-empty gradient array must be
within the relative L2 bound on its own, so that a field with a small norm
cannot hide behind a large one (d_translation holds W*H*2 entries and would
dominate a norm over the concatenated vector).  A field whose reference norm
is about zero is compared against a floor of FLOOR times the norm of the
whole reference vector instead of its own norm.
"""
import numpy as np

from pydvg.grad import DLinearGradient, DSolid
from pydvg.scene import GRAD_FIELDS

FLOOR = 1e-6


def image_report(img, ref, tol=1e-5):
    """Fraction of pixels whose every channel is within `tol`, and max error."""
    a = np.asarray(img, np.float32).reshape(-1, 4)
    b = np.asarray(ref, np.float32).reshape(-1, 4)
    bad = np.abs(a - b).max(axis=1) > tol
    return float(bad.mean()), float(np.abs(a - b).max())


def grad_report(g, gref, fields=GRAD_FIELDS):
    """Relative L2 error of the whole gradient vector and of each field.

    Each field's error is |a - b| / max(|b|, FLOOR * |b_total|); empty fields
    are skipped."""
    fields = [k for k in fields if np.asarray(gref[k]).size > 0 or np.asarray(g[k]).size > 0]
    a = np.concatenate([np.asarray(g[k], np.float64).ravel() for k in fields]) if fields else np.zeros(0)
    b = np.concatenate([np.asarray(gref[k], np.float64).ravel() for k in fields]) if fields else np.zeros(0)
    nb_total = float(np.linalg.norm(b))
    out = {}
    for k in fields:
        ak, bk = np.asarray(g[k], np.float64).ravel(), np.asarray(gref[k], np.float64).ravel()
        assert ak.shape == bk.shape, (k, ak.shape, bk.shape)
        den = max(float(np.linalg.norm(bk)), FLOOR * nb_total, 1e-30)
        out[k] = float(np.linalg.norm(ak - bk) / den)
    total = float(np.linalg.norm(a - b) / max(nb_total, 1e-30))
    return total, out


def assert_grads(g, gref, tol, fields=GRAD_FIELDS, label=""):
    """Assert the whole vector and every field within `tol` relative L2."""
    total, per = grad_report(g, gref, fields)
    bad = {k: v for k, v in per.items() if not v <= tol}
    print(f"{label} gradient rel L2 {total:.3g}; per field "
          + ", ".join(f"{k} {v:.3g}" for k, v in per.items()), flush=True)
    assert total <= tol and not bad, f"{label}: fields beyond {tol}: {bad} (total {total:.3g})"
    return total, per


def tree_leaves(t):
    """Every leaf of a SceneGrad as one f64 vector, in declaration order (per
    shape the geometry's fields, transform, stroke width; per group the
    shape_to_canvas, fill, stroke; then filter radius, background, background
    image, translation)."""
    out = []
    for s in t.shapes:
        g = s.geometry
        for v in vars(g).values():
            if v is not None:
                out.append(np.asarray(v, np.float64).ravel())
        out += [np.asarray(s.transform, np.float64).ravel(), np.array([s.stroke_width])]
    for g in t.shape_groups:
        out.append(np.asarray(g.shape_to_canvas, np.float64).ravel())
        for p in (g.fill, g.stroke):
            if p is None:
                continue
            if isinstance(p, DSolid):
                out.append(np.array(p.color, np.float64))
            else:
                for v in (p.start, p.end) if isinstance(p, DLinearGradient) else (p.center, p.radius):
                    out.append(np.asarray(v, np.float64))
                for st in p.stops:
                    out.append(np.array([st.offset, *st.color], np.float64))
    out += [np.array([t.filter.radius]), np.array(t.background, np.float64)]
    out += [np.asarray(a, np.float64).ravel() for a in (t.background_image, t.translation) if a is not None]
    return np.concatenate(out)
