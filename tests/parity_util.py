"""Comparison helpers for the GPU-vs-oracle parity tests."""
import numpy as np

from pydvg.scene import GRAD_FIELDS


def image_report(img, ref, tol=1e-5):
    """Fraction of pixels whose every channel is within `tol`, and max error."""
    a = np.asarray(img, np.float32).reshape(-1, 4)
    b = np.asarray(ref, np.float32).reshape(-1, 4)
    bad = np.abs(a - b).max(axis=1) > tol
    return float(bad.mean()), float(np.abs(a - b).max())


def grad_report(g, gref, fields=GRAD_FIELDS):
    """Relative L2 error of the whole gradient vector and per field."""
    out = {}
    for k in fields:
        a, b = g[k], gref[k]
        if b.size == 0:
            continue
        nb = float(np.linalg.norm(b))
        out[k] = float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))
    a = np.concatenate([g[k] for k in fields])
    b = np.concatenate([gref[k] for k in fields])
    total = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
    return total, out
