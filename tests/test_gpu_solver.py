"""GPU: the cubic winding solve's fast path against the libm formula.

The walks evaluate the winding of cubic segments through a per-segment record
(coefficients, root mask) and, on well-conditioned segments, fast acos / cos /
cube root.  dvg_selftest_winding draws random segments and points inside
their y-range (the cases the walk solves) and counts the pairs whose winding
differs from the curves.rs formula evaluated with libm.  Measured: 2 in
16.7M, both near-ties at a segment endpoint; the bound is 5e-7 of the pairs.
"""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2])
def test_fast_winding_agrees_with_libm_formula(renderer, seed):
    n = 1 << 24
    m = C.c_uint32(0xFFFFFFFF)
    print(f"case winding self-test seed {seed}", flush=True)
    st = renderer.lib.dvg_selftest_winding(renderer.ctx, n, seed, C.byref(m))
    assert st == 0
    print(f"seed {seed}: {m.value} of {n} pairs differ")
    assert m.value <= n * 5e-7


# dvg_selftest_inside counts, per family of 2^22 pairs, (a) the pairs where
# the checked inside test of the default boundary pass differs from the f64
# formula without having flagged itself uncertain (the pass evaluates flagged
# samples again in f64), (b) the flagged pairs, (c) the pairs where the
# unchecked f32 pass's test differs.  Measured: (a) 0, 2, 0, 1-3
# of 4.19M (family 1: near-quadratic cubics, family 3: points within 0.01 px
# of the curve), bound 4e-6 of a family; (b) a small share of a family except
# next to the curve (1.6 %) and next to a y-extremum (0.15 %).
UNFLAGGED_BOUND = 4e-6
FLAGGED_BOUNDS = (1e-4, 1e-4, 5e-3, 5e-2)
# (c): random pairs, near-quadratic cubics and points near a y-extremum agree
# up to rare ties; points within 0.01 px of the curve differ where f32 and f64
# place the crossing on different sides of the point (the f32 root's error is
# about 1e-6 of the segment's extent).
DEFAULT_BOUNDS = (2e-6, 2e-6, 1e-4, 2e-3)


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_inside_test_follows_f64_winding(renderer, seed):
    """The backward's cubic inside tests (the record solve, with or without
    its error checks; the monotone pieces where Cardano cannot settle the
    count) against the reference CPU's f64 winding
    (distance/winding.rs:166-188), restated on the device."""
    n = 1 << 24
    counts = (C.c_uint32 * 12)()
    print(f"case inside self-test seed {seed}", flush=True)
    st = renderer.lib.dvg_selftest_inside(renderer.ctx, n, seed, counts)
    assert st == 0
    per = n // 4
    print(f"seed {seed}: checked test, unflagged differing pairs per family {list(counts[:4])}, "
          f"flagged {list(counts[4:8])}; default test, differing {list(counts[8:])}")
    for f in range(4):
        assert counts[f] <= per * UNFLAGGED_BOUND, f
        assert counts[4 + f] <= per * FLAGGED_BOUNDS[f], f
        assert counts[8 + f] <= per * DEFAULT_BOUNDS[f], f
