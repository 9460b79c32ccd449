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
