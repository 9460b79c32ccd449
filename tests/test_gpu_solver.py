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


# Bounds per family of dvg_selftest_inside (2^22 pairs a family over two
# seeds): random pairs, near-quadratic cubics and points near a y-extremum
# should agree with the f64 formula up to rare ties; points within 0.01 px of
# the curve differ where f32 and f64 place the crossing on different sides of
# the point (the f32 root's error is about 1e-6 of the segment's extent).
INSIDE_BOUNDS = (2e-6, 2e-6, 1e-4, 2e-3)


@pytest.mark.parametrize("seed", [3, 4])
def test_backward_inside_test_follows_f64_winding(renderer, seed):
    """The backward's cubic inside test (the record solve, or the monotone
    pieces on ill-conditioned segments) against the reference CPU's f64
    winding (distance/winding.rs:166-188), restated on the device."""
    n = 1 << 24
    counts = (C.c_uint32 * 8)()
    print(f"case inside self-test seed {seed}", flush=True)
    st = renderer.lib.dvg_selftest_inside(renderer.ctx, n, seed, counts)
    assert st == 0
    per = n // 4
    print(f"seed {seed}: differing pairs per family {list(counts[:4])}, solved on pieces {list(counts[4:])}")
    assert counts[5] > per // 2  # the near-quadratic family takes the piece path
    for f in range(4):
        assert counts[f] <= per * INSIDE_BOUNDS[f], f
