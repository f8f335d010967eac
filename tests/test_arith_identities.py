"""Floating-point identities the kernels rely on for bit-exact results,
checked on the host by brute force.

The distance-driven FP (csrc/fp_dd.hip) forms the float quotients
tsum / lsum of a row from one FP64 reciprocal: RN_f32(a * RN_f64(1 / b))
must equal the float division RN_f32(a / b) for all float a, b (b > 0).
The argument: the FP64 product is within about 2^-52 (relative) of a / b,
while a quotient of two floats is never an exact float midpoint and stays at
least 2^-49 (relative) away from every midpoint.  The samples below include
quotients placed next to midpoints, where a wrong identity would fail first.
"""
import numpy as np


def _quotients(a, b):
    want = a / b
    got = (a.astype(np.float64) * (1.0 / b.astype(np.float64))).astype(np.float32)
    return want, got


def test_float_division_from_fp64_reciprocal_random():
    rng = np.random.default_rng(20261203)
    n = 2_000_000
    a = ((rng.random(n) * 2 - 1) * np.exp2(rng.integers(-30, 30, n))).astype(np.float32)
    b = (rng.random(n) * np.exp2(rng.integers(-30, 30, n))).astype(np.float32)
    b = np.where(b > 0, b, np.float32(1))
    want, got = _quotients(a, b)
    np.testing.assert_array_equal(got, want)


def test_float_division_from_fp64_reciprocal_near_midpoints():
    rng = np.random.default_rng(20261204)
    n = 2_000_000
    b = (rng.random(n) + 0.5).astype(np.float32)
    m = (rng.random(n) + 0.5).astype(np.float32)
    mid = m.astype(np.float64) + np.spacing(m).astype(np.float64) / 2
    a = (mid * b.astype(np.float64)).astype(np.float32)  # a / b next to a midpoint
    want, got = _quotients(a, b)
    np.testing.assert_array_equal(got, want)


def test_float_division_identity_is_sensitive():
    """A plainly wrong variant (the reciprocal rounded to float first) fails
    on the same samples, so the checks above can tell the difference."""
    rng = np.random.default_rng(20261204)
    n = 200_000
    b = (rng.random(n) + 0.5).astype(np.float32)
    a = (rng.random(n) + 0.5).astype(np.float32)
    wrong = (a * (np.float32(1) / b)).astype(np.float32)
    assert (wrong != a / b).any()


# The single-view BP (csrc/par_svbp.hip, div_rn) divides by the channel
# pitch b with y = RN(1/b): q0 = RN(a y), r = a - b q0 (exact in one FMA),
# q1 = RN(q0 + r y) must be the correctly rounded a / b (Markstein's
# theorem) for every |a| >= 2^-960.  The FMAs are emulated exactly with
# rationals; the samples include exact multiples of b and values one ulp
# away from them, where a wrong correction step would fail first.
def _fma(x, y, z):
    from fractions import Fraction
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def _div_rn(a, b, y):
    q0 = a * y
    return _fma(_fma(-q0, b, a), y, q0)


def test_fp64_division_from_reciprocal_and_two_fmas():
    rng = np.random.default_rng(20261205)
    pitches = [1.0, 1.1, 0.65, 1.3, 0.7, 1.0 / 3.0, 0.1, 2.0, 0.9999999999999999]
    pitches += list((rng.random(20) + 0.5) * np.exp2(rng.integers(-6, 6, 20)))
    for b in pitches:
        y = 1.0 / b
        a = list((rng.random(300) * 2 - 1) * 1000.0)
        k = rng.integers(-100000, 100000, 300).astype(np.float64)
        a += list(k * b)  # exact multiples (rounded products)
        a += list(np.nextafter(k * b, np.inf)) + list(np.nextafter(k * b, -np.inf))
        for x in a:
            if abs(x) < 2.0 ** -960:
                continue
            assert _div_rn(x, b, y) == x / b, (x, b)
