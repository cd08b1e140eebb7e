"""Soundness of the FP32 distance pre-test of the 3D candidate scans, on the CPU.

The 3D gather and bond kernels form |a|^2 from FP32 copies of the
coordinates (difference, minimum image by fmaf(-L, rintf(d / L), d), sum of
squares) and compare it with two FP32 thresholds set by the engine's host
code (fp32_band in ik_band.hpp, exported as pgt_pretest_band and checked
bit for bit against the restatement here): below in32 the candidate is taken as
inside the search radius r, above out32 as outside, and only in between is
the FP64 test run.  The thresholds are r^2 -+ B with
    E = (4M + 2L + 4r) 2^-23,  B = 2 (D E (2.02 r + 2E) + 3 2^-24 1.03 r^2),
M the largest coordinate magnitude (box and grid), rounded outwards to FP32.

This emulates the device arithmetic in numpy (FP32 operations, an fmaf with
one rounding, the FP64 minimum image fma(-L, rint(d / L), d) with one
rounding) on adversarial samples: separations within a few band widths of r,
coordinates up to many box lengths away, boxes from 3 to 400 wide.  Every
pre-test verdict must agree with the FP64 test.  The restatement of B here is
what is checked for soundness: on these samples a band 50 times narrower
gives wrong verdicts (the worst-case bound is about an order of magnitude
wider than the errors the samples reach).
"""
from fractions import Fraction

import numpy as np
import pytest

F32 = np.float32


def thresholds(r, M, Lmax, D=3):
    E = (4 * M + 2 * Lmax + 4 * r) * 2.0 ** -23
    B = 2 * (D * E * (2.02 * r + 2 * E) + 3 * 2.0 ** -24 * 1.03 * r * r)
    lo, hi = F32(r * r - B), F32(r * r + B)
    if float(lo) > r * r - B:
        lo = np.nextafter(lo, F32(-np.inf))
    if float(hi) < r * r + B:
        hi = np.nextafter(hi, F32(np.inf))
    return lo, hi, B


def fmaf_neg(L32, k32, d32):
    """fmaf(-L, k, d) with one rounding: the product of two FP32 values and
    the sum are exact in FP64 here (|k| < 2^10), then rounded once."""
    return F32(-float(L32) * float(k32) + float(d32))


def fma64_neg(L, k, d):
    """fma(-L, k, d) for doubles with one rounding (exact rational, then
    rounded to the nearest double)."""
    return float(Fraction(d) - Fraction(L) * Fraction(int(k)))


def a2_fp32(q, x, L):
    s = F32(0)
    for a in range(3):
        d = F32(q[a]) - F32(x[a])
        L32 = F32(L[a])
        inv = F32(1.0 / L[a])
        d = fmaf_neg(L32, np.rint(F32(d * inv)), d)
        s = F32(s + F32(d * d))
    return s


def a2_fp64(q, x, L):
    s = 0.0
    for a in range(3):
        d = q[a] - x[a]
        d = fma64_neg(L[a], np.rint(d * (1.0 / L[a])), d)
        s += d * d
    return s


@pytest.mark.parametrize("Lbox,far", [(3.2, 0.0), (14.4, 0.0), (50.8, 0.0), (256.0, 0.0), (400.0, 0.0),
                                      (14.4, 8.0), (50.8, 3.0)])
def test_fp32_pretest_never_disagrees_with_fp64(Lbox, far):
    rng = np.random.default_rng(int(Lbox * 10 + far))
    r = 2.0 if Lbox > 5 else 1.2
    L = np.array([Lbox, Lbox * 1.1, Lbox * 0.9])
    lo = -L / 2
    # grid points up to `far` box lengths outside; particles wrapped into the box
    M = max(np.abs(lo).max(), np.abs(lo + L).max()) * (1 + 2 * far) + 1e-9
    in32, out32, B = thresholds(r, M, L.max())
    assert 0 < in32 < r * r < out32
    n = 4000
    nin = nout = nband = 0
    for i in range(n):
        x = rng.uniform(-M, M, 3) if far else rng.uniform(lo, lo + L)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        # separations within a few band widths of r (in |a|^2), some exactly at r
        rel = rng.choice([0.0, 1.0]) * rng.uniform(-4, 4) * B / (r * r)
        q = x + r * np.sqrt(1 + rel) * u
        q = q - L * np.floor((q - lo) / L)  # wrapped like the cell list does
        s32 = a2_fp32(q, x, L)
        s64 = a2_fp64(q, x, L)
        if s32 < in32:
            nin += 1
            assert s64 < r * r, (Lbox, far, i, float(s32), s64)
        elif s32 > out32:
            nout += 1
            assert s64 >= r * r, (Lbox, far, i, float(s32), s64)
        else:
            nband += 1
    # the samples straddle the band: all three verdicts occur
    assert nin > 0 and nout > 0 and nband > 0


def test_band_is_narrow_at_the_benchmark_sizes():
    """At the 3D benchmark box (L = 50.8, grid inside the box) the band is
    +-1.5e-4 of r_cg^2 (+-8e-5 of r_cg), so a candidate falls in it about once
    in several thousand and is settled in FP32 otherwise."""
    L = 50.8
    in32, out32, B = thresholds(2.0, L / 2, L)
    assert B / 4.0 < 2e-4
    assert float(out32) - float(in32) < 2e-3


def _restated_band(r, M, Lmax, Lmin, D, ok):
    """The thresholds with the engine's gating: (-1, inf) without a bound."""
    lo, hi, B = thresholds(r, M, Lmax, D)
    E = (4 * M + 2 * Lmax + 4 * r) * 2.0 ** -23
    if not (ok and r * r - B > 0 and 2 * r + 8 * E < Lmin):
        return F32(-1.0), F32(np.inf)
    return lo, hi


def test_engine_band_is_the_restated_band_bit_for_bit():
    """The thresholds the engine's host code sets (pgt_pretest_band, the
    function make_geom calls) equal the restatement whose soundness the test
    above checks, bit for bit, over radii, coordinate ranges, box shapes and
    both dimensions, including the cases without a bound (an open axis, a
    radius near half the box, a band reaching zero)."""
    from pyglasstools_amd import engine

    rng = np.random.default_rng(11)
    cases = [(2.0, 25.4, 50.8, 50.8, 3, True), (2.0, 25.4, 50.8, 50.8, 2, True), (1.2, 1.6, 3.2, 3.2, 3, True),
             (2.0, 25.4, 50.8, 50.8, 3, False), (1.6, 1.7, 3.3, 3.2, 3, True), (1e-3, 1e6, 2e6, 2e6, 3, True)]
    for _ in range(300):
        L = rng.uniform(3.0, 500.0, 3)
        far = rng.choice([0.0, rng.uniform(0.0, 10.0)])
        M = L.max() / 2 * (1 + 2 * far)
        cases.append((rng.uniform(0.05, 0.45) * L.min(), M, L.max(), L.min(), int(rng.choice([2, 3])), True))
    for r, M, Lmax, Lmin, D, ok in cases:
        got = engine.pretest_band(r, M, Lmax, Lmin, D, ok)
        want = _restated_band(r, M, Lmax, Lmin, D, ok)
        assert np.float32(got[0]).tobytes() == np.float32(want[0]).tobytes(), (r, M, Lmax, Lmin, D, ok, got, want)
        assert np.float32(got[1]).tobytes() == np.float32(want[1]).tobytes(), (r, M, Lmax, Lmin, D, ok, got, want)
    # the gating cases do occur among the fixed ones
    assert engine.pretest_band(2.0, 25.4, 50.8, 50.8, 3, False)[0] == -1.0
    assert engine.pretest_band(1.6, 1.7, 3.3, 3.2, 3, True)[0] == -1.0
