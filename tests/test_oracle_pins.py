"""Pins of the FP64 CPU oracle against what the mathematics fixes.

The reference ships no tests or golden outputs for this path, so each oracle
function is checked against an independent statement of its value: library
or high-precision nodes, closed forms, exact integrals, derivatives taken by
sympy, known values, and brute force on tiny systems.  A test that plants an
error in the oracle's arithmetic fails at least one of these.
"""
import json
import math
import os

import mpmath as mp
import numpy as np
import pytest
import sympy as sp

import oracle
from conftest import GOLDEN

RC = 2.0


# ---------------------------------------------------------------- quadrature
@pytest.mark.parametrize("n", list(range(1, 41)) + [64])
def test_gl_nodes_match_numpy(n):
    x, w = oracle.gl_nodes(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    m = (n + 1) // 2
    # numpy returns ascending nodes on [-1, 1]; take the non-negative half
    xr, wr = xr[n - m:], wr[n - m:]
    # numpy's own weights carry ~1e-13 relative error at larger n; the
    # 40-digit check below is the tight one
    np.testing.assert_allclose(x, xr, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, wr, rtol=2e-12, atol=0)


@pytest.mark.parametrize("n", [2, 5, 7, 20, 30, 64])
def test_gl_nodes_match_mpmath_40_digits(n):
    mp.mp.dps = 40
    x, w = oracle.gl_nodes(n)
    for xi, wi in zip(x, w):
        r = mp.findroot(lambda z: mp.legendre(n, z), mp.mpf(xi)) if xi != 0 else mp.mpf(0)
        dp = mp.diff(lambda z: mp.legendre(n, z), r)
        wr = 2 / ((1 - r * r) * dp * dp)
        assert abs(xi - float(r)) <= 1e-16 + 2e-16 * abs(float(r))
        assert abs(wi - float(wr)) <= 4e-16 * float(wr)


def test_gl_rule_integrates_polynomials_exactly():
    # n nodes integrate degree 2n-1 exactly on [0,1]; degree 2n does not
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 8):
        x, w = oracle.gl_nodes(n)
        s = np.concatenate([0.5 - 0.5 * x[::-1], 0.5 + 0.5 * x[(n & 1):]])
        ws = np.concatenate([0.5 * w[::-1], 0.5 * w[(n & 1):]])
        c = rng.normal(size=2 * n)
        exact = sum(c[k] / (k + 1) for k in range(2 * n))
        assert abs(np.polyval(c[::-1], s) @ ws - exact) < 1e-13
        c2 = np.zeros(2 * n + 1)
        c2[-1] = 1.0
        assert abs(np.polyval(c2[::-1], s) @ ws - 1.0 / (2 * n + 1)) > 1e-12


# ---------------------------------------------------------------- kernels
def test_octic_closed_form_values():
    norm = 15.0 / (8 * math.pi * RC ** 2)
    assert oracle.cg_delta("octic", RC, 0.0) == pytest.approx(norm, rel=1e-15)
    assert oracle.cg_delta("octic", RC, RC ** 2) == 0.0
    assert oracle.cg_delta("octic", RC, 2 * RC ** 2) == 0.0
    for r in (0.3, 1.0, 1.7, 1.999):
        t = (r / RC) ** 2
        # 1 - 2 t^2 + t^4 cancels near the edge: absolute error ~ eps * norm
        assert oracle.cg_delta("octic", RC, r * r) == pytest.approx(norm * (1 - t * t) ** 2, rel=1e-13, abs=1e-15 * norm)


@pytest.mark.parametrize("kind", ["octic", "rect", "mollifier"])
def test_kernels_are_normalised_in_2d(kind):
    from scipy.integrate import quad

    val, err = quad(lambda r: 2 * math.pi * r * oracle.cg_delta(kind, RC, r * r), 0, RC, epsabs=1e-14, limit=200)
    assert val == pytest.approx(1.0, abs=1e-9)


def test_octic_is_c1_at_the_edge():
    h = 1e-6
    d_in = (oracle.cg_delta("octic", RC, (RC - h) ** 2) - oracle.cg_delta("octic", RC, (RC - 2 * h) ** 2)) / h
    assert abs(d_in) < 1e-4


def test_mollifier_and_rect_values():
    assert oracle.cg_delta("rect", RC, 1.0) == pytest.approx(1 / (math.pi * RC ** 2), rel=1e-15)
    c = 13.46842098743079272180855772032086697297031125109170011102 / (2.0 * math.pi * RC * RC)
    assert oracle.cg_delta("mollifier", RC, 0.0) == pytest.approx(c * math.exp(-1.0), rel=1e-15)


# ---------------------------------------------------------------- bond integral
def _golden():
    with open(os.path.join(GOLDEN, "bond_vectors.json")) as f:
        return json.load(f)


def test_bond_integral_matches_mpmath_golden_vectors():
    d = _golden()
    assert any(v["crosses"] for v in d["vectors"]) and not all(v["crosses"] for v in d["vectors"])
    for v in d["vectors"]:
        for n in d["orders"]:
            ref = float(v["gl"][str(n)])
            got = oracle.bond_fixed("octic", v["rc"], n, v["x"], v["ri"], v["rij"])
            assert got == pytest.approx(ref, rel=5e-14, abs=5e-17), (n, v)


def test_gl5_equals_exact_integral_inside_support():
    d = _golden()
    for v in d["vectors"]:
        if v["crosses"]:
            continue
        got = oracle.bond_fixed("octic", v["rc"], 5, v["x"], v["ri"], v["rij"])
        assert got == pytest.approx(float(v["exact"]), rel=1e-13, abs=1e-17)


def test_gl_converges_slowly_on_crossing_segments():
    # GL-n is the oracle, not the exact integral: on segments that cross the
    # support edge GL5 differs from the integral at the percent level
    d = _golden()
    rel = [abs(float(v["gl"]["5"]) / float(v["exact"]) - 1) for v in d["vectors"]
           if v["crosses"] and float(v["exact"]) > 1e-6]
    assert np.median(rel) > 1e-3


def test_bond_integral_exact_polynomial_inside():
    # exact integral of (1 - t(s)^2)^2, t quadratic in s, by numpy polynomial algebra
    rng = np.random.default_rng(5)
    P = np.polynomial.Polynomial
    for _ in range(200):
        x = rng.uniform(-1, 1, 3)
        ri = x + rng.uniform(-0.6, 0.6, 3)
        rij = rng.uniform(-0.6, 0.6, 3)
        a = ri - x
        t = P([a @ a, 2 * a @ rij, rij @ rij]) / RC ** 2
        poly = (1 - t * t) ** 2
        exact = 15 / (8 * math.pi * RC ** 2) * (poly.integ()(1.0) - poly.integ()(0.0))
        assert oracle.bond_fixed("octic", RC, 5, x, ri, rij) == pytest.approx(exact, rel=1e-13)


def test_bond_integral_symmetric_under_reversal():
    # s -> 1-s maps the rule onto itself: bond(x; ri -> rj) == bond(x; rj -> ri)
    rng = np.random.default_rng(6)
    for _ in range(100):
        x = rng.uniform(-1, 1, 3)
        ri = x + rng.normal(0, 1.2, 3)
        rij = rng.normal(0, 1.0, 3)
        a = oracle.bond_fixed("octic", RC, 5, x, ri, rij)
        b = oracle.bond_fixed("octic", RC, 5, x, ri + rij, -rij)
        assert a == pytest.approx(b, rel=1e-13, abs=1e-17)


def test_adaptive_qag_converges_to_the_integral():
    d = _golden()
    for v in d["vectors"][:30]:
        got = oracle.qag("octic", v["rc"], v["x"], v["ri"], v["rij"], 1e-12, 1e-12, 200)
        assert got == pytest.approx(float(v["exact"]), rel=1e-10, abs=1e-14)


# ---------------------------------------------------------------- pair forces
FAMILY = {
    # name: (m, n terms of the repulsion, coefficient builder, scaled rcut, eps)
    "poly12": (lambda v0, rc: [v0, 0.2, -28.0 * v0 / rc ** 12, 48.0 * v0 / rc ** 14, -21.0 * v0 / rc ** 16], 1.25),
    "poly18": (lambda v0, rc: [v0, 0.0, -55.0 * v0 / rc ** 18, 99.0 * v0 / rc ** 20, -45.0 * v0 / rc ** 22], 1.25),
    "poly10": (lambda v0, rc: [v0, 0.0416667, -56.0 * v0 / rc ** 10, 140.0 * v0 / rc ** 12, -120.0 * v0 / rc ** 14,
                               35.0 * v0 / rc ** 16], 1.48),
    "poly106": (lambda v0, rc: [v0, 0.1, (-21 + 10 * rc ** 4) * v0 / rc ** 10, -((5 * (-7 + 3 * rc ** 4) * v0) / rc ** 12),
                                (3 * (-5 + 2 * rc ** 4) * v0) / rc ** 14], 3.0),
    "polylj": (lambda v0, rc: [v0, 0.2, -28.0 * v0 / rc ** 12 + 10.0 / rc ** 6, 48.0 * v0 / rc ** 14 - 15.0 / rc ** 8,
                               -21.0 * v0 / rc ** 16 + 6.0 / rc ** 10], 2.5),
}


def _sympy_energy(kind, p):
    r, s = sp.symbols("r s", positive=True)
    x = (r / s) ** 2
    v0, c0, c1, c2 = [sp.Float(v, 30) for v in (p[0], p[2], p[3], p[4])]
    rep = {"poly12": v0 * (s / r) ** 12, "poly18": v0 * (s / r) ** 18, "poly10": v0 * (s / r) ** 10,
           "poly106": v0 * ((s / r) ** 10 - (s / r) ** 6), "polylj": v0 * ((s / r) ** 12 - (s / r) ** 6)}[kind]
    phi = rep + c0 + c1 * x + c2 * x ** 2
    if kind == "poly10":
        phi += sp.Float(p[5], 30) * x ** 3
    return r, s, phi


def test_poly12_known_value():
    p = oracle.poly12_params()
    assert oracle.force_divr("poly12", p, 1.25, 1.0, 1.0, 1.0) == pytest.approx(10.142265153709670, rel=1e-15)


@pytest.mark.parametrize("kind", sorted(FAMILY))
def test_force_is_minus_dphi_dr_over_r(kind):
    build, rc = FAMILY[kind]
    p = build(1.0, rc)
    r, s, phi = _sympy_energy(kind, p)
    fdr = sp.lambdify((r, s), -sp.diff(phi, r) / r, "mpmath")
    en = sp.lambdify((r, s), phi, "mpmath")
    rng = np.random.default_rng(7)
    mp.mp.dps = 30
    for _ in range(40):
        di, dj = rng.uniform(0.73, 1.62, 2)
        sig = 0.5 * (di + dj) * (1 - p[1] * abs(di - dj))
        rr = rng.uniform(0.7, 0.999 * rc) * sig
        got = oracle.force_divr(kind, p, rc, rr * rr, di, dj)
        want = float(fdr(mp.mpf(rr), mp.mpf(sig)))
        assert got == pytest.approx(want, rel=1e-11, abs=1e-11 * (1 + abs(want)))
        assert oracle.energy(kind, p, rc, rr * rr, di, dj) == pytest.approx(float(en(mp.mpf(rr), mp.mpf(sig))),
                                                                         rel=1e-10, abs=1e-11)


@pytest.mark.parametrize("kind", sorted(FAMILY))
def test_potential_is_smooth_at_cutoff(kind):
    build, rc = FAMILY[kind]
    p = build(1.0, rc)
    r, s, phi = _sympy_energy(kind, p)
    at = {r: rc, s: 1}
    assert abs(float(phi.subs(at))) < 1e-6
    assert abs(float(sp.diff(phi, r).subs(at))) < 1e-6


@pytest.mark.parametrize("kind", sorted(FAMILY))
def test_force_cutoff_and_symmetry(kind):
    build, rc = FAMILY[kind]
    p = build(1.0, rc)
    rng = np.random.default_rng(8)
    for _ in range(50):
        di, dj = rng.uniform(0.73, 1.62, 2)
        rsq = rng.uniform(0.5, 4.0)
        assert oracle.force_divr(kind, p, rc, rsq, di, dj) == oracle.force_divr(kind, p, rc, rsq, dj, di)
        sig = 0.5 * (di + dj) * (1 - p[1] * abs(di - dj))
        cut = rc * rc * sig * sig
        assert oracle.force_divr(kind, p, rc, cut, di, dj) == 0.0
        assert oracle.force_divr(kind, p, rc, cut * 1.0001, di, dj) == 0.0
        assert oracle.force_divr(kind, p, rc, cut * 0.9999, di, dj) != 0.0
