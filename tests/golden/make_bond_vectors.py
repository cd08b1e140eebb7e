"""Generate tests/golden/bond_vectors.json: 50-digit reference values of the
coarse-graining bond integral.

For random segments (x, r_i, r_ij) near an Octic kernel of width r_cg this
records, evaluated entirely in mpmath at 50 significant digits:
  gl[n]  the n-point Gauss-Legendre sum on [0,1] of Delta(|x - (r_i + s r_ij)|^2)
         (nodes: roots of P_n found by mpmath.findroot, weights from P_n')
  exact  the integral itself, by mpmath.quad split at the points where the
         segment crosses the kernel support
Segments are drawn so that about half stay inside the support (where GL-n
with n >= 5 is exact) and half cross its edge.

Run:  python tests/golden/make_bond_vectors.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
ORDERS = (1, 2, 3, 5, 8, 13)
RC = mp.mpf(2)


def gl_nodes(n):
    xs, ws = [], []
    guesses = np.polynomial.legendre.leggauss(n)[0]
    for g in guesses:
        r = mp.findroot(lambda z: mp.legendre(n, z), mp.mpf(float(g)))
        dp = mp.diff(lambda z: mp.legendre(n, z), r)
        xs.append(r)
        ws.append(2 / ((1 - r * r) * dp * dp))
    return xs, ws


def octic(dsq):
    rcsq = RC * RC
    if dsq >= rcsq:
        return mp.mpf(0)
    t = dsq / rcsq
    return 15 / (8 * mp.pi * rcsq) * (1 - t * t) ** 2


def integrand(s, x, ri, rij):
    d = [x[a] - (ri[a] + s * rij[a]) for a in range(3)]
    return octic(sum(v * v for v in d))


def crossings(x, ri, rij):
    """s in (0,1) where |x - r_i - s r_ij| = r_cg."""
    a = [ri[k] - x[k] for k in range(3)]
    A = sum(v * v for v in a) - RC * RC
    B = 2 * sum(a[k] * rij[k] for k in range(3))
    C = sum(v * v for v in rij)
    disc = B * B - 4 * A * C
    pts = []
    if disc > 0 and C > 0:
        for sgn in (-1, 1):
            s = (-B + sgn * mp.sqrt(disc)) / (2 * C)
            if 0 < s < 1:
                pts.append(s)
    return sorted(pts)


def main():
    rng = np.random.default_rng(20240601)
    nodes = {n: gl_nodes(n) for n in ORDERS}
    out = []
    for k in range(60):
        x = rng.uniform(-1, 1, 3)
        x[2] = 0.0 if k % 3 else x[2]
        inside = k % 2 == 0
        dim = 3 if k % 3 == 0 else 2
        u = rng.normal(0, 1, 3)
        v = rng.normal(0, 1, 3)
        if dim == 2:
            u[2] = v[2] = 0.0
        u /= np.linalg.norm(u)
        v /= np.linalg.norm(v)
        if inside:  # whole segment well inside the support
            ri = x + u * rng.uniform(0.0, 0.9)
            rij = v * rng.uniform(0.3, 0.9)
        else:  # one endpoint near the edge of the support, bond of typical length
            ri = x + u * rng.uniform(1.5, 2.3)
            rij = v * rng.uniform(0.8, 1.6)
        if dim == 2:
            x[2] = ri[2] = rij[2] = 0.0
        X = [mp.mpf(float(v)) for v in x]
        RI = [mp.mpf(float(v)) for v in ri]
        RIJ = [mp.mpf(float(v)) for v in rij]
        f = lambda s: integrand(s, X, RI, RIJ)
        rec = {"x": [float(v) for v in x], "ri": [float(v) for v in ri], "rij": [float(v) for v in rij],
               "rc": 2.0, "gl": {}}
        for n in ORDERS:
            xs, ws = nodes[n]
            val = sum(w / 2 * f((1 + z) / 2) for z, w in zip(xs, ws))
            rec["gl"][str(n)] = mp.nstr(val, 40)
        pts = [mp.mpf(0)] + crossings(X, RI, RIJ) + [mp.mpf(1)]
        rec["exact"] = mp.nstr(mp.quad(f, pts), 40)
        rec["crosses"] = len(pts) > 2
        out.append(rec)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bond_vectors.json")
    with open(path, "w") as fh:
        json.dump({"kernel": "octic", "orders": list(ORDERS), "digits": 50, "vectors": out}, fh, indent=1)
    print("wrote %d vectors to %s" % (len(out), path))


if __name__ == "__main__":
    main()
