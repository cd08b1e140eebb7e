"""FP64 CPU oracle for the Irving-Kirkwood stress-field path.

TEST INFRASTRUCTURE ONLY.  This package is the checker the HIP engine is
compared against and the CPU baseline bench.py times beside it.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
it; the product package (pyglasstools_amd) never does.

The arithmetic lives in ik_oracle.c (a plain-C restatement of
pyglasstools/irvingkirkwood/IrvingKirkwood.h:74-110 and the functions it
calls; each C function cites the reference lines it follows).  This module is
the ctypes binding.

Parity pins (tests/test_oracle_*.py; the surrogate network's checker oracle/mlp.py
is pinned in tests/test_surrogate_cpu.py):
  gl_nodes        numpy.polynomial.legendre.leggauss and 40-digit mpmath roots
  cg_delta        closed-form values, 2D normalisation integral, support
  bond_fixed      exact polynomial integral on inside segments (GL-n is exact
                  there for n >= 5) and an mpmath 50-digit GL-n sum on
                  crossing segments (tests/golden/bond_vectors.json)
  force_divr      F/r(r=sigma=1)=10.142265153709670 for poly12, F/r = -phi'/r
                  by sympy, phi=phi'=0 at the cutoff, symmetry in (di, dj)
  qag             converges to the mpmath integral of the same integrand
  ik_compute      an independent numpy brute force on tiny systems, a
                  two-particle closed form (weights 1 / 1/2 / 0), rotation,
                  translation and permutation invariance, cell list == brute
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

POT_KINDS = {"poly12": 0, "poly18": 1, "poly10": 2, "poly106": 3, "polylj": 4}
CG_KINDS = {"octic": 0, "rect": 1, "mollifier": 2}


class _Args(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("L", ctypes.c_double * 3),
        ("periodic", ctypes.c_int * 3),
        ("N", ctypes.c_int),
        ("pos", ctypes.c_void_p),
        ("diam", ctypes.c_void_p),
        ("mass", ctypes.c_void_p),
        ("vel", ctypes.c_void_p),
        ("disp", ctypes.c_void_p),
        ("G", ctypes.c_int),
        ("grid", ctypes.c_void_p),
        ("pot_kind", ctypes.c_int),
        ("params", ctypes.c_double * 6),
        ("scaled_rcut", ctypes.c_double),
        ("max_rcut", ctypes.c_double),
        ("cg_kind", ctypes.c_int),
        ("cg_rcut", ctypes.c_double),
        ("quad_mode", ctypes.c_int),
        ("quad_order", ctypes.c_int),
        ("adapt_limit", ctypes.c_int),
        ("adapt_epsabs", ctypes.c_double),
        ("adapt_epsrel", ctypes.c_double),
        ("mode", ctypes.c_int),
        ("brute", ctypes.c_int),
        ("nthreads", ctypes.c_int),
    ]


class _Out(ctypes.Structure):
    _fields_ = [
        ("virial", ctypes.c_void_p),
        ("kinetic", ctypes.c_void_p),
        ("elastic", ctypes.c_void_p),
        ("virial_abs", ctypes.c_void_p),
        ("n_ball", ctypes.c_int64),
        ("n_triples", ctypes.c_int64),
        ("n_nonzero", ctypes.c_int64),
    ]


def build():
    """Compile ik_oracle.c with the Makefile beside it."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def build_native(outdir):
    """Compile ik_oracle.c for the host it runs on (-march=native, like the
    reference's CompilerFlagsSetup.cmake:14) into outdir; returns the library
    path.  Only the CPU-baseline leg of the benchmark uses this build."""
    import subprocess

    out = os.path.join(outdir, "libik_oracle_native.so")
    subprocess.run(["gcc", "-O3", "-march=native", "-ffp-contract=off", "-fopenmp", "-fPIC", "-std=c11", "-shared",
                    "-o", out, os.path.join(_HERE, "ik_oracle.c"), "-lm"], check=True, capture_output=True)
    return out


def use_library(path):
    """Bind the oracle functions to the library at path (e.g. build_native's)."""
    global _LIB
    _LIB = _bind(path)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libik_oracle.so")
        if not os.path.exists(path):
            build()
        _LIB = _bind(path)
    return _LIB


def _bind(path):
    L = ctypes.CDLL(path)
    d, i, p = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    L.oracle_gl_nodes.argtypes = [i, p, p]
    L.oracle_gl_nodes.restype = i
    L.oracle_cg_delta.argtypes = [i, d, d]
    L.oracle_cg_delta.restype = d
    L.oracle_bond_fixed.argtypes = [i, d, i, p, p, p]
    L.oracle_bond_fixed.restype = d
    L.oracle_qag.argtypes = [i, d, p, p, p, d, d, i]
    L.oracle_qag.restype = d
    L.oracle_force_divr.argtypes = [i, p, d, d, d, d]
    L.oracle_force_divr.restype = d
    L.oracle_energy.argtypes = [i, p, d, d, d, d]
    L.oracle_energy.restype = d
    L.oracle_ik_compute.argtypes = [ctypes.POINTER(_Args), ctypes.POINTER(_Out)]
    L.oracle_ik_compute.restype = i
    L.oracle_max_threads.restype = i
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _v3(v):
    a = np.zeros(3, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64).ravel()
    a[: v.size] = v
    return a


def gl_nodes(n):
    """Non-negative Gauss-Legendre abscissae on [-1,1] (ascending) + weights."""
    m = (n + 1) // 2
    x = np.zeros(m)
    w = np.zeros(m)
    if lib().oracle_gl_nodes(n, _ptr(x), _ptr(w)) != 0:
        raise ValueError("bad order %d" % n)
    return x, w


def cg_delta(kind, rcut, dr_sq):
    return lib().oracle_cg_delta(CG_KINDS[kind], rcut, dr_sq)


def bond_fixed(kind, rcut, order, x, ri, rij):
    x, ri, rij = _v3(x), _v3(ri), _v3(rij)
    return lib().oracle_bond_fixed(CG_KINDS[kind], rcut, order, _ptr(x), _ptr(ri), _ptr(rij))


def qag(kind, rcut, x, ri, rij, epsabs, epsrel, limit):
    x, ri, rij = _v3(x), _v3(ri), _v3(rij)
    return lib().oracle_qag(CG_KINDS[kind], rcut, _ptr(x), _ptr(ri), _ptr(rij), epsabs, epsrel, limit)


def _params6(params):
    p = np.zeros(6)
    p[: len(params)] = params
    return p


def force_divr(kind, params, scaled_rcut, rsq, di, dj):
    p = _params6(params)
    return lib().oracle_force_divr(POT_KINDS[kind], _ptr(p), scaled_rcut, rsq, di, dj)


def energy(kind, params, scaled_rcut, rsq, di, dj):
    p = _params6(params)
    return lib().oracle_energy(POT_KINDS[kind], _ptr(p), scaled_rcut, rsq, di, dj)


def max_threads():
    return lib().oracle_max_threads()


def ik_compute(system, grid, *, pot="poly12", params=None, scaled_rcut=1.25, max_rcut=None,
               cg="octic", cg_rcut=2.0, order=5, adaptive=None, fields=("virial",),
               mode=1, brute=False, nthreads=0, with_abs=False):
    """Run the oracle's field computation.

    system: dict with dim, L (3,), periodic (3,), pos (N,3), diam (N,), mass (N,),
            vel (N,3) and optionally disp (N,3).
    grid:   (G,3) float64 grid points.
    adaptive: None for the fixed-order rule, else (limit, epsrel, epsabs).
    Returns dict of (G,D,D) arrays (flat index a + D*b, i.e. [g, b, a]) and counts.
    """
    if params is None:
        params = poly12_params()
    dim = int(system["dim"])
    pos = np.ascontiguousarray(system["pos"], dtype=np.float64).reshape(-1, 3)
    N = pos.shape[0]
    diam = np.ascontiguousarray(system["diam"], dtype=np.float64)
    mass = np.ascontiguousarray(system.get("mass", np.ones(N)), dtype=np.float64)
    vel = np.ascontiguousarray(system.get("vel", np.zeros((N, 3))), dtype=np.float64).reshape(-1, 3)
    disp = system.get("disp")
    disp = None if disp is None else np.ascontiguousarray(disp, dtype=np.float64).reshape(-1, 3)
    grid = np.ascontiguousarray(grid, dtype=np.float64).reshape(-1, 3)
    G = grid.shape[0]
    if max_rcut is None:
        dmax = float(diam.max())
        max_rcut = max(scaled_rcut * dmax, dmax)
    A = _Args()
    A.dim = dim
    A.L[:] = [float(v) for v in system["L"]]
    A.periodic[:] = [int(v) for v in system["periodic"]]
    A.N, A.G = N, G
    A.pos, A.diam, A.mass, A.vel = _ptr(pos), _ptr(diam), _ptr(mass), _ptr(vel)
    A.disp = _ptr(disp)
    A.grid = _ptr(grid)
    A.pot_kind = POT_KINDS[pot]
    A.params[:] = list(_params6(params))
    A.scaled_rcut = scaled_rcut
    A.max_rcut = max_rcut
    A.cg_kind = CG_KINDS[cg]
    A.cg_rcut = cg_rcut
    if adaptive is None:
        A.quad_mode, A.quad_order = 0, order
    else:
        A.quad_mode = 1
        A.adapt_limit, A.adapt_epsrel, A.adapt_epsabs = int(adaptive[0]), adaptive[1], adaptive[2]
    A.mode, A.brute, A.nthreads = mode, int(bool(brute)), nthreads
    D = dim
    outs = {}
    O = _Out()
    for name in ("virial", "kinetic", "elastic"):
        if name in fields:
            outs[name] = np.zeros((G, D, D))
            setattr(O, name, _ptr(outs[name]))
    if with_abs:
        outs["virial_abs"] = np.zeros((G, D, D))
        O.virial_abs = _ptr(outs["virial_abs"])
    rc = lib().oracle_ik_compute(ctypes.byref(A), ctypes.byref(O))
    if rc != 0:
        raise ValueError("oracle_ik_compute failed with code %d" % rc)
    outs["n_ball"] = O.n_ball
    outs["n_triples"] = O.n_triples
    outs["n_nonzero"] = O.n_nonzero
    return outs


def poly12_params(v0=1.0, eps=0.2, rcut=1.25):
    """Coefficients exactly as potential/__init__.py:81-88 builds them."""
    c0 = -28.0 * v0 / rcut ** 12
    c1 = 48.0 * v0 / rcut ** 14
    c2 = -21.0 * v0 / rcut ** 16
    return [v0, eps, c0, c1, c2]

