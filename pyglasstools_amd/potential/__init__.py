R"""Pair potentials of the polydisperse m-n family (mirrors pyglasstools/potential/__init__.py).

Coefficients c0, c1, c2 (and c3 for poly10) are built exactly as the
reference builds them (potential/__init__.py:78-133), so the smoothed
potential and its first two derivatives vanish at the cutoff.  Lennard-Jones
and Yukawa forms (potential/__init__.py:58-75,135-143) are outside the
polydisperse m-n + even-polynomial family this engine implements.
"""
import numpy as np

from .. import _native
from .. import set_potential

_potential = _native._potential


class pairpotential(object):
    R"""Common methods of a pair potential (potential/__init__.py:8-56)."""

    def __init__(self):
        self.cpppairpotential = None
        self.name = None
        self.dict_params = None

    def get_potentialname(self):
        return self.name

    def set_scaledrcut(self, rcut):
        self.cpppairpotential.scaled_rcut = rcut

    def get_scaledrcut(self):
        return self.cpppairpotential.scaled_rcut

    def set_params(self, val, num):
        if isinstance(num, str):
            self.cpppairpotential.setParams(val, self.dict_params[num])
        elif isinstance(num, int):
            self.cpppairpotential.setParams(val, num)

    def get_params(self, num=None):
        if num is None:
            return self.cpppairpotential.getParams()
        elif isinstance(num, str):
            return self.cpppairpotential.getParams()[self.dict_params[num]]
        elif isinstance(num, int):
            return self.cpppairpotential.getParams()[num]

    def print_params(self):
        params = self.cpppairpotential.getParams()
        for key in self.dict_params:
            print("{}: {} ".format(key, params[self.dict_params[key]]))

    def get_pairforce(self, rij, di, dj):
        """|F| = (F/r) * |r| for the bond vector rij (2 or 3 components)."""
        rij = np.asarray(rij, dtype=np.float64)
        if len(rij) == 2:
            rij = np.append(rij, 0)
        return self.cpppairpotential.getPairForceDivR(rij, di, dj) * np.linalg.norm(rij)

    def get_pairenergy(self, rij, di, dj):
        rij = np.asarray(rij, dtype=np.float64)
        if len(rij) == 2:
            rij = np.append(rij, 0)
        return self.cpppairpotential.getPairEnergy(rij, di, dj)


class lj(pairpotential):
    def __init__(self, eps, vareps, rcut, name="lennard-jones", mode=None):
        raise NotImplementedError("Lennard-Jones is not a polydisperse m-n potential; this engine implements "
                                  "polydisperse12/18/10/106/lj")


class polydisperse12(pairpotential):
    R"""v0 (sigma/r)^12 + c0 + c1 (r/sigma)^2 + c2 (r/sigma)^4 below rcut*sigma_ij."""

    def __init__(self, v0=1.0, eps=0.2, rcut=1.25):
        self.name = "polydisperse-12"
        c0 = -28.0 * v0 / rcut ** 12
        c1 = 48.0 * v0 / rcut ** 14
        c2 = -21.0 * v0 / rcut ** 16
        self.cpppairpotential = _potential.PairPotentialPoly12(rcut, [v0, eps, c0, c1, c2])
        self.dict_params = {"v0": 0, "eps": 1, "c0": 2, "c1": 3, "c2": 4}
        set_potential(self)


class polydisperse18(pairpotential):
    def __init__(self, v0=1.0, eps=0.0, rcut=1.25):
        self.name = "polydisperse-18"
        c0 = -55.0 * v0 / rcut ** 18
        c1 = 99.0 * v0 / rcut ** 20
        c2 = -45.0 * v0 / rcut ** 22
        self.cpppairpotential = _potential.PairPotentialPoly18(rcut, [v0, eps, c0, c1, c2])
        self.dict_params = {"v0": 0, "eps": 1, "c0": 2, "c1": 3, "c2": 4}
        set_potential(self)


class polydisperselj(pairpotential):
    def __init__(self, v0=1.0, eps=0.2, rcut=2.5):
        self.name = "polydisperse-lj"
        c0 = -28.0 * v0 / rcut ** 12 + 10.0 / rcut ** 6
        c1 = 48.0 * v0 / rcut ** 14 - 15.0 / rcut ** 8
        c2 = -21.0 * v0 / rcut ** 16 + 6.0 / rcut ** 10
        self.cpppairpotential = _potential.PairPotentialPolyLJ(rcut, [v0, eps, c0, c1, c2])
        self.dict_params = {"v0": 0, "eps": 1, "c0": 2, "c1": 3, "c2": 4}
        set_potential(self)


class polydisperse10(pairpotential):
    def __init__(self, v0=1.0, eps=0.0416667, rcut=1.48):
        self.name = "polydisperse-10"
        c0 = -(56.0) * v0 / (rcut ** 10)
        c1 = (140.0) * v0 / (rcut ** 12)
        c2 = -(120.0) * v0 / (rcut ** 14)
        c3 = (35.0) * v0 / (rcut ** 16)
        self.cpppairpotential = _potential.PairPotentialPoly10(rcut, [v0, eps, c0, c1, c2, c3])
        self.dict_params = {"v0": 0, "eps": 1, "c0": 2, "c1": 3, "c2": 4, "c3": 5}
        set_potential(self)


class polydisperse106(pairpotential):
    def __init__(self, v0=1.0, eps=0.1, rcut=3.0):
        self.name = "polydisperse-106"
        c0 = (-21 + 10 * rcut ** 4) * v0 / rcut ** 10
        c1 = -((5 * (-7 + 3 * rcut ** 4) * v0) / rcut ** 12)
        c2 = (3 * (-5 + 2 * rcut ** 4) * v0) / rcut ** 14
        self.cpppairpotential = _potential.PairPotentialPoly106(rcut, [v0, eps, c0, c1, c2])
        self.dict_params = {"v0": 0, "eps": 1, "c0": 2, "c1": 3, "c2": 4}
        set_potential(self)


class polydisperseyukawa(pairpotential):
    def __init__(self, v0=10.0, eps=0.0, rcut=4.0, kappa=1.0):
        raise NotImplementedError("the Yukawa form is not a polydisperse m-n potential; this engine implements "
                                  "polydisperse12/18/10/106/lj")
