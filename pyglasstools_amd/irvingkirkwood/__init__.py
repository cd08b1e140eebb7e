R"""Irving-Kirkwood coarse-grained fields (mirrors pyglasstools/irvingkirkwood/__init__.py).

`ikcalculator` builds the reference's cartesian grid and `radialcalculator`
its polar grid; both run IrvingKirkwood.compute, which here is the GPU engine.
With one rank the scatter/bcast of the reference (__init__.py:53-63) are the
identity.
"""
import numpy as np

from .. import _native
from .. import comm, rank, size, solvers_list, get_sysdata, get_potential

_irvingkirkwood = _native._irvingkirkwood
_pyglasstools = _native._pyglasstools


def initialize_field(names, dim, gridpoints):
    """Field objects for the requested observable names (irvingkirkwood/__init__.py:8-32)."""
    list_obs = {}
    if any("virialstress" in s for s in names):
        if any("elastic" in s for s in names):
            if dim == 2:
                list_obs["elasticvirialstress"] = _irvingkirkwood.ElasticVirialStressField2D(
                    "elasticvirialstress", "2-TENSOR", False, len(gridpoints))
                list_obs["elasticvirialstress"].addGridpoints(gridpoints)
            elif dim == 3:
                list_obs["elasticvirialstress"] = _irvingkirkwood.ElasticVirialStressField3D(
                    "elasticvirialstress", "2-TENSOR", False, len(gridpoints))
                list_obs["elasticvirialstress"].addGridpoints(gridpoints)
        if all("elastic" in s for s in names) is False:
            if dim == 2:
                list_obs["virialstress"] = _irvingkirkwood.VirialStressField2D(
                    "virialstress", "2-TENSOR", False, len(gridpoints))
                list_obs["virialstress"].addGridpoints(gridpoints)
            elif dim == 3:
                list_obs["virialstress"] = _irvingkirkwood.VirialStressField3D(
                    "virialstress", "2-TENSOR", False, len(gridpoints))
                list_obs["virialstress"].addGridpoints(gridpoints)
    if any("kineticstress" in s for s in names):
        if dim == 2:
            list_obs["kineticstress"] = _irvingkirkwood.KineticStressField2D(
                "kineticstress", "2-TENSOR", True, len(gridpoints))
            list_obs["kineticstress"].addGridpoints(gridpoints)
        elif dim == 3:
            list_obs["kineticstress"] = _irvingkirkwood.KineticStressField3D(
                "kineticstress", "2-TENSOR", True, len(gridpoints))
            list_obs["kineticstress"].addGridpoints(gridpoints)
    return list_obs


def cartesian_grid(boxsize, dx):
    """The reference grid: nmax = int(L/dx) points per axis from -L/2 to +L/2
    (both ends included), x outermost, z = 0 (irvingkirkwood/__init__.py:44-51)."""
    nmax = int(boxsize / dx)
    points = np.linspace(-boxsize / 2.0, +boxsize / 2.0, nmax)
    n = len(points)
    grid = np.zeros((n * n, 3), dtype=np.float64)
    grid[:, 0] = np.repeat(points, n)
    grid[:, 1] = np.tile(points, n)
    return grid


class ikcalculator(object):
    R"""IK fields on a cartesian grid of spacing ~dx (irvingkirkwood/__init__.py:34-77)."""

    def __init__(self, cgfunc, dx):
        self.cgfunc = cgfunc
        self.manager = _pyglasstools.Manager()
        self.coordmode = "cartesian"
        globalsize = 0
        if rank == 0:
            gridpoints = cartesian_grid(get_sysdata().pysimbox.boxsize[0], dx)
            globalsize = len(gridpoints)
            self.gridpoints = np.array_split(gridpoints, size)
        else:
            self.gridpoints = np.array_split([np.zeros(size)], size)
        self.gridpoints = comm.scatter_v(self.gridpoints, 0)
        self.gridpoints = np.reshape(self.gridpoints, (len(self.gridpoints), 3))
        self.gridsize = comm.bcast(globalsize, 0)
        self.calculator = _irvingkirkwood.IrvingKirkwood(get_sysdata().cppparticledata,
                                                         get_potential().cpppairpotential,
                                                         cgfunc._getCGFunc(), comm)
        solvers_list.append(self)

    def add_observables(self, observables):
        for name in observables:
            self.calculator.addObservable(observables[name])

    def run(self):
        self.calculator.compute(self.gridpoints)

    def update(self, frame_num):
        from .. import update_sysdata
        update_sysdata(frame_num)
        self.calculator.setSystemData(get_sysdata().cppparticledata)


def polar_grid(center, spacingtype="normal", rmax=10, rmin=0, dr=0.1, dlnr=0.05):
    """The reference's polar grid (irvingkirkwood/__init__.py:86-124): the centre,
    then rings of radius r with int(2 pi / (dr/r)) points each.
    Returns (gridid, r, theta, gridpoints)."""
    if spacingtype == "normal":
        r = np.linspace(0, rmax, int(rmax / dr))[1:]
    elif spacingtype == "log":
        r = np.exp(np.linspace(np.log(rmin), np.log(rmax), int((np.log(rmax) - np.log(rmin)) / dlnr)))
    else:
        raise ValueError("spacingtype must be 'normal' or 'log'")
    gridid, globalr, globaltheta, gridpoints = [0], [0], [0], [np.asarray(center, dtype=np.float64)]
    for i in range(len(r)):
        dtheta = dr / r[i]
        theta = np.linspace(0, 2 * np.pi, int(2 * np.pi / dtheta))
        x = r[i] * np.cos(theta)
        y = r[i] * np.sin(theta)
        for j in range(len(theta)):
            gridid.append(i + 1)
            globalr.append(r[i])
            globaltheta.append(theta[j])
            gridpoints.append(np.asarray([x[j], y[j], 0], dtype=np.float64) + center)
    return (np.asarray(gridid, dtype=np.float64), np.asarray(globalr, dtype=np.float64),
            np.asarray(globaltheta, dtype=np.float64), np.asarray(gridpoints, dtype=np.float64))


class radialcalculator(object):
    R"""IK fields on a polar grid around `center` (irvingkirkwood/__init__.py:80-166)."""

    def __init__(self, cgfunc, center=np.array([0, 0, 0]), spacingtype="normal", rmax=10, rmin=0, dr=0.1,
                 dlnr=0.05):
        self.cgfunc = cgfunc
        self.manager = _pyglasstools.Manager()
        self.coordmode = "polar"
        self.center = center
        gid, gr, gth, gp = polar_grid(center, spacingtype, rmax, rmin, dr, dlnr)
        globalsize = len(gp)
        self.gridid = np.reshape(comm.scatter_v(np.array_split(gid, size), 0), (-1,))
        self.globalr = np.reshape(comm.scatter_v(np.array_split(gr, size), 0), (-1,))
        self.globaltheta = np.reshape(comm.scatter_v(np.array_split(gth, size), 0), (-1,))
        self.gridpoints = np.reshape(comm.scatter_v(np.array_split(gp, size), 0), (-1, 3))
        self.gridsize = comm.bcast(globalsize, 0)
        self.calculator = _irvingkirkwood.IrvingKirkwood(get_sysdata().cppparticledata,
                                                         get_potential().cpppairpotential,
                                                         cgfunc._getCGFunc(), comm)
        solvers_list.append(self)

    def set_center(self, center):
        self.center = center
        self.gridpoints[0:] += center

    def add_observables(self, observables):
        for name in observables:
            self.calculator.addObservable(observables[name])

    def run(self):
        self.calculator.compute(self.gridpoints)

    def update(self, frame_num):
        from .. import update_sysdata
        update_sysdata(frame_num)
        self.calculator.setSystemData(get_sysdata().cppparticledata)


from . import cgfunc  # noqa: E402,F401
