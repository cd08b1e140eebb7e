R"""Coarse-graining functions (mirrors pyglasstools/irvingkirkwood/cgfunc/__init__.py).

mode="fixedpoint" takes `order` (Gauss-Legendre nodes of the bond integral);
mode="adaptive" takes `maxiter`, `relerr`, `abserr` (QAG, 15-point
Gauss-Kronrod).  The kernels are evaluated on the GPU.
"""
from ... import _native

_cgfunc = _native._cgfunc


class basecgfunc(object):
    def _getCGFunc(self):
        return self.cgfunc

    def set_rcut(self, rcut):
        self.cgfunc.setRcut(rcut)

    def get_rcut(self):
        return self.cgfunc.getRcut()

    def get_deltafunc(self, x, ri):
        return self.cgfunc.getDeltaFunc(x, ri)

    def get_objfunc(self, s, x, ri, rij):
        return self.cgfunc.getObjFunc(s, x, ri, rij)

    def get_bondfunc(self, x, ri, rij):
        return self.cgfunc.getBondFunc(x, ri, rij)


def _make(fixed_cls, adaptive_cls, rcut, mode, kwargs):
    # keyword lookup as in the reference (cgfunc/__init__.py:31-44): a missing
    # keyword is a NameError there; here it is a TypeError naming the keyword
    if mode == "fixedpoint":
        if "order" not in kwargs:
            raise TypeError("mode='fixedpoint' needs order=")
        return fixed_cls(kwargs["order"], rcut)
    elif mode == "adaptive":
        for k in ("relerr", "abserr", "maxiter"):
            if k not in kwargs:
                raise TypeError("mode='adaptive' needs %s=" % k)
        return adaptive_cls(kwargs["maxiter"], kwargs["relerr"], kwargs["abserr"], rcut)
    raise ValueError("mode must be 'fixedpoint' or 'adaptive'")


class octic(basecgfunc):
    R"""Octic kernel 15/(8 pi rc^2) (1 - (r/rc)^4)^2, zero beyond rc."""

    def __init__(self, rcut=2.5, mode=None, **kwargs):
        self.cgfunc = _make(_cgfunc.CGFuncOcticFixedPoint, _cgfunc.CGFuncOcticAdaptive, rcut, mode, kwargs)


class rect(basecgfunc):
    R"""Top-hat kernel 1/(pi rc^2) inside rc."""

    def __init__(self, rcut=2.5, mode=None, **kwargs):
        self.cgfunc = _make(_cgfunc.CGFuncRectFixedPoint, _cgfunc.CGFuncRectAdaptive, rcut, mode, kwargs)


class mollifier(basecgfunc):
    R"""Bump kernel C exp(-1/(1 - (r/rc)^2)) inside rc."""

    def __init__(self, rcut=2.5, mode=None, **kwargs):
        self.cgfunc = _make(_cgfunc.CGFuncMollifierFixedPoint, _cgfunc.CGFuncMollifierAdaptive, rcut, mode, kwargs)
