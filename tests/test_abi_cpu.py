"""The C-ABI library loads and exports every entry point include/pgt_ik.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "pgt_ik.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pgt_[a-z_]+)\s*\(", src)))


def test_header_declares_the_bound_symbols():
    from pyglasstools_amd import engine

    assert sorted(engine.EXPORTED) == declared_functions()


def test_library_exports_every_declared_symbol():
    from pyglasstools_amd import engine

    lib = ctypes.CDLL(engine.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_no_device_fails_loudly():
    from pyglasstools_amd import engine

    if engine.device_count() > 0:
        return  # only meaningful on a host without a GPU
    import pytest

    with pytest.raises(engine.PGTError) as ex:
        engine.Engine(dim=2, max_rcut=2.0)
    assert "no HIP device" in str(ex.value)
    with pytest.raises(engine.PGTError):
        engine.eval_pair_force([[1.0, 0, 0]], 1.0, 1.0)


def test_package_imports_and_alias():
    import pyglasstools
    import pyglasstools.io
    import pyglasstools.irvingkirkwood.cgfunc
    import pyglasstools_amd

    assert pyglasstools is pyglasstools_amd
    assert pyglasstools.io is pyglasstools_amd.io
    from pyglasstools.irvingkirkwood import _irvingkirkwood
    from pyglasstools.potential import _potential

    assert hasattr(_irvingkirkwood, "IrvingKirkwood") and hasattr(_potential, "PairPotentialPoly12")
