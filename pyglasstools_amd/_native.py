"""Loader for the in-tree native libraries.

The IK path has no CPU implementation in this package: if libpgt_ik.so or the
_pgt extension is missing the import fails loudly, and with no GPU every
compute call raises (the engine refuses to start without a HIP device).
"""
import ctypes
import importlib
import os
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpgt_ik.so")
MODULE_PATH = os.path.join(HERE, "_pgt" + sysconfig.get_config_var("EXT_SUFFIX"))


class NativeLibraryMissing(ImportError):
    pass


def _require(path):
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            "%s is not built; run `python -m pyglasstools_amd.csrc.build` (or __graft_entry__.build())" % path)


_require(LIB_PATH)
_require(MODULE_PATH)
# load the engine first (global), so the extension's rpath lookup is satisfied
_engine_cdll = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
_pgt = importlib.import_module("pyglasstools_amd._pgt")

_pyglasstools = _pgt._pyglasstools
_potential = _pgt._potential
_cgfunc = _pgt._cgfunc
_irvingkirkwood = _pgt._irvingkirkwood

# the reference's extension-module paths, importable as modules
sys.modules.setdefault("pyglasstools_amd._pyglasstools", _pyglasstools)
sys.modules.setdefault("pyglasstools_amd.potential._potential", _potential)
sys.modules.setdefault("pyglasstools_amd.irvingkirkwood._irvingkirkwood", _irvingkirkwood)
sys.modules.setdefault("pyglasstools_amd.irvingkirkwood.cgfunc._cgfunc", _cgfunc)


def device_count():
    return _pgt.device_count()
