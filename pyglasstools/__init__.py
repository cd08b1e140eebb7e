"""Import alias: `import pyglasstools` gives the MI355X engine package
`pyglasstools_amd`, so scripts written for pyglasstools run unchanged.

Every module of pyglasstools_amd is registered under the matching
`pyglasstools.*` name (the same module objects, so global state such as the
solver and logger lists is shared).
"""
import importlib as _importlib
import sys as _sys

_real = _importlib.import_module("pyglasstools_amd")
for _name, _mod in list(_sys.modules.items()):
    if _name == "pyglasstools_amd" or _name.startswith("pyglasstools_amd."):
        _sys.modules["pyglasstools" + _name[len("pyglasstools_amd"):]] = _mod
_sys.modules[__name__] = _real
