"""Build the in-tree native libraries of pyglasstools_amd.

  libpgt_ik.so  HIP engine + C ABI (hipcc, gfx950)
  _pgt*.so      pybind11 module over the C ABI (g++)

Both land in the package directory (pyglasstools_amd/), which is where the
package loads them from.  Run:  python -m pyglasstools_amd.csrc.build
"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
INCLUDE = os.path.join(ROOT, "include")
ARCH = os.environ.get("PGT_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC",
    "--offload-arch=" + ARCH,
    # keep FP32 arithmetic off the packed-FP32 VALU instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def lib_path():
    return os.path.join(PKG, "libpgt_ik.so")


def module_path():
    return os.path.join(PKG, "_pgt" + sysconfig.get_config_var("EXT_SUFFIX"))


def _run(cmd):
    print("+ " + " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


# translation units of libpgt_ik.so, compiled in parallel: the host engine
# with the cell-list, bond-list and surrogate kernels, and the field gathers
ENGINE_SOURCES = ["ik_gather_3d64.hip", "ik_gather_3d32.hip", "ik_gather_3d2.hip", "ik_gather_2d64.hip",
                  "ik_gather_2d32.hip", "ik_engine.hip", "ik_gather_adaptive.hip", "ik_sweep_gather.hip"]


def build_engine():
    import shutil
    import tempfile
    from concurrent.futures import ThreadPoolExecutor

    out = lib_path()
    tmp = out + ".tmp"
    objdir = tempfile.mkdtemp(prefix="pgt_build_")
    try:
        objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in ENGINE_SOURCES]
        jobs = [["hipcc"] + HIP_FLAGS + ["-I" + INCLUDE, "-c", "-o", o, os.path.join(HERE, s)]
                for s, o in zip(ENGINE_SOURCES, objs)]
        workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS") or 0) or os.cpu_count() or 1))
        with ThreadPoolExecutor(max_workers=workers) as ex:
            list(ex.map(_run, jobs))
        _run(["hipcc", "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs)
    finally:
        shutil.rmtree(objdir, ignore_errors=True)
    os.replace(tmp, out)
    return out


def build_module():
    import pybind11

    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = [
        "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden",
        "-I" + pybind11.get_include(), "-I" + py_inc, "-I" + INCLUDE,
        os.path.join(HERE, "pgt_module.cpp"),
        "-L" + PKG, "-lpgt_ik", "-Wl,-rpath,$ORIGIN",
        "-o", tmp,
    ]
    _run(cmd)
    os.replace(tmp, out)
    return out


def build_all():
    build_engine()
    build_module()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "engine" in what:
        build_engine()
    if "module" in what:
        build_module()
    if "all" in what:
        build_all()
