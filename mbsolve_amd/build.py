"""Builds the native parts of mbsolve_amd in-tree.

* ``libmbsolve_hip.so`` -- gfx950 kernels and the C ABI of
  ``include/mbsolve_hip.h`` (hipcc).
* ``libmbsolve.so``     -- the C++ mbsolve API, host set-up numerics and the
  ``hip-fdtd-*`` solver classes (g++, links the C ABI library).
* ``_mbsolve*.so``      -- Python bindings (pybind11).

Run ``python -m mbsolve_amd.build``; object files go to ``mbsolve_amd/_build``.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
BUILD = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["mbh_capi.hip", "mbh_inst_basic.hip", "mbh_inst_rodr.hip",
               "mbh_inst_zform.hip", "mbh_inst_zmfma.hip", "mbh_inst_red.hip"]
CPP_SOURCES = ["lib_model.cpp", "numerics.cpp", "cv_representation.cpp",
               "fdtd_setup.cpp", "solver_hip_fdtd.cpp"]

# FP64 reference-order kernels must not contract mul+add into FMA (explicit
# fma() calls are kept); packed FP32 instructions are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-I" + os.path.join(ROOT, "include")]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
             "-Wno-sign-compare",
             "-I" + os.path.join(PKG, "cpp", "include"),
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    out = "\n".join(l for l in p.stdout.splitlines()
                    if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if p.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def _hip_object(src):
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    extra = os.environ.get("MBSOLVE_HIP_EXTRA_FLAGS", "").split()
    _run([HIPCC] + HIP_FLAGS + extra +
         ["-c", os.path.join(PKG, "csrc", src), "-o", obj])
    return obj


def _cpp_object(src):
    obj = os.path.join(BUILD, src.replace(".cpp", ".o"))
    _run([CXX] + CXX_FLAGS + ["-c", os.path.join(PKG, "cpp", "src", src),
                              "-o", obj])
    return obj


def _pybind_includes():
    import pybind11
    return ["-I" + pybind11.get_include(),
            "-I" + sysconfig.get_paths()["include"]]


def build(verbose=True, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    log = print if verbose else (lambda *a, **k: None)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        hip_objs = list(ex.map(_hip_object, HIP_SOURCES))
        cpp_objs = list(ex.map(_cpp_object, CPP_SOURCES))
    hip_lib = os.path.join(PKG, "libmbsolve_hip.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + hip_objs +
         ["-o", hip_lib])
    log("built", hip_lib)
    lib = os.path.join(PKG, "libmbsolve.so")
    _run([CXX, "-shared", "-fPIC"] + cpp_objs +
         ["-L" + PKG, "-lmbsolve_hip", "-Wl,-rpath,$ORIGIN", "-o", lib])
    log("built", lib)
    ext = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    mod = os.path.join(PKG, "_mbsolve" + ext)
    _run([CXX] + CXX_FLAGS + _pybind_includes() +
         ["-shared", os.path.join(PKG, "cpp", "python", "bindings.cpp"),
          "-L" + PKG, "-lmbsolve", "-lmbsolve_hip", "-Wl,-rpath,$ORIGIN",
          "-o", mod])
    log("built", mod)
    out = [hip_lib, lib, mod]
    out += _build_writer_hdf5(ext, log)
    return out


def _build_writer_hdf5(ext, log):
    """Optional HDF5 writer (needs the HDF5 C library and headers)."""
    root = os.environ.get("MBSOLVE_HDF5_ROOT", "/opt/conda")
    inc, libdir = os.path.join(root, "include"), os.path.join(root, "lib")
    if not (os.path.exists(os.path.join(inc, "hdf5.h")) and
            os.path.exists(os.path.join(libdir, "libhdf5.so"))):
        log("HDF5 not found under %s: writer-hdf5 not built" % root)
        return []
    obj = os.path.join(BUILD, "writer_hdf5.o")
    _run([CXX] + CXX_FLAGS + ["-isystem", inc, "-c",
                              os.path.join(PKG, "cpp", "src",
                                           "writer_hdf5.cpp"), "-o", obj])
    wlib = os.path.join(PKG, "libmbsolve_writer_hdf5.so")
    _run([CXX, "-shared", "-fPIC", obj, "-L" + PKG, "-lmbsolve",
          "-L" + libdir, "-lhdf5", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath," + libdir, "-o", wlib])
    log("built", wlib)
    mod = os.path.join(PKG, "_writerhdf5" + ext)
    _run([CXX] + CXX_FLAGS + _pybind_includes() +
         ["-shared", os.path.join(PKG, "cpp", "python", "writerhdf5.cpp"),
          "-L" + PKG, "-lmbsolve_writer_hdf5", "-lmbsolve",
          "-Wl,-rpath,$ORIGIN", "-o", mod])
    log("built", mod)
    return [wlib, mod]


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
