"""Build the native parts in-tree, for gfx950:

* ``libdescriptor_amd/libldsf.so``   -- HIP kernels, the C-ABI (include/ldsf.h),
  the drop-in C++ API (include/Descriptors.hpp, SymmetryFunctions.hpp) and the
  neighbour-list producer;
* ``libdescriptor_amd/libdescriptor<ext>.so`` -- the pybind11 module, linked to
  libldsf.so with an $ORIGIN rpath.

Usage: ``python libdescriptor_amd/build.py`` (or ``build()`` from Python). It is run
as a file: the package refuses to import until this has built it.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = "gfx950"

HIP_SOURCES = ["sf_general.hip", "sf_wave.hip", "sf_wave_fp64.hip", "sf_wave_fp64_fused.hip",
               "sf_plan.hip", "ldsf_runtime.hip", "sf_server.hip", "sf_layout.hip",
               "ldsf_noalloc.hip", "sf_stage.hip"]
# per-source options: the wave kernels take the iterative ILP scheduler (on the
# 262,144-atom system the fused kernel 2.3 % faster in FP32-bulk and 1.5 % in
# FP64, neither with scratch), except the separate FP64 forward and adjoint,
# which were 1-2.5 % slower under it
_ILP = ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]
HIP_EXTRA = {"sf_wave.hip": _ILP, "sf_wave_fp64_fused.hip": _ILP}
CXX_SOURCES = ["sf_params.cpp", "neighbor_list.cpp", "descriptors.cpp"]

# Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", INCLUDE,
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument", "-Wno-unknown-pragmas"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-I", INCLUDE, "-Wno-unknown-pragmas"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _obj(src):
    return os.path.join(BUILD, os.path.splitext(src)[0] + ".o")


def ext_suffix():
    return sysconfig.get_config_var("EXT_SUFFIX") or ".so"


def lib_path():
    return os.path.join(HERE, "libldsf.so")


def module_path():
    return os.path.join(HERE, "libdescriptor" + ext_suffix())


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC] + HIP_FLAGS + HIP_EXTRA.get(s, []) +
                    ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    for s in CXX_SOURCES:
        jobs.append([CXX] + CXX_FLAGS + ["-c", os.path.join(CSRC, s), "-o", _obj(s)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        outs = list(ex.map(_run, jobs))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    objs = [_obj(s) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = lib_path() + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs +
         ["-Wl,--no-undefined", "-Wl,-soname,libldsf.so"])
    os.replace(tmp, lib_path())
    import pybind11
    py_inc = sysconfig.get_paths()["include"]
    tmp = module_path() + ".tmp"
    _run([CXX] + CXX_FLAGS + ["-shared", "-I", pybind11.get_include(), "-I", py_inc,
                              os.path.join(CSRC, "python_bindings.cpp"), "-o", tmp,
                              "-L", HERE, "-lldsf", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, module_path())
    return lib_path(), module_path()


if __name__ == "__main__":
    for p in build(verbose="-v" in sys.argv):
        print("built", p)
