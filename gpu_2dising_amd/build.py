"""Builds the engine for gfx950, in-tree:

  gpu_2dising_amd/libgasp_amd.so      the C-ABI library (include/gasp_amd.h)
  gpu_2dising_amd/bin/gasp            the command line front-end
  gpu_2dising_amd/gasp.<abi>.so       the PyGASP-compatible extension module `gasp`

Run as `python -m gpu_2dising_amd.build` (or `python gpu_2dising_amd/build.py`).
Every call recompiles.  hipcc cross-compiles without a GPU.
"""
from __future__ import annotations

import os
import shlex
import subprocess
import sys
import sysconfig

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
BIN = os.path.join(PKG, "bin")
ARCH = os.environ.get("GASP_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # no packed-FP32 instruction selection in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-Wno-missing-field-initializers"]

LIB = os.path.join(PKG, "libgasp_amd.so")
HDF5_LIB = os.path.join(PKG, "libgasp_hdf5.so")
# HDF5 C library (1.10) used for the Python module's grid files
HDF5_ROOT = os.environ.get("GASP_HDF5_ROOT", "/opt/conda")
CLI = os.path.join(BIN, "gasp")
EXT_SUFFIX = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
PYEXT = os.path.join(PKG, "gasp" + EXT_SUFFIX)


def _run(cmd: list[str]) -> None:
    print("+", " ".join(shlex.quote(c) for c in cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_lib() -> str:
    os.makedirs(BUILD, exist_ok=True)
    eng_o = os.path.join(BUILD, "gasp_engine.o")
    cnn_o = os.path.join(BUILD, "gasp_cnn.o")
    host_o = os.path.join(BUILD, "gasp_host.o")
    cpu_o = os.path.join(BUILD, "gasp_cpu.o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "gasp_engine.hip"), "-o", eng_o])
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "gasp_cnn.hip"), "-o", cnn_o])
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "gasp_host.cpp"), "-o", host_o])
    _run([CXX, *CXX_FLAGS, "-c", os.path.join(CSRC, "gasp_cpu.cpp"), "-o", cpu_o])
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, eng_o, cnn_o, host_o, cpu_o])
    os.replace(tmp, LIB)
    return LIB


def build_hdf5() -> str:
    """libgasp_hdf5.so: create/write_ising_grids_hdf5 (include/gasp_hdf5.h)."""
    tmp = HDF5_LIB + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "gasp_hdf5.cpp"),
          # after the system directories, so that only hdf5*.h come from there
          "-idirafter", os.path.join(HDF5_ROOT, "include"),
          f"-L{os.path.join(HDF5_ROOT, 'lib')}", "-lhdf5", f"-Wl,-rpath,{os.path.join(HDF5_ROOT, 'lib')}"])
    os.replace(tmp, HDF5_LIB)
    return HDF5_LIB


def build_cli() -> str:
    os.makedirs(BIN, exist_ok=True)
    _run([CXX, *CXX_FLAGS, "-o", CLI, os.path.join(CSRC, "gasp_cli.cpp"),
          f"-L{PKG}", "-lgasp_amd", "-Wl,-rpath,$ORIGIN/.."])
    return CLI


def build_pyext() -> str:
    import numpy as np

    py_inc = sysconfig.get_paths()["include"]
    tmp = PYEXT + ".tmp"
    _run([CXX, *CXX_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "gasp_pymod.cpp"),
          f"-I{py_inc}", f"-I{np.get_include()}", f"-L{PKG}", "-lgasp_amd", "-lgasp_hdf5",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, PYEXT)
    return PYEXT


def build_all() -> None:
    build_lib()
    build_hdf5()
    if os.path.exists(os.path.join(CSRC, "gasp_cli.cpp")):
        build_cli()
    if os.path.exists(os.path.join(CSRC, "gasp_pymod.cpp")):
        build_pyext()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "lib": build_lib, "hdf5": build_hdf5, "cli": build_cli, "pyext": build_pyext}[w]()
