"""Build the native parts in-tree.

* ``kaiju_amd/lib/libgamera_hip.so`` -- the HIP step for gfx950 (hipcc).
* ``kaiju_amd/lib/libgamera_net.so`` -- the learned-correction network
  (BF16 MFMA convolutions, backward and Adam) for gfx950.
* ``kaiju_amd/lib/libgamera_io.so`` -- the reference's HDF5 grid and restart
  files on the host (g++ against the HDF5 C library; built when HDF5 is found).
* ``kaiju_amd/bin/gamerax_gpu`` and ``kaiju_amd/lib/libgamhost.so`` -- the
  Fortran host (amdflang) that keeps the reference's Model/Grid/State types and
  calls the C-ABI.

Usage: ``python -m kaiju_amd.build [hip|net|io|fortran|all]``.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
HIP_LIB = os.path.join(LIBDIR, "libgamera_hip.so")
NET_LIB = os.path.join(LIBDIR, "libgamera_net.so")
IO_LIB = os.path.join(LIBDIR, "libgamera_io.so")
# HDF5 1.10 C library and headers of the image
HDF5_PREFIXES = [os.environ.get("HDF5_DIR", ""), "/opt/conda", "/usr"]

# one translation unit: gam_api.hip includes gam_step.hip and gam_setup.hip
HIP_SOURCES = ["gam_api.hip"]
HIP_DEPS = ["gam_api.hip", "gam_step.hip", "gam_setup.hip", "gam_common.h"]

# -ffp-contract=off keeps every multiply and add separately rounded, as the
# reference's CPU build does; packed FP32 ops are disabled on this toolchain.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _flang() -> str:
    for c in (os.environ.get("FC"), shutil.which("amdflang"), "/opt/rocm/bin/amdflang", shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("amdflang not found")


def _build_so(srcs, out: str, verbose: bool) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = out + ".tmp"
    cmd = [_hipcc(), *HIPCC_FLAGS, f"-Wl,-soname,{os.path.basename(out)}", "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           *[os.path.join(CSRC, s) for s in srcs], "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)
    return out


def build_hip(verbose: bool = False) -> str:
    return _build_so(HIP_SOURCES, HIP_LIB, verbose)


def build_net(verbose: bool = False) -> str:
    return _build_so(["gam_net.hip"], NET_LIB, verbose)


def hdf5_prefix():
    """Directory holding include/hdf5.h and lib/libhdf5_hl.so, or None."""
    for p in HDF5_PREFIXES:
        if p and os.path.exists(os.path.join(p, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(p, "lib", "libhdf5_hl.so")):
            return p
    return None


def build_io(verbose: bool = False) -> str:
    """libgamera_io.so: host C++ over the HDF5 C API (no GPU code)."""
    pre = hdf5_prefix()
    if pre is None:
        raise RuntimeError("HDF5 C library not found (set HDF5_DIR)")
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = IO_LIB + ".tmp"
    cxx = os.environ.get("CXX") or shutil.which("g++") or "g++"
    cmd = [cxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I", os.path.join(ROOT, "include"),
           "-I", os.path.join(pre, "include"), os.path.join(CSRC, "gam_io.cpp"), "-o", tmp,
           "-L", os.path.join(pre, "lib"), "-lhdf5_hl", "-lhdf5", "-Wl,-rpath," + os.path.join(pre, "lib"),
           "-Wl,-soname,libgamera_io.so"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, IO_LIB)
    return IO_LIB


FORTRAN_MODULES = ["gamera_hip_c.F90", "basetypes_gpu.F90", "gamtypes_gpu.F90", "gamapp_gpu.F90",
                   "gamtypessub_gpu.F90", "gamprob_gpu.F90"]


def build_fortran(verbose: bool = False, lib: str = HIP_LIB, exe: str = None) -> str:
    """Fortran host (amdflang): the modules and the gamerax_gpu driver, linked
    against a build of the C-ABI (the HIP library by default)."""
    os.makedirs(BINDIR, exist_ok=True)
    objdir = os.path.join(PKG, "_fobj")
    os.makedirs(objdir, exist_ok=True)
    fc = _flang()
    fflags = ["-O2", "-cpp", "-ffree-form", "-module-dir", objdir, "-I", objdir]
    objs = []
    for m in FORTRAN_MODULES:
        src = os.path.join(FSRC, m)
        obj = os.path.join(objdir, m.replace(".F90", ".o"))
        cmd = [fc, *fflags, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    exe = exe or os.path.join(BINDIR, "gamerax_gpu")
    libdir, libfile = os.path.split(lib)
    libname = libfile[3:-3] if libfile.startswith("lib") and libfile.endswith(".so") else libfile
    drv = os.path.join(FSRC, "gamerax_gpu.F90")
    cmd = [fc, *fflags, drv, *objs, "-o", exe, "-L", libdir, "-l" + libname,
           "-Wl,-rpath,$ORIGIN/" + os.path.relpath(libdir, os.path.dirname(exe))]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    return exe


def build_all(verbose: bool = False):
    build_hip(verbose)
    build_net(verbose)
    if hdf5_prefix() is not None:
        build_io(verbose)
    if os.path.isdir(FSRC) and all(os.path.exists(os.path.join(FSRC, m)) for m in FORTRAN_MODULES):
        build_fortran(verbose)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what == "hip":
        build_hip(True)
    elif what == "net":
        build_net(True)
    elif what == "io":
        build_io(True)
    elif what == "fortran":
        build_fortran(True)
    else:
        build_all(True)
