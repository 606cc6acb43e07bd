"""Build the CPU emulation of libgamera_hip.so from the same kernel sources.

TEST INFRASTRUCTURE ONLY.  ``libgamera_emu.so`` runs every launch as nested
CPU loops (tests/emu/gam_emu.h); ``libgamera_emu_asan.so`` adds
AddressSanitizer so any out-of-bounds access of a kernel aborts on the CPU.
"""
import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "kaiju_amd", "csrc")
OUT = os.path.join(HERE, "_build")
# every source the step library's translation unit can include
SRCS = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hip", ".h"))) + \
       [os.path.join(HERE, "gam_emu.h"), os.path.join(ROOT, "include", "gamera_hip.h")]


def _cxx():
    return os.environ.get("CXX") or shutil.which("g++") or "g++"


def build(asan: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    lib = os.path.join(OUT, "libgamera_emu_asan.so" if asan else "libgamera_emu.so")
    if os.path.exists(lib) and os.path.getmtime(lib) >= max(os.path.getmtime(s) for s in SRCS):
        return lib
    cmd = [_cxx(), "-std=c++17", "-O1", "-g", "-fPIC", "-shared", "-DGAM_EMU", "-ffp-contract=off",
           "-fno-fast-math", "-x", "c++", "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-I", HERE,
           os.path.join(CSRC, "gam_api.hip"), "-o", lib + ".tmp"]
    if asan:
        cmd[1:1] = ["-fsanitize=address", "-fno-omit-frame-pointer"]
    subprocess.run(cmd, check=True)
    os.replace(lib + ".tmp", lib)
    return lib


def asan_runtime() -> str:
    return subprocess.run([_cxx(), "-print-file-name=libasan.so"], capture_output=True, text=True,
                          check=True).stdout.strip()


if __name__ == "__main__":
    print(build(False))
    print(build(True))
