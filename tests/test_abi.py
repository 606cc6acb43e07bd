"""The C-ABI libraries load and export every entry point include/gamera_hip.h
and include/gamera_net.h declare (no GPU needed: nothing is called)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gamera_hip.h")
LIB = os.path.join(ROOT, "kaiju_amd", "lib", "libgamera_hip.so")
NET_HEADER = os.path.join(ROOT, "include", "gamera_net.h")
NET_LIB = os.path.join(ROOT, "kaiju_amd", "lib", "libgamera_net.so")
IO_HEADER = os.path.join(ROOT, "include", "gamera_io.h")
IO_LIB = os.path.join(ROOT, "kaiju_amd", "lib", "libgamera_io.so")


def declared_functions(header=HEADER):
    src = open(header).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gam_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("gam_create", "gam_destroy", "gam_set_grid", "gam_put_state", "gam_get_state",
                     "gam_advance_mhd", "gam_calc_dt", "gam_enforce_bcs", "gam_step", "gam_divb",
                     "gam_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build first: python -m kaiju_amd.build hip"
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    data = open(LIB, "rb").read()
    assert b"gfx950" in data


def test_create_rejects_bad_dims_without_touching_the_gpu():
    from kaiju_amd import gamera
    L = gamera.lib()
    d = gamera.GamDims(16, 16, 16, 3)            # nG must be 4
    p = gamera.GamParams()
    h = ctypes.c_void_p()
    assert L.gam_create(ctypes.byref(d), ctypes.byref(p), ctypes.byref(h)) < 0
    assert not h.value
    assert L.gam_num_stages() > 0
    assert L.gam_stage_name(0) == b"clock"


def test_net_library_exports_every_declared_symbol():
    assert os.path.exists(NET_LIB), "build first: python -m kaiju_amd.build net"
    lib = ctypes.CDLL(NET_LIB)
    names = declared_functions(NET_HEADER)
    assert "gam_net_forward" in names and "gam_net_backward" in names and "gam_net_adam" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in open(NET_LIB, "rb").read()
    assert lib.gam_net_num_params() == 27 * (32 * 8 + 32 * 32 + 8 * 32) + 32 + 32 + 8


def test_io_library_exports_every_declared_symbol():
    from kaiju_amd import build
    if build.hdf5_prefix() is None:
        import pytest
        pytest.skip("no HDF5 C library in this image")
    assert os.path.exists(IO_LIB), "build first: python -m kaiju_amd.build io"
    lib = ctypes.CDLL(IO_LIB)
    names = declared_functions(IO_HEADER)
    assert "gam_io_read_restart" in names and "gam_io_read_grid" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
