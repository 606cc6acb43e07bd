"""Python side of the learned-correction network (libgamera_net.so).

``GamNet`` wraps one network instance sized for a block's ghost-inclusive
cell grid; the forward, backward and Adam kernels are the library's own (BF16
MFMA, FP32 accumulation).  ``init_params`` draws He-scaled weights from a seed.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from kaiju_amd.gamera import HipStepError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libgamera_net.so")
_fp = C.POINTER(C.c_float)
_lib = None

CO, CI, TAPS = (32, 32, 8), (8, 32, 32), 27


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HipStepError(f"{LIB_PATH} is missing; run __graft_entry__.build() (or python -m kaiju_amd.build net)")
        L = C.CDLL(LIB_PATH)
        H = C.c_void_p
        sig = {
            "gam_net_num_params": (C.c_int, []),
            "gam_net_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
            "gam_net_destroy": (C.c_int, [H]),
            "gam_net_error": (C.c_char_p, [H]),
            "gam_net_stream": (C.c_void_p, [H]),
            "gam_net_set_stream": (C.c_int, [H, C.c_void_p]),
            "gam_net_sync": (C.c_int, [H]),
            "gam_net_set_params": (C.c_int, [H, _fp]),
            "gam_net_get_params": (C.c_int, [H, _fp, _fp]),
            "gam_net_set_scales": (C.c_int, [H] + [C.c_float] * 5),
            "gam_net_features": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int]),
            "gam_net_set_features": (C.c_int, [H, _fp]),
            "gam_net_forward": (C.c_int, [H]),
            "gam_net_infer": (C.c_int, [H, C.c_void_p, C.c_void_p, C.c_int, C.c_int]),
            "gam_net_infer_event": (C.c_void_p, [H]),
            "gam_net_output": (C.c_void_p, [H]),
            "gam_net_get_output": (C.c_int, [H, _fp]),
            "gam_net_zero_grad": (C.c_int, [H]),
            "gam_net_backward": (C.c_int, [H, C.c_void_p]),
            "gam_net_backward_host": (C.c_int, [H, _fp]),
            "gam_net_adam": (C.c_int, [H, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
            "gam_net_get_act": (C.c_int, [H, C.c_int, C.POINTER(C.c_ushort)]),
            "gam_net_grad_buffer": (C.c_void_p, [H]),
            "gam_net_set_offsets": (C.c_int, [H, C.c_float, C.c_float]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def num_params() -> int:
    return lib().gam_net_num_params()


def init_params(seed: int = 0, out_scale: float = 0.1) -> np.ndarray:
    """He-normal weights for the (leaky) ReLU layers, a small last layer, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (co, ci) in enumerate(zip(CO, CI)):
        std = np.sqrt(2.0 / (TAPS * ci)) * (out_scale if l == 2 else 1.0)
        parts.append(rng.normal(0.0, std, size=TAPS * co * ci))
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


class GamNet:
    def __init__(self, Ni: int, Nj: int, Nk: int):
        h = C.c_void_p()
        rc = lib().gam_net_create(Ni, Nj, Nk, C.byref(h))
        if rc != 0 or not h.value:
            raise HipStepError(f"gam_net_create rc={rc}")
        self.h = h
        self.G = (Ni, Nj, Nk)
        self.nc = Ni * Nj * Nk

    def _chk(self, rc, what):
        if rc < 0:
            raise HipStepError(f"{what}: rc={rc} {lib().gam_net_error(self.h).decode()}")
        return rc

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().gam_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, w: np.ndarray):
        w = np.ascontiguousarray(w, dtype=np.float32)
        assert w.size == num_params()
        self._chk(lib().gam_net_set_params(self.h, w.ctypes.data_as(_fp)), "gam_net_set_params")

    def get_params(self):
        w = np.zeros(num_params(), np.float32)
        g = np.zeros(num_params(), np.float32)
        self._chk(lib().gam_net_get_params(self.h, w.ctypes.data_as(_fp), g.ctypes.data_as(_fp)), "gam_net_get_params")
        return w, g

    def set_scales(self, rho=1.0, v=1.0, p=1.0, b=1.0, gamma=5.0 / 3.0):
        self._chk(lib().gam_net_set_scales(self.h, rho, v, p, b, gamma), "gam_net_set_scales")

    def set_offsets(self, rho0=0.0, p0=0.0):
        self._chk(lib().gam_net_set_offsets(self.h, rho0, p0), "gam_net_set_offsets")

    def features_from_device(self, Gas_ptr: int, B_ptr: int, precision: int):
        self._chk(lib().gam_net_features(self.h, Gas_ptr, B_ptr, precision), "gam_net_features")

    def set_features(self, feat: np.ndarray):
        feat = np.ascontiguousarray(feat, dtype=np.float32)
        assert feat.shape == (self.nc, 8)
        self._chk(lib().gam_net_set_features(self.h, feat.ctypes.data_as(_fp)), "gam_net_set_features")

    def forward(self):
        self._chk(lib().gam_net_forward(self.h), "gam_net_forward")

    def infer(self, Gas_ptr: int, B_ptr: int, precision: int, halo: int):
        """Features and the three layers, output stored on the cells at
        least `halo` from the grid faces (equal to forward's there)."""
        self._chk(lib().gam_net_infer(self.h, Gas_ptr, B_ptr, precision, halo), "gam_net_infer")

    def infer_event(self) -> int:
        """The hipEvent_t recorded after each infer (for gam_set_correction_wait)."""
        return lib().gam_net_infer_event(self.h)

    def output_ptr(self) -> int:
        return lib().gam_net_output(self.h)

    def get_output(self) -> np.ndarray:
        o = np.zeros((self.nc, 8), np.float32)
        self._chk(lib().gam_net_get_output(self.h, o.ctypes.data_as(_fp)), "gam_net_get_output")
        return o

    def get_act(self, layer: int) -> np.ndarray:
        a = np.zeros((self.nc, 32), np.uint16)
        self._chk(lib().gam_net_get_act(self.h, layer, a.ctypes.data_as(C.POINTER(C.c_ushort))), "gam_net_get_act")
        return (a.astype(np.uint32) << 16).view(np.float32).astype(np.float64)

    def grad_buffer_ptr(self) -> int:
        return lib().gam_net_grad_buffer(self.h)

    def zero_grad(self):
        self._chk(lib().gam_net_zero_grad(self.h), "gam_net_zero_grad")

    def backward(self, dout_ptr: int):
        self._chk(lib().gam_net_backward(self.h, dout_ptr), "gam_net_backward")

    def backward_host(self, dout: np.ndarray):
        dout = np.ascontiguousarray(dout, dtype=np.float32)
        assert dout.shape == (self.nc, 8)
        self._chk(lib().gam_net_backward_host(self.h, dout.ctypes.data_as(_fp)), "gam_net_backward_host")

    def adam(self, lr: float, t: int, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8):
        self._chk(lib().gam_net_adam(self.h, lr, b1, b2, eps, t), "gam_net_adam")

    def set_stream(self, stream_ptr):
        """Queue the net's work on another hipStream_t (None: its own)."""
        self._chk(lib().gam_net_set_stream(self.h, stream_ptr), "gam_net_set_stream")

    def sync(self):
        self._chk(lib().gam_net_sync(self.h), "gam_net_sync")
