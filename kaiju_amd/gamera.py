"""Host-side mirror of the reference's Gamera app interface over the HIP C-ABI.

``GamApp`` keeps the shape of the reference's ``gamApp_T``
(reference ``src/base/types/gamtypes.F90:294-314``): a ``Model`` with the time
bookkeeping, the grid sizes, ``State``/``oState`` access, and the per-step
calls ``AdvanceModel(dt)`` (``gamtypessub.F90:78-94``) and ``stepGamera()``
(``gamapp.F90:73-94``).  All arithmetic runs in ``libgamera_hip.so``; this module
only moves plain arrays across the C boundary and fails loudly when the library
is missing.

Arrays are numpy arrays in Fortran order with the reference's shapes
(``Gas`` is ``(Ni, Nj, Nk, 5)``, ``magFlux`` is ``(Ni+1, Nj+1, Nk+1, 3)``).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Optional

import numpy as np

from . import configs

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libgamera_hip.so")

_dp = C.POINTER(C.c_double)


class GamDims(C.Structure):
    _fields_ = [("Nip", C.c_int32), ("Njp", C.c_int32), ("Nkp", C.c_int32), ("nG", C.c_int32)]


class GamParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("gamma", "CFL", "pdmb", "Vd0", "cHogH", "cHogM", "Ca", "dFloor",
                                          "pFloor", "tFin", "limDT0", "limCPR")] + \
               [(n, C.c_int32) for n in ("rmeth", "doLFMLim", "doBoris", "doMHD", "do25D", "doCPR",
                                         "fixedTimestep", "precision")]


class GamClock(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("t", "dt", "dt0", "stateTime", "oStateTime")] + \
               [(n, C.c_int32) for n in ("ts", "status", "dt0_after", "ncpr")]


class HipStepError(RuntimeError):
    pass


class DeviceArray:
    """A device buffer of the library's GPU holding a numpy array's bytes
    (for correction fields and gradient buffers; no torch needed)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape, self.dtype = tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        rc = lib().gam_device_alloc(self.nbytes, C.byref(p))
        if rc != 0 or not p.value:
            raise HipStepError(f"gam_device_alloc({self.nbytes}) rc={rc}")
        self.ptr = p.value

    @classmethod
    def of(cls, a: np.ndarray, dtype=np.float32):
        d = cls(a.shape, dtype)
        d.set(a)
        return d

    def set(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype).reshape(self.shape)
        if lib().gam_memcpy(self.ptr, a.ctypes.data, self.nbytes) != 0:
            raise HipStepError("gam_memcpy to device failed")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if lib().gam_memcpy(out.ctypes.data, self.ptr, self.nbytes) != 0:
            raise HipStepError("gam_memcpy from device failed")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().gam_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


_lib = None


def use_library(path: str):
    """Point the binding at another build of the same C-ABI (the tests use the
    CPU emulation build to check kernels before they reach a GPU)."""
    global LIB_PATH, _lib
    LIB_PATH = path
    _lib = None


def lib():
    """Load libgamera_hip.so (no fallback: the step exists only as HIP code)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipStepError(f"{LIB_PATH} is missing; run __graft_entry__.build() (or python -m kaiju_amd.build hip)")
    L = C.CDLL(LIB_PATH)
    H = C.c_void_p
    sig = {
        "gam_create": (C.c_int, [C.POINTER(GamDims), C.POINTER(GamParams), C.POINTER(H)]),
        "gam_destroy": (C.c_int, [H]),
        "gam_last_error": (C.c_char_p, [H]),
        "gam_sync": (C.c_int, [H]),
        "gam_stream": (C.c_void_p, [H]),
        "gam_set_grid": (C.c_int, [H] + [_dp] * 12),
        "gam_set_corners": (C.c_int, [H, _dp]),
        "gam_get_grid": (C.c_int, [H] + [_dp] * 12),
        "gam_put_state": (C.c_int, [H, C.c_int, C.c_double, _dp, _dp, _dp]),
        "gam_get_state": (C.c_int, [H, C.c_int, C.POINTER(C.c_double), _dp, _dp, _dp]),
        "gam_set_clock": (C.c_int, [H, C.POINTER(GamClock)]),
        "gam_get_clock": (C.c_int, [H, C.POINTER(GamClock)]),
        "gam_ic_orszag_tang": (C.c_int, [H] + [C.c_double] * 4),
        "gam_ic_blast_wave": (C.c_int, [H] + [C.c_double] * 9),
        "gam_hatch": (C.c_int, [H]),
        "gam_advance_mhd": (C.c_int, [H]),
        "gam_calc_dt": (C.c_int, [H, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "gam_enforce_bcs": (C.c_int, [H]),
        "gam_step": (C.c_int, [H, C.c_int]),
        "gam_divb": (C.c_int, [H, C.POINTER(C.c_double)]),
        "gam_divb_noalloc": (C.c_int, [H, C.POINTER(C.c_double)]),
        "gam_get_work": (C.c_int, [H, C.c_int, _dp]),
        "gam_set_timing": (C.c_int, [H, C.c_int]),
        "gam_get_timing": (C.c_int, [H, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
        "gam_stage_name": (C.c_char_p, [C.c_int]),
        "gam_num_stages": (C.c_int, []),
        "gam_stage_bytes": (C.c_int, [H, C.c_int, C.POINTER(C.c_double)]),
        "gam_set_tile": (C.c_int, [H, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
        "gam_get_tile": (C.c_int, [H, C.c_int, C.POINTER(C.c_int)]),
        "gam_set_background": (C.c_int, [H, _dp, _dp, _dp, _dp, _dp]),
        "gam_get_background": (C.c_int, [H, _dp, _dp, _dp, _dp, _dp]),
        "gam_background_dipole": (C.c_int, [H, C.c_double, _dp, _dp]),
        "gam_graph_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
        "gam_graph_create_packed": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
        "gam_graph_nodes": (C.c_int, [C.c_void_p, C.POINTER(C.c_int)]),
        "gam_graph_launch": (C.c_int, [C.c_void_p, C.c_int]),
        "gam_graph_sync": (C.c_int, [C.c_void_p]),
        "gam_graph_destroy": (C.c_int, [C.c_void_p]),
        "gam_graph_error": (C.c_char_p, [C.c_void_p]),
        "gam_state_device": (C.c_int, [H, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                       C.POINTER(C.c_void_p)]),
        "gam_set_correction": (C.c_int, [H, C.c_void_p, C.c_double]),
        "gam_set_correction_wait": (C.c_int, [H, C.c_void_p]),
        "gam_set_bcs": (C.c_int, [H, C.POINTER(C.c_int)]),
        "gam_get_bcs": (C.c_int, [H, C.POINTER(C.c_int)]),
        "gam_set_ring": (C.c_int, [H, C.c_int, C.POINTER(C.c_int)]),
        "gam_set_wind": (C.c_int, [H, C.c_double, C.c_double, _dp, _dp]),
        "gam_set_wind_series": (C.c_int, [H, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double,
                                          C.c_double]),
        "gam_get_ring": (C.c_int, [H, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "gam_device_alloc": (C.c_int, [C.c_size_t, C.POINTER(C.c_void_p)]),
        "gam_device_free": (C.c_int, [C.c_void_p]),
        "gam_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
        "gam_correction_target": (C.c_int, [H, _dp, _dp]),
        "gam_correction_target_device": (C.c_int, [H, C.c_void_p, C.c_void_p]),
        "gam_put_state_device": (C.c_int, [H, C.c_int, C.c_double, C.c_void_p, C.c_void_p]),
        "gam_correction_grad": (C.c_int, [H, C.c_void_p, C.c_double, C.c_double, _dp, C.c_void_p,
                                          C.POINTER(C.c_double)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a: Optional[np.ndarray]):
    if a is None:
        return None
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim <= 1):
        raise ValueError("arrays must be float64 in Fortran order")
    return a.ctypes.data_as(_dp)


def params_from_spec(spec: configs.CaseSpec) -> GamParams:
    m = spec.model
    p = GamParams()
    for k in ("gamma", "CFL", "pdmb", "Vd0", "cHogH", "cHogM", "Ca", "dFloor", "pFloor", "tFin", "limDT0", "limCPR"):
        setattr(p, k, float(getattr(m, k)))
    p.rmeth = 1 if m.rmeth.upper() == "7UP" else 0
    p.doLFMLim = int(m.doLFMLim)
    p.doBoris = int(m.doBoris)
    p.doMHD = int(m.doMHD)
    p.do25D = int(spec.do25D)
    p.doCPR = int(m.doCPR)
    p.fixedTimestep = int(m.fixedTimestep)
    p.precision = int(m.precision)
    return p


def gen_corners(spec: configs.CaseSpec) -> np.ndarray:
    """Corner array xyz(isg:ieg+1, jsg:jeg+1, ksg:keg+1, 3) for the spec's grid
    (reference genGridXML, init.F90:669-731, and the warp of configs.blast_wave)."""
    nG = spec.nG
    Nip, Njp, Nkp = spec.N
    x0, x1, y0, y1, z0, z1 = spec.bounds
    dx, dy, dz = (x1 - x0) / Nip, (y1 - y0) / Njp, (z1 - z0) / Nkp
    ii = np.arange(1 - nG, Nip + nG + 2)
    jj = np.arange(1 - nG, Njp + nG + 2)
    kk = np.arange(1 - nG, Nkp + nG + 2)
    X = x0 + (ii - 1) * dx
    Y = y0 + (jj - 1) * dy
    Z = z0 + (kk - 1) * dz
    xyz = np.empty((len(ii), len(jj), len(kk), 3), order="F")
    xyz[..., 0] = X[:, None, None]
    xyz[..., 1] = Y[None, :, None]
    xyz[..., 2] = Z[None, None, :]
    if spec.grid == "warp3":
        Lx, Ly, Lz = x1 - x0, y1 - y0, z1 - z0
        s = spec.warp_amp * (np.sin(2 * math.pi * ((X - x0) / Lx))[:, None, None]
                             * np.sin(2 * math.pi * ((Y - y0) / Ly))[None, :, None]
                             * np.sin(2 * math.pi * ((Z - z0) / Lz))[None, None, :])
        xyz[..., 0] += s * Lx
        xyz[..., 1] += s * Ly
        xyz[..., 2] += s * Lz
    elif spec.grid == "cyl":   # init.F90:712-716: (x1, x2, x3) = (R, phi / 2 pi, z)
        R, ph = X[:, None], 2 * math.pi * Y[None, :]
        xyz[..., 0] = (R * np.cos(ph))[..., None]
        xyz[..., 1] = (R * np.sin(ph))[..., None]
        xyz[..., 2] = Z[None, None, :]
    elif spec.grid == "sph":
        from kaiju_amd import msph
        p = spec.ic_params
        return msph.corners(spec.N, nG, p["r_in"], p["r_out"], p["theta0"], p.get("axis", False))
    elif spec.grid != "cart":
        raise ValueError(f"grid {spec.grid} is generated by the caller; use set_corners()")
    return xyz


class Model:
    """Model_T time bookkeeping mirrored from the device clock."""

    def __init__(self, spec: configs.CaseSpec):
        self.spec = spec.model
        self.t = 0.0
        self.dt = 0.0
        self.dt0 = 0.0
        self.ts = 0
        self.tFin = spec.model.tFin
        self.status = 0
        self.nOut = 0    # Model%IO counters a restart file carries
        self.nRes = 0


class GamApp:
    """gamApp_T over the HIP step (one handle = one block on one GPU)."""

    def __init__(self, spec: configs.CaseSpec):
        L = lib()
        self.spec = spec
        self.Model = Model(spec)
        Nip, Njp, Nkp = spec.N
        self.nG = spec.nG
        self.Ni, self.Nj, self.Nk = Nip + 2 * spec.nG, Njp + 2 * spec.nG, Nkp + 2 * spec.nG
        self.G = (self.Ni, self.Nj, self.Nk)
        self.F = (self.Ni + 1, self.Nj + 1, self.Nk + 1)
        d = GamDims(Nip, Njp, Nkp, spec.nG)
        self._params = params_from_spec(spec)
        h = C.c_void_p()
        rc = L.gam_create(C.byref(d), C.byref(self._params), C.byref(h))
        if rc != 0 or not h.value:
            raise HipStepError(f"gam_create failed rc={rc}")
        self.h = h
        if spec.wind_series is not None:
            self.set_wind_series(spec.wind_series)
        elif spec.wind is not None:
            self.set_wind(*spec.wind)
        if any(b != "periodic" for b in spec.bcs):
            self.set_bcs(spec.bcs)
        if spec.ring:
            self.set_ring(spec.ring)

    # ---------------------------------------------------------------- utils
    def _chk(self, rc: int, what: str) -> int:
        if rc < 0:
            msg = lib().gam_last_error(self.h).decode()
            raise HipStepError(f"{what}: rc={rc} {msg}")
        return rc

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            rc = lib().gam_destroy(self.h)
            if rc < 0:   # a captured graph still holds the handle
                raise HipStepError(f"gam_destroy: rc={rc} {lib().gam_last_error(self.h).decode()}")
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

    # ---------------------------------------------------------------- setup
    def set_corners(self, xyz: np.ndarray):
        xyz = np.asfortranarray(xyz, dtype=np.float64)
        assert xyz.shape == self.F + (3,)
        self._chk(lib().gam_set_corners(self.h, _p(xyz)), "gam_set_corners")
        self._xyz = xyz

    def set_grid(self, grid: dict):
        names = ("xyz", "xyzcc", "xfc", "face", "Tf", "edge", "Te", "Teb", "volume", "di", "dj", "dk")
        arrs = [np.asfortranarray(grid[n], dtype=np.float64) for n in names]
        self._keep = arrs
        self._chk(lib().gam_set_grid(self.h, *[_p(a) for a in arrs]), "gam_set_grid")
        self._xyz = arrs[0]

    def get_grid(self, only: Optional[tuple] = None) -> dict:
        """The geometry arrays (only: the names to fetch, default all)."""
        G, F = self.G, self.F
        shapes = dict(xyz=F + (3,), xyzcc=G + (3,), xfc=F + (3, 3), face=F + (3,), Tf=F + (9, 3),
                      edge=F + (3,), Te=F + (9, 3), Teb=F + (4, 3), volume=G, di=G, dj=G, dk=G)
        names = ("xyz", "xyzcc", "xfc", "face", "Tf", "edge", "Te", "Teb", "volume", "di", "dj", "dk")
        out = {k: np.zeros(s, order="F") for k, s in shapes.items() if only is None or k in only}
        self._chk(lib().gam_get_grid(self.h, *[_p(out.get(n)) for n in names]), "gam_get_grid")
        return out

    def init_ic(self):
        p, sp = self.spec.ic_params, self.spec
        L = lib()
        if sp.ic == "ot":
            self._chk(L.gam_ic_orszag_tang(self.h, p["rho0"], p["p0"], p["b0"], p["kscale"]), "ic_ot")
        elif sp.ic == "bw":
            self._chk(L.gam_ic_blast_wave(self.h, p["x0"], p["y0"], p["z0"], p["D0"], p["P0"], p["pRat"],
                                          p["dRat"], p["rC"], p["B0"]), "ic_bw")
        elif sp.ic == "msph":   # uniform wind on the spherical grid (host-built, exact face fluxes)
            from kaiju_amd import msph
            gas, mag = msph.uniform_state(gen_corners(sp), p["rho"], p["v"], p["p"], p["b"], sp.model.gamma)
            self.put_state(0, 0.0, gas, mag, None)
        else:
            raise ValueError(sp.ic)

    def hatch(self):
        """init.F90:204-244: first Bxyz, BCs, oState = State, first dt."""
        c = self.get_clock()
        c.dt0_after = self.spec.dt0_after
        self.set_clock(c)
        self._chk(lib().gam_hatch(self.h), "gam_hatch")
        self._sync_model()

    def InitModel(self, corners: Optional[np.ndarray] = None):
        """Reference Hatch (init.F90:38-59) for the spec: corners -> metric on
        the GPU, background field (AddB0, init.F90:182-184) if the spec has one,
        initial condition on the GPU, hatch."""
        self.set_corners(gen_corners(self.spec) if corners is None else corners)
        if self.spec.background is not None:
            b = self.spec.background
            self.background_dipole(b.get("M0", 0.0), b.get("xc", (0.0, 0.0, 0.0)), b.get("bu", (0.0, 0.0, 0.0)))
        self.init_ic()
        self.hatch()

    # ---------------------------------------------------------------- background
    BG_SHAPES = dict(fcB0=(3, 3), edgB0=(2, 3), B0=(3,), bFlux0=(3,), dpB0=(3,))

    def background_dipole(self, M0: float, xc, bu):
        xc = np.ascontiguousarray(xc, dtype=np.float64)
        bu = np.ascontiguousarray(bu, dtype=np.float64)
        self._chk(lib().gam_background_dipole(self.h, float(M0), _p(xc), _p(bu)), "gam_background_dipole")

    def set_background(self, bg: dict):
        arrs = [np.asfortranarray(bg[k], dtype=np.float64) for k in self.BG_SHAPES]
        self._chk(lib().gam_set_background(self.h, *[_p(a) for a in arrs]), "gam_set_background")

    def get_background(self) -> dict:
        out = {k: np.zeros(self.G + shp, order="F") for k, shp in self.BG_SHAPES.items()}
        self._chk(lib().gam_get_background(self.h, *[_p(out[k]) for k in self.BG_SHAPES]), "gam_get_background")
        return out

    # ---------------------------------------------------------------- state
    def put_state(self, which: int, time: float, Gas=None, magFlux=None, Bxyz=None):
        self._chk(lib().gam_put_state(self.h, which, time, _p(Gas), _p(magFlux), _p(Bxyz)), "gam_put_state")

    def put_state_device(self, which: int, time: float, gas_ptr: int, mag_ptr: int):
        """State (0) or oState (1) from FP64 device arrays; Bxyz recomputed."""
        self._chk(lib().gam_put_state_device(self.h, which, time, gas_ptr, mag_ptr), "gam_put_state_device")

    def get_state(self, which: int = 0):
        Gas = np.zeros(self.G + (5,), order="F")
        mag = np.zeros(self.F + (3,), order="F")
        B = np.zeros(self.G + (3,), order="F")
        t = C.c_double()
        self._chk(lib().gam_get_state(self.h, which, C.byref(t), _p(Gas), _p(mag), _p(B)), "gam_get_state")
        return dict(time=t.value, Gas=Gas, magFlux=mag, Bxyz=B)

    # ---------------------------------------------------------------- restart
    CLOCK_FIELDS = ("t", "dt", "dt0", "stateTime", "oStateTime", "ts", "status", "dt0_after", "ncpr")

    def save_restart(self, path: str):
        """Write State, oState and the clock to an .npz file: the content of the
        reference's restart files (gioH5.F90:650-872: both states, ts, dt, dt0),
        enough for a bit-exact resume of this block."""
        c = self.get_clock()
        arrays = {}
        for which, tag in ((0, "State"), (1, "oState")):
            s = self.get_state(which)
            for k in ("Gas", "magFlux", "Bxyz"):
                arrays[f"{tag}_{k}"] = s[k]
            arrays[f"{tag}_time"] = np.float64(s["time"])
        for f in self.CLOCK_FIELDS:
            arrays[f"clock_{f}"] = np.array(getattr(c, f))
        arrays["dims"] = np.array(self.spec.N + (self.nG,))
        arrays["precision"] = np.array(self.spec.model.precision)
        np.savez(path, **arrays)

    def load_restart(self, path: str, corners: Optional[np.ndarray] = None):
        """Resume from save_restart's file (same block shape and precision): the
        grid from the spec's corners (or the given ones) and its background
        field, as Hatch builds them, then both states and the clock."""
        with np.load(path, allow_pickle=False) as z:
            if tuple(z["dims"]) != self.spec.N + (self.nG,) or int(z["precision"]) != self.spec.model.precision:
                raise ValueError(f"{path}: restart of another block shape or precision")
            self.set_corners(gen_corners(self.spec) if corners is None else corners)
            if self.spec.background is not None:
                b = self.spec.background
                self.background_dipole(b.get("M0", 0.0), b.get("xc", (0.0, 0.0, 0.0)), b.get("bu", (0.0, 0.0, 0.0)))
            for which, tag in ((0, "State"), (1, "oState")):
                self.put_state(which, float(z[f"{tag}_time"]), np.asfortranarray(z[f"{tag}_Gas"]),
                               np.asfortranarray(z[f"{tag}_magFlux"]), np.asfortranarray(z[f"{tag}_Bxyz"]))
            c = self.get_clock()
            for f in self.CLOCK_FIELDS:
                setattr(c, f, type(getattr(c, f))(z[f"clock_{f}"].item()))
            self.set_clock(c)
        self._sync_model()

    def write_restart_h5(self, path: str):
        """writeH5Res (gioH5.F90:650-704): the reference's restart file, with
        State, oState, the corners and the nOut/nRes/ts/dt/dt0 attributes."""
        from kaiju_amd import gamio
        c = self.get_clock()
        xyz = getattr(self, "_xyz", None)
        if xyz is None:
            xyz = self.get_grid()["xyz"]
        ioc = gamio.IOClock(nOut=self.Model.nOut, nRes=self.Model.nRes, ts=c.ts, t=c.stateTime, dt=c.dt,
                            dt0=c.dt0, ot=c.oStateTime)
        gamio.write_restart(path, self.spec.N, self.nG, ioc, xyz, self.get_state(0), self.get_state(1),
                            doMHD=bool(self._params.doMHD))

    def read_restart_h5(self, path: str):
        """readH5Restart (gioH5.F90:707-872) plus the grid read from the same
        file (readH5Grid, gioH5.F90:36-124, as a restarting run does): the
        corners, the spec's background field, State and oState (Bxyz derived
        from magFlux), and the clock: t = State%time, oState%time, dt, dt0, ts;
        nRes one past the file's."""
        from kaiju_amd import gamio
        N = gamio.grid_dims(path, self.nG)
        if tuple(N) != tuple(self.spec.N):
            raise ValueError(f"{path}: restart of a {N} block, this block is {self.spec.N}")
        self.set_corners(gamio.read_grid(path, self.nG, N))
        if self.spec.background is not None:
            b = self.spec.background
            self.background_dipole(b.get("M0", 0.0), b.get("xc", (0.0, 0.0, 0.0)), b.get("bu", (0.0, 0.0, 0.0)))
        mhd = bool(self._params.doMHD)
        ioc, st, ost = gamio.read_restart(path, N, self.nG, doMHD=mhd)
        self.put_state(0, ioc.t, st["Gas"], st["magFlux"] if mhd else None, None)
        self.put_state(1, ioc.ot, ost["Gas"], ost["magFlux"] if mhd else None, None)
        c = self.get_clock()
        c.t, c.dt, c.dt0, c.ts = ioc.t, ioc.dt, ioc.dt0, ioc.ts
        c.stateTime, c.oStateTime = ioc.t, ioc.ot
        c.status, c.ncpr = 0, 0
        c.dt0_after = self.spec.dt0_after
        self.set_clock(c)
        self.Model.nOut, self.Model.nRes = ioc.nOut, ioc.nRes
        self._sync_model()
        return ioc

    def get_clock(self) -> GamClock:
        c = GamClock()
        self._chk(lib().gam_get_clock(self.h, C.byref(c)), "gam_get_clock")
        return c

    def set_clock(self, c: GamClock):
        self._chk(lib().gam_set_clock(self.h, C.byref(c)), "gam_set_clock")

    def _sync_model(self):
        c = self.get_clock()
        M = self.Model
        M.t, M.dt, M.dt0, M.ts, M.status = c.t, c.dt, c.dt0, c.ts, c.status
        return c

    # ---------------------------------------------------------------- stepping
    def stepGamera(self, fused: bool = True) -> int:
        """gamapp.F90:73-94.  fused=False issues AdvanceMHD, CalcDT and
        EnforceBCs as separate C-ABI calls, like the reference's sequence."""
        L = lib()
        if fused:
            st = self._chk(L.gam_step(self.h, 1), "gam_step")
        else:
            self._chk(L.gam_advance_mhd(self.h), "gam_advance_mhd")
            dt, stc = C.c_double(), C.c_int()
            st = self._chk(L.gam_calc_dt(self.h, C.byref(dt), C.byref(stc)), "gam_calc_dt")
            if st == 0:
                self._chk(L.gam_enforce_bcs(self.h), "gam_enforce_bcs")
        self._sync_model()
        return st

    def steps(self, n: int) -> int:
        st = self._chk(lib().gam_step(self.h, n), "gam_step")
        self._sync_model()
        return st

    def AdvanceModel(self, dt: float) -> int:
        """gamtypessub.F90:78-94: step at least once, then until t >= t0 + dt."""
        target = self.Model.t + dt
        st = self.stepGamera()
        while st == 0 and self.Model.t < target:
            st = self.stepGamera()
        return st

    # ---------------------------------------------------------------- boundaries
    def set_bcs(self, bcs):
        """Boundary type per face (INI, OUTI, INJ, OUTJ, INK, OUTK), names of configs.BC_TYPES."""
        t = (C.c_int * 6)(*[configs.BC_TYPES.index(b) for b in bcs])
        self._chk(lib().gam_set_bcs(self.h, t), "gam_set_bcs")

    def get_bcs(self):
        t = (C.c_int * 6)()
        self._chk(lib().gam_get_bcs(self.h, t), "gam_get_bcs")
        return tuple(configs.BC_TYPES[x] for x in t)

    def set_wind(self, rho, p, v, b):
        """The constant solar wind of a "wind" outer i boundary (before set_bcs)."""
        vv, bb = (C.c_double * 3)(*v), (C.c_double * 3)(*b)
        self._chk(lib().gam_set_wind(self.h, rho, p, vv, bb), "gam_set_wind")

    def set_wind_series(self, ws):
        """The solar wind of a "wind" outer i boundary as a time series (a
        configs.WindSeries in code units; before set_bcs): each boundary point
        takes the sample holding at the clock's time less the front's travel
        time from the series' marker (GetWindAt, apps/wind.F90:160-177)."""
        t, rho, p, v, b, xyzBS = ws.arrays()
        d = lambda a: None if a is None else a.ctypes.data_as(_dp)
        self._chk(lib().gam_set_wind_series(self.h, len(t), d(t), d(rho), d(p), d(v), d(b), d(xyzBS),
                                            float(ws.ByC), float(ws.BzC), float(ws.Bx0)), "gam_set_wind_series")

    def set_ring(self, chunks):
        """Ring average about the LFM axis with chunks[r] chunks in ring r+1
        (empty: off); before set_corners, which applies RingGridFix."""
        n = len(chunks)
        t = (C.c_int * max(n, 1))(*chunks)
        self._chk(lib().gam_set_ring(self.h, n, t), "gam_set_ring")

    def get_ring(self):
        n, t = C.c_int(), (C.c_int * 32)()
        self._chk(lib().gam_get_ring(self.h, C.byref(n), t), "gam_get_ring")
        return tuple(t[:n.value])

    # ---------------------------------------------------------------- learned correction
    def state_device(self, which: int = 0):
        """Device addresses (Gas, magFlux, Bxyz) of State (0) or oState (1)."""
        g, m, b = C.c_void_p(), C.c_void_p(), C.c_void_p()
        self._chk(lib().gam_state_device(self.h, which, C.byref(g), C.byref(m), C.byref(b)), "gam_state_device")
        return g.value, m.value, b.value

    def set_correction(self, o_ptr: Optional[int], scale: float = 1.0):
        """Apply the device FP32 correction field o [G][8] in every following step (None: off)."""
        self._chk(lib().gam_set_correction(self.h, o_ptr, float(scale)), "gam_set_correction")

    def set_correction_wait(self, event_ptr: Optional[int]):
        """A hipEvent_t each following step waits for before its first kernel
        that reads the correction (None: no wait)."""
        self._chk(lib().gam_set_correction_wait(self.h, event_ptr), "gam_set_correction_wait")

    def correction_target(self, dGas: np.ndarray, dMag: np.ndarray):
        """Target one-step change (Gas (G,5), magFlux (F,3)) for correction_grad."""
        dGas = np.asfortranarray(dGas, dtype=np.float64)
        dMag = np.asfortranarray(dMag, dtype=np.float64)
        assert dGas.shape == self.G + (5,) and dMag.shape == self.F + (3,)
        self._chk(lib().gam_correction_target(self.h, _p(dGas), _p(dMag)), "gam_correction_target")

    def correction_target_device(self, gas_ptr: int, mag_ptr: int):
        self._chk(lib().gam_correction_target_device(self.h, gas_ptr, mag_ptr), "gam_correction_target_device")

    def correction_grad(self, o_ptr: int, scale: float, dt: float, weights, dout_ptr: int) -> float:
        """Loss of the linearised one-step change of o against the target; dL/do into dout_ptr."""
        w = np.ascontiguousarray(weights, dtype=np.float64)
        assert w.size == 4
        loss = C.c_double()
        self._chk(lib().gam_correction_grad(self.h, o_ptr, float(scale), float(dt), _p(w), dout_ptr, C.byref(loss)),
                  "gam_correction_grad")
        return loss.value

    def get_work(self, which: int) -> np.ndarray:
        """Solver work arrays of the last step: 0 fluxes (G,8,3), 1 EMFs (G,3), 2 Vf (G,3,3)."""
        shape = {0: self.G + (8, 3), 1: self.G + (3,), 2: self.G + (3, 3)}[which]
        out = np.zeros(shape, order="F")
        self._chk(lib().gam_get_work(self.h, which, _p(out)), "gam_get_work")
        return out

    def divb(self) -> float:
        v = C.c_double()
        self._chk(lib().gam_divb(self.h, C.byref(v)), "gam_divb")
        return v.value

    def divb_noalloc(self) -> float:
        """The same total from the entry point that allocates nothing."""
        v = C.c_double()
        self._chk(lib().gam_divb_noalloc(self.h, C.byref(v)), "gam_divb_noalloc")
        return v.value

    def sync(self):
        self._chk(lib().gam_sync(self.h), "gam_sync")

    # ---------------------------------------------------------------- timers
    def set_timing(self, on: bool):
        self._chk(lib().gam_set_timing(self.h, int(on)), "gam_set_timing")

    def stage_id(self, name: str) -> int:
        L = lib()
        for s in range(L.gam_num_stages()):
            if L.gam_stage_name(s).decode() == name:
                return s
        raise KeyError(name)

    def set_tile(self, stage: str, ti: int, tj: int, tk: int, fast: int):
        """Launch brick of a stencil stage (ti*tj*tk == 256); speed only."""
        self._chk(lib().gam_set_tile(self.h, self.stage_id(stage), ti, tj, tk, fast), "gam_set_tile")

    def get_tile(self, stage: str) -> tuple:
        t = (C.c_int * 4)()
        if lib().gam_get_tile(self.h, self.stage_id(stage), t) != 0:
            raise KeyError(f"stage {stage} takes no tile")
        return tuple(t)

    def timings(self) -> dict:
        L = lib()
        out = {}
        for s in range(L.gam_num_stages()):
            ms, n, b = C.c_double(), C.c_int64(), C.c_double()
            self._chk(L.gam_get_timing(self.h, s, C.byref(ms), C.byref(n)), "gam_get_timing")
            self._chk(L.gam_stage_bytes(self.h, s, C.byref(b)), "gam_stage_bytes")
            out[L.gam_stage_name(s).decode()] = dict(ms=ms.value, launches=n.value, bytes=b.value)
        return out
