"""ctypes front end of the oracle: the CPU restatement of the GAMERA step.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s CPU-baseline leg, and only there as the checker or the timed CPU
baseline.  The product path in ``kaiju_amd`` never imports this package.

Arrays are numpy arrays in Fortran order with the reference's shapes, e.g.
``Gas`` is ``(Ni, Nj, Nk, 5)`` and ``magFlux`` is ``(Ni+1, Nj+1, Nk+1, 3)``,
so index ``[i - isg, j - jsg, k - ksg, v]`` is the reference's ``Gas(i,j,k,v)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)


class OrcModel(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("gamma", "CFL", "pdmb", "Vd0", "cHogH", "cHogM", "Ca", "dFloor", "pFloor",
                 "t", "tFin", "dt", "dt0", "limDT0", "limCPR")] + \
               [(n, C.c_int) for n in
                ("ts", "rmeth", "doLFMLim", "doBoris", "doMHD", "do25D", "doCPR",
                 "fixedTimestep", "status", "doBackground")] + \
               [("bc", C.c_int * 6), ("ncpr", C.c_int)] + \
               [("wind_rho", C.c_double), ("wind_p", C.c_double), ("wind_v", C.c_double * 3), ("wind_b", C.c_double * 3),
                ("wind_nt", C.c_int), ("wind_ser", C.POINTER(C.c_double))] + \
               [(n, C.c_double) for n in ("wind_tMin", "wind_tMax", "wind_dtW", "wind_ByC", "wind_BzC", "wind_Bx0")]


class OrcGrid(C.Structure):
    _fields_ = [(n, C.c_int) for n in
                ("Nip", "Njp", "Nkp", "nG", "Ni", "Nj", "Nk", "is_", "ie", "js", "je", "ks", "ke",
                 "isg", "ieg", "jsg", "jeg", "ksg", "keg",
                 "isDT", "ieDT", "jsDT", "jeDT", "ksDT", "keDT")] + \
               [("nc", C.c_size_t), ("nf", C.c_size_t)] + \
               [(n, _dp) for n in ("xyz", "xyzcc", "xfc", "face", "Tf", "edge", "Te", "Teb",
                                   "volume", "di", "dj", "dk", "fcB0", "edgB0", "B0", "bFlux0", "dpB0")] + \
               [("ring", C.c_int), ("nring", C.c_int), ("nch", C.c_int * 32)]


class OrcState(C.Structure):
    _fields_ = [("time", C.c_double)] + [(n, _dp) for n in ("Gas", "magFlux", "Bxyz", "Efld")]


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "gamera_oracle.c")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        M, G, S = C.POINTER(OrcModel), C.POINTER(OrcGrid), C.POINTER(OrcState)
        sig = {
            "orc_grid_derive": (None, [G]),
            "orc_gen_corners_cart": (None, [G, _dp, C.c_int]),
            "orc_gen_corners_warp3": (None, [G, _dp, C.c_double]),
            "orc_gen_corners_cyl": (None, [G, _dp]),
            "orc_gen_corners_refwarp": (None, [G, _dp, C.c_double, C.c_double, C.c_int, C.c_int]),
            "orc_corners2grid": (None, [G]),
            "orc_ic_ot": (None, [M, G, S, C.c_double, C.c_double, C.c_double, C.c_double]),
            "orc_ic_bw": (None, [M, G, S] + [C.c_double] * 9),
            "orc_solver_new": (C.c_void_p, [G]),
            "orc_solver_free": (None, [C.c_void_p]),
            "orc_solver_array": (_dp, [C.c_void_p, C.c_int]),
            "orc_hatch_finish": (None, [M, G, S, S]),
            "orc_advance_mhd": (None, [M, G, S, S, C.c_void_p, C.c_double]),
            "orc_calc_dt": (C.c_double, [M, G, S]),
            "orc_enforce_bcs": (None, [M, G, S]),
            "orc_bflux2fld": (None, [G, _dp, _dp]),
            "orc_step": (None, [M, G, S, S, C.c_void_p]),
            "orc_run_steps": (C.c_int, [M, G, S, S, C.c_void_p, C.c_int, C.c_int]),
            "orc_divb": (C.c_double, [G, S]),
            "orc_predictor": (None, [M, G, S, S, S, C.c_double]),
            "orc_calc_efield": (None, [M, G, S, S, _dp, _dp]),
            "orc_calc_stress": (None, [M, G, S, S, _dp, _dp, _dp, _dp]),
            "orc_cell_c2p": (None, [M, _dp, _dp]),
            "orc_cell_p2c": (None, [M, _dp, _dp]),
            "orc_pdm": (C.c_double, [C.c_double] * 5),
            "orc_pdmlr": (None, [C.c_double, _dp, C.c_double, _dp, _dp]),
            "orc_up7": (C.c_double, [_dp]),
            "orc_central8": (C.c_double, [_dp]),
            "orc_gaskinflux": (None, [M, C.c_int, _dp, _dp, _dp, _dp]),
            "orc_magkinflux": (None, [M, C.c_int, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]),
            "orc_cell_boris": (None, [M, _dp, _dp, _dp, _dp, _dp, _dp, C.c_double, C.c_double]),
            "orc_cell_maxwell": (None, [M, _dp, _dp, C.c_double]),
            "orc_cell_dt": (C.c_double, [M, G, S, C.c_int, C.c_int, C.c_int]),
            "orc_magkinflux_b0": (None, [M, C.c_int, _dp, _dp, C.c_double, _dp, _dp, C.c_double, _dp, _dp, _dp]),
            "orc_add_b0": (None, [M, G, C.c_double, _dp, _dp]),
            "orc_ring_average": (None, [M, G, S, _dp]),
            "orc_reconstruct_ring": (None, [_dp, C.c_int, C.c_int]),
            "orc_ppm_intlr": (None, [C.c_double] * 5 + [_dp, _dp]),
            "orc_lfm_map": (None, [G, C.c_int, C.c_int, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 3),
            "orc_wgt_wind": (C.c_double, [M, _dp]),
            "orc_wind_at": (None, [M, _dp, C.c_double, _dp]),
            "orc_omp_threads": (C.c_int, []),
            "orc_nuke_count": (C.c_long, [C.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def omp_threads() -> int:
    """OpenMP threads of the oracle's parallel loops."""
    return int(lib().orc_omp_threads())


def ptr(a: np.ndarray):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1)
    return a.ctypes.data_as(_dp)


def model_struct(spec_model, do25D: bool) -> OrcModel:
    m = OrcModel()
    for k in ("gamma", "CFL", "pdmb", "Vd0", "cHogH", "cHogM", "Ca", "dFloor", "pFloor",
              "tFin", "limDT0", "limCPR"):
        setattr(m, k, float(getattr(spec_model, k)))
    m.rmeth = 1 if spec_model.rmeth.upper() == "7UP" else 0
    m.doLFMLim = int(spec_model.doLFMLim)
    m.doBoris = int(spec_model.doBoris)
    m.doMHD = int(spec_model.doMHD)
    m.doCPR = int(spec_model.doCPR)
    m.do25D = int(do25D)
    m.t = 0.0
    m.ts = 0
    return m


class OracleState:
    """State_T with numpy storage."""

    def __init__(self, G, F):
        self.Gas = np.zeros(G + (5,), order="F")
        self.magFlux = np.zeros(F + (3,), order="F")
        self.Bxyz = np.zeros(G + (3,), order="F")
        self.Efld = np.zeros(G + (3,), order="F")
        self.s = OrcState()
        self.s.time = 0.0
        self.s.Gas, self.s.magFlux, self.s.Bxyz, self.s.Efld = (
            ptr(self.Gas), ptr(self.magFlux), ptr(self.Bxyz), ptr(self.Efld))

    @property
    def time(self):
        return self.s.time

    def copy_from(self, other: "OracleState"):
        self.Gas[...] = other.Gas
        self.magFlux[...] = other.magFlux
        self.Bxyz[...] = other.Bxyz
        self.Efld[...] = other.Efld
        self.s.time = other.s.time


class OracleCase:
    """A whole reference-style app (Model, Grid, State, oState, Solver) on the CPU."""

    GRID_ARRAYS = ("xyz", "xyzcc", "xfc", "face", "Tf", "edge", "Te", "Teb", "volume", "di", "dj", "dk")

    def __init__(self, spec, setup: bool = True):
        L = lib()
        self.spec = spec
        g = OrcGrid()
        g.Nip, g.Njp, g.Nkp = spec.N
        g.nG = spec.nG
        L.orc_grid_derive(C.byref(g))
        ring = getattr(spec, "ring", None)
        if ring:
            g.ring, g.nring = 1, len(ring)
            for r, n in enumerate(ring):
                g.nch[r] = n
        self.g = g
        G = (g.Ni, g.Nj, g.Nk)
        F = (g.Ni + 1, g.Nj + 1, g.Nk + 1)
        self.G, self.F = G, F
        shapes = dict(xyz=F + (3,), xyzcc=G + (3,), xfc=F + (3, 3), face=F + (3,), Tf=F + (9, 3),
                      edge=F + (3,), Te=F + (9, 3), Teb=F + (4, 3), volume=G, di=G, dj=G, dk=G)
        self.grid = {}
        for k, shp in shapes.items():
            a = np.zeros(shp, order="F")
            self.grid[k] = a
            setattr(g, k, ptr(a))
        # background field arrays (G-sized, reference gamtypes.F90:231-238)
        bshapes = dict(fcB0=G + (3, 3), edgB0=G + (2, 3), B0=G + (3,), bFlux0=G + (3,), dpB0=G + (3,))
        self.bg = {}
        for k, shp in bshapes.items():
            a = np.zeros(shp, order="F")
            self.bg[k] = a
            setattr(g, k, ptr(a))
        self.m = model_struct(spec.model, spec.do25D)
        self.m.doBackground = int(spec.background is not None)
        from kaiju_amd.configs import BC_TYPES
        for f, b in enumerate(getattr(spec, "bcs", ("periodic",) * 6)):
            self.m.bc[f] = BC_TYPES.index(b)
        wind = getattr(spec, "wind", None)
        if wind is not None:
            rho, p, v, b = wind
            self.m.wind_rho, self.m.wind_p = rho, p
            for x in range(3):
                self.m.wind_v[x], self.m.wind_b[x] = v[x], b[x]
        if getattr(spec, "wind_series", None) is not None:
            self.set_wind_series(spec.wind_series)
        self.state = OracleState(G, F)
        self.ostate = OracleState(G, F)
        self.solver = L.orc_solver_new(C.byref(g))
        if setup:
            self.setup()

    def set_wind_series(self, ws):
        """The wind as a time series (a configs.WindSeries): readWind's table
        (apps/wind.F90:667-702) with dtW = the smallest sample spacing."""
        t, rho, p, v, b, xyzBS = ws.arrays()
        nt = len(t)
        tab = np.zeros((nt, 12))
        tab[:, 0], tab[:, 1], tab[:, 2], tab[:, 3:6], tab[:, 6:9] = t, rho, p, v, b
        if xyzBS is not None:
            tab[:, 9:12] = xyzBS
        self._wind_tab = np.ascontiguousarray(tab).ravel()    # row n at [12 n, 12 n + 12)
        m = self.m
        m.wind_nt, m.wind_ser = nt, ptr(self._wind_tab)
        m.wind_tMin, m.wind_tMax, m.wind_dtW = float(t.min()), float(t.max()), float(np.diff(t).min())
        m.wind_ByC, m.wind_BzC, m.wind_Bx0 = ws.ByC, ws.BzC, ws.Bx0

    def wind_at(self, xyz, t):
        """GetWindAt after RefreshWind at t: (rho, p, v[3], B[3]) at point xyz."""
        q = (C.c_double * 8)()
        lib().orc_wind_at(C.byref(self.m), (C.c_double * 3)(*xyz), t, q)
        return np.array(q[:])

    def __del__(self):
        try:
            if getattr(self, "solver", None):
                lib().orc_solver_free(self.solver)
                self.solver = None
        except Exception:
            pass

    # -- setup producers (reference init.F90 Hatch) --
    def gen_corners(self):
        L, sp = lib(), self.spec
        bds = (C.c_double * 6)(*sp.bounds)
        if sp.grid == "cart":
            L.orc_gen_corners_cart(C.byref(self.g), bds, int(sp.do25D))
        elif sp.grid == "warp3":
            L.orc_gen_corners_warp3(C.byref(self.g), bds, sp.warp_amp)
        elif sp.grid == "refwarp":
            L.orc_gen_corners_refwarp(C.byref(self.g), bds, sp.warp_amp, 1.0, 1, int(sp.do25D))
        elif sp.grid == "cyl":
            L.orc_gen_corners_cyl(C.byref(self.g), bds)
        else:
            raise ValueError(sp.grid)

    def metric(self):
        lib().orc_corners2grid(C.byref(self.g))

    def ic(self):
        L, sp = lib(), self.spec
        p = sp.ic_params
        if sp.ic == "ot":
            L.orc_ic_ot(C.byref(self.m), C.byref(self.g), C.byref(self.state.s),
                        p["rho0"], p["p0"], p["b0"], p["kscale"])
        elif sp.ic == "bw":
            L.orc_ic_bw(C.byref(self.m), C.byref(self.g), C.byref(self.state.s),
                        p["x0"], p["y0"], p["z0"], p["D0"], p["P0"], p["pRat"], p["dRat"], p["rC"], p["B0"])
        elif sp.ic != "none":
            raise ValueError(sp.ic)

    def add_b0(self):
        """background.F90:20-167 AddB0 for the spec's analytic background."""
        b = self.spec.background
        if b is None:
            return
        xc = (C.c_double * 3)(*b.get("xc", (0.0, 0.0, 0.0)))
        bu = (C.c_double * 3)(*b.get("bu", (0.0, 0.0, 0.0)))
        lib().orc_add_b0(C.byref(self.m), C.byref(self.g), float(b.get("M0", 0.0)), xc, bu)

    def setup(self):
        """Reference Hatch: corners, Corners2Grid, background field, IC, first BCs and dt."""
        self.gen_corners()
        self.metric()
        self.add_b0()
        self.ic()
        self.hatch_finish()

    def hatch_finish(self):
        lib().orc_hatch_finish(C.byref(self.m), C.byref(self.g), C.byref(self.state.s), C.byref(self.ostate.s))

    # -- stepping --
    def step(self, n: int = 1) -> int:
        return lib().orc_run_steps(C.byref(self.m), C.byref(self.g), C.byref(self.state.s),
                                   C.byref(self.ostate.s), self.solver, n, self.spec.dt0_after)

    def timed_steps(self, n: int) -> float:
        t0 = time.perf_counter()
        st = self.step(n)
        if st:
            raise RuntimeError(f"oracle step status {st}")
        return time.perf_counter() - t0

    def solver_array(self, which: int, shape):
        p = lib().orc_solver_array(self.solver, which)
        n = int(np.prod(shape))
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(shape, order="F")

    def divb(self) -> float:
        return lib().orc_divb(C.byref(self.g), C.byref(self.state.s))

    @property
    def model(self) -> OrcModel:
        return self.m
