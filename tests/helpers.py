"""Helpers shared by the test modules (oracle <-> HIP transfers, error norms)."""
from __future__ import annotations

import sys

import numpy as np

from kaiju_amd import configs


def relerr(a: np.ndarray, b: np.ndarray) -> float:
    """max|a - b| / max|b| (normwise relative error of one array)."""
    d = float(np.max(np.abs(a - b)))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


def relerr_slabs(a: np.ndarray, b: np.ndarray) -> float:
    """Worst normwise relative error over the trailing-axis components."""
    if a.ndim <= 3:
        return relerr(a, b)
    a2 = a.reshape(-1, *a.shape[3:], order="F").reshape(a.shape[0] * a.shape[1] * a.shape[2], -1, order="F")
    b2 = b.reshape(-1, *b.shape[3:], order="F").reshape(b.shape[0] * b.shape[1] * b.shape[2], -1, order="F")
    return max(relerr(a2[:, c], b2[:, c]) for c in range(a2.shape[1]))


def say(*msg):
    print(*msg, flush=True)
    sys.stdout.flush()


def small_cases():
    """Small configurations (seconds on the oracle) covering every code path."""
    ot = configs.orszag_tang(32, 4, steps=8)
    ot8 = configs.orszag_tang(32, 4, steps=8, rmeth="8CENT")
    ot25 = configs.orszag_tang(32, 1, steps=8)
    bw = configs.blast_wave(16, steps=8)
    bwb = configs.blast_wave(16, steps=8, doBoris=True, Ca=6.0)
    hyd = configs.blast_wave(16, steps=8, doMHD=False)
    nolfm = configs.orszag_tang(24, 4, steps=6, doLFMLim=False)
    thin = configs.blast_wave((24, 2, 3), steps=6)     # owned extents below the ghost depth
    thin.bounds = (0.0, 1.0, 0.4, 0.6, 0.3, 0.7)
    # background field split: a dipole outside the box plus a uniform part
    dip = configs.blast_wave(16, steps=8)
    dip.background = DIPOLE
    dipb = configs.blast_wave(16, steps=8, doBoris=True, Ca=6.0)
    dipb.background = DIPOLE
    # non-periodic boundaries: every type on every axis once (Cartesian grid, as
    # the reflecting walls require)
    bcA = configs.orszag_tang(24, 4, steps=6)
    bcA.bcs = ("zeroexten", "zeroexten", "reflect", "reflect", "zerograd", "zerograd")
    bcB = configs.orszag_tang(24, 4, steps=6, doBoris=True, Ca=3.0)
    bcB.bcs = ("zerograd", "zerograd", "zeroexten", "zeroexten", "reflect", "reflect")
    bcC = configs.orszag_tang(24, 4, steps=6)
    bcC.bcs = ("reflect", "reflect", "zerograd", "zerograd", "zeroexten", "zeroexten")
    bcS = configs.orszag_tang(24, 4, steps=6)
    bcS.bcs = ("sphere", "sphere", "periodic", "periodic", "periodic", "periodic")
    # CalcDT repairs (StateCPR): a weak-field blast wave whose dt stalls, so a tight
    # limCPR triggers them from the third step on
    cpr = configs.blast_wave(16, steps=8, B0=0.5, doCPR=True, limCPR=1.05)
    # Boris with Ca below the flow speed: |v| reaches 1.5 Ca (the nuke of
    # stress.F90:399-414 fires) and alpha = |B|^2/(rho Ca^2) ~ 1
    nuke = configs.orszag_tang(32, 4, steps=8, doBoris=True, Ca=0.6)
    # cylindrical shell with the through-pole inner BC (cylpole) and zero gradient outside
    cyl = configs.cylinder_blast((12, 16, 8), steps=6)
    # odd i extents: the paired j/k interfaces end on an unpaired one, and every
    # other row of the cell arrays starts on an odd index (8-byte aligned pairs)
    odd = configs.blast_wave((15, 12, 10), steps=6)
    for s, n in ((ot, "ot32x32x4"), (ot8, "ot32x32x4_8cent"), (ot25, "ot32x32x1_25d"), (bw, "bw16_warp"),
                 (bwb, "bw16_warp_boris"), (hyd, "bw16_hydro"), (nolfm, "ot24x24x4_nolfmlim"),
                 (thin, "bw24x2x3_thin"), (dip, "bw16_warp_dipole"), (dipb, "bw16_warp_dipole_boris"),
                 (bcA, "ot24_bc_exten_reflect_grad"), (bcB, "ot24_bc_grad_exten_reflect_boris"),
                 (bcC, "ot24_bc_reflect_grad_exten"), (bcS, "ot24_bc_sphere"), (cpr, "bw16_warp_cpr"),
                 (nuke, "ot32x32x4_boris_nuke"), (cyl, "cyl12x16x8_cylpole"), (odd, "bw15x12x10_odd")):
        s.name = n
    return [ot, ot8, ot25, bw, bwb, hyd, nolfm, thin, dip, dipb, bcA, bcB, bcC, bcS, cpr, nuke, cyl, odd]


# B0 = (0.5, -1, 0.25) + dipole of moment (0,0,8) at (-0.7, 0.45, 0.6): |B0| ~ 1-20 in the unit cube
DIPOLE = dict(M0=8.0, xc=(-0.7, 0.45, 0.6), bu=(0.5, -1.0, 0.25))


def app_from_oracle(oc):
    """A GamApp holding exactly the oracle's geometry, State, oState and clock."""
    from kaiju_amd.gamera import GamApp
    app = GamApp(oc.spec)
    app.set_grid(oc.grid)
    if oc.spec.background is not None:
        app.set_background(oc.bg)
    st, os_ = oc.state, oc.ostate
    app.put_state(0, st.time, st.Gas, st.magFlux, st.Bxyz)
    app.put_state(1, os_.time, os_.Gas, os_.magFlux, os_.Bxyz)
    c = app.get_clock()
    m = oc.m
    c.t, c.dt, c.dt0, c.ts = m.t, m.dt, m.dt0, m.ts
    c.stateTime, c.oStateTime = st.time, os_.time
    c.dt0_after = oc.spec.dt0_after
    c.status = 0
    app.set_clock(c)
    app._sync_model()
    return app


def field_err(a: np.ndarray, b: np.ndarray, groups) -> float:
    """Worst normwise relative error over physical fields.  ``groups`` lists
    component index tuples of one field; a vector field is scaled by the
    largest magnitude of any of its components, so a component that is zero
    up to rounding (e.g. z-momentum of a 2D flow) is judged against the field,
    not against its own noise."""
    worst = 0.0
    for grp in groups:
        d = max(float(np.max(np.abs(a[..., c] - b[..., c]))) for c in grp)
        s = max(float(np.max(np.abs(b[..., c]))) for c in grp)
        worst = max(worst, d / s if s > 0 else d)
    return worst


GAS_FIELDS = ((0,), (1, 2, 3), (4,))   # density, momentum, energy
VEC_FIELD = ((0, 1, 2),)


def compare_state(app, oc, which: int = 0):
    """Worst relative error of Gas, magFlux, Bxyz between the GPU and the oracle."""
    s = app.get_state(which)
    o = oc.state if which == 0 else oc.ostate
    return dict(Gas=field_err(s["Gas"], o.Gas, GAS_FIELDS), magFlux=field_err(s["magFlux"], o.magFlux, VEC_FIELD),
                Bxyz=field_err(s["Bxyz"], o.Bxyz, VEC_FIELD), time=abs(s["time"] - o.time))


def on_emulation() -> bool:
    """True when the binding points at the CPU emulation build (tests/emu)."""
    import os
    from kaiju_amd import gamera
    return "emu" in os.path.basename(gamera.LIB_PATH)


def DevF32(a: np.ndarray):
    """An FP32 device copy of a (plain host memory on the CPU emulation)."""
    from kaiju_amd.gamera import DeviceArray
    return DeviceArray.of(a, np.float32)
