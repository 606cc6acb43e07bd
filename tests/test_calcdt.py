"""CalcDT's failure branches (reference step.F90:111-156) on the device, against
the oracle: a sudden 10x drop of the timestep and a timestep at or below TINY
(the "disaster", which wins over "bad").  The reference prints, writes a black
box and stops; the library returns the status, keeps the computed dt, and a
fused gam_step stops at the step that failed.  GPU tests; they also run on the
CPU emulation (tests/test_emu.py)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from kaiju_amd import configs
from tests.helpers import app_from_oracle, say


def _case():
    return configs.orszag_tang(16, 4, steps=2)


def _oracle_dt(oc):
    L = oracle.lib()
    dt = L.orc_calc_dt(C.byref(oc.m), C.byref(oc.g), C.byref(oc.state.s))
    return dt, oc.m.status


def _device_dt(app):
    from kaiju_amd.gamera import lib
    dt, st = C.c_double(), C.c_int()
    rc = lib().gam_calc_dt(app.h, C.byref(dt), C.byref(st))
    assert rc >= 0
    return dt.value, st.value


@pytest.mark.gpu
def test_tenfold_drop_is_bad():
    """dtOld / dtMin >= 10 with ts > 0: status 1 ("bad"), dt = the new minimum."""
    say("[calcdt] 10x drop")
    oc = oracle.OracleCase(_case())
    assert oc.step(1) == 0
    oc.m.dt = 12.0 * oc.m.dt               # the previous step was 12x longer than CFL allows now
    with app_from_oracle(oc) as app:
        got = _device_dt(app)
        want = _oracle_dt(oc)
        say(f"   device {got}  oracle {want}")
        assert want[1] == 1 and got[1] == 1
        assert abs(got[0] - want[0]) <= 1e-14 * want[0]
        assert app.get_clock().status == 1
    # just under the threshold it is fine
    oc = oracle.OracleCase(_case())
    assert oc.step(1) == 0
    oc.m.dt = 9.0 * oc.m.dt
    with app_from_oracle(oc) as app:
        got, want = _device_dt(app), _oracle_dt(oc)
        assert want[1] == 0 and got[1] == 0
        assert abs(got[0] - want[0]) <= 1e-14 * want[0]


@pytest.mark.gpu
def test_tiny_timestep_is_a_disaster_and_halts_the_fused_step():
    """One cell moving at 1e14: its CFL step is below TINY (status 2, which
    wins over the 10x drop it also causes); a fused gam_step stops after the
    step whose CalcDT failed, and later steps do nothing until the clock is
    reset."""
    say("[calcdt] dt <= TINY")
    oc = oracle.OracleCase(_case())
    assert oc.step(1) == 0
    nG = oc.spec.nG
    c = (nG + 5, nG + 7, nG + 1)
    for st in (oc.state,):
        st.Gas[c + (1,)] = 1e14 * st.Gas[c + (0,)]
        st.Gas[c + (4,)] += 0.5 * 1e28 * st.Gas[c + (0,)]
    with app_from_oracle(oc) as app:
        got, want = _device_dt(app), _oracle_dt(oc)
        say(f"   device {got}  oracle {want}")
        assert want[1] == 2 and got[1] == 2
        assert want[0] <= 1e-12 and abs(got[0] - want[0]) <= 1e-14 * want[0]
    # through the fused step: the first step's CalcDT sees the fast cell
    oc = oracle.OracleCase(_case())
    assert oc.step(1) == 0
    with app_from_oracle(oc) as app:
        s = app.get_state(0)
        g = s["Gas"]
        g[c + (1,)] = 1e14 * g[c + (0,)]
        g[c + (4,)] += 0.5 * 1e28 * g[c + (0,)]
        app.put_state(0, s["time"], g, s["magFlux"], s["Bxyz"])
        ts0 = app.Model.ts
        assert app.steps(3) == 2
        assert app.Model.ts == ts0 + 1
        s1 = app.get_state(0)
        assert app.steps(2) == 2 and app.Model.ts == ts0 + 1       # halted: nothing moves
        s2 = app.get_state(0)
        assert np.array_equal(s1["Gas"], s2["Gas"])
