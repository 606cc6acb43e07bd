"""HIP step vs the oracle on identical inputs (FP64 path).

Bar: <= 1e-12 normwise relative error per array after one step (the
reference's parity target), the dt identical to 1e-12, and div.B unchanged
to rounding.  Geometry and initial-condition kernels are compared at 1e-13.
Every case prints its name before it touches the GPU.
"""
import numpy as np
import pytest

import oracle
from kaiju_amd import configs
from tests.helpers import (GAS_FIELDS, VEC_FIELD, app_from_oracle, compare_state, field_err, relerr_slabs,
                           say, small_cases)

pytestmark = pytest.mark.gpu

CASES = small_cases()
TOL_STEP = 1e-12
TOL_SETUP = 1e-13


@pytest.fixture(scope="module", params=CASES, ids=[c.name for c in CASES])
def case(request):
    oracle.build()
    return request.param


def test_metric_from_corners(case):
    """gam_set_corners (Corners2Grid + ebGeom on the GPU) vs orc_corners2grid."""
    say(f"[metric] {case.name}")
    from kaiju_amd.gamera import GamApp
    oc = oracle.OracleCase(case)
    with GamApp(case) as app:
        app.set_corners(oc.grid["xyz"])
        g = app.get_grid()
    for k, ref in oc.grid.items():
        err = relerr_slabs(g[k], ref)
        assert err <= TOL_SETUP, f"{k}: {err:.3e}"


def test_initial_condition(case):
    """GPU initial condition on the oracle's geometry vs the oracle's."""
    say(f"[ic] {case.name}")
    oc = oracle.OracleCase(case, setup=False)
    oc.gen_corners()
    oc.metric()
    from kaiju_amd.gamera import GamApp
    with GamApp(case) as app:
        app.set_grid(oc.grid)
        app.init_ic()
        s = app.get_state(0)
    oc.ic()
    eg = field_err(s["Gas"], oc.state.Gas, GAS_FIELDS)
    em = field_err(s["magFlux"], oc.state.magFlux, VEC_FIELD)
    assert eg <= TOL_SETUP and em <= TOL_SETUP, (eg, em)


def test_hatch(case):
    """gam_hatch (first Bxyz, BCs, oState, first dt) vs orc_hatch_finish."""
    say(f"[hatch] {case.name}")
    oc = oracle.OracleCase(case, setup=False)
    oc.gen_corners()
    oc.metric()
    oc.ic()
    from kaiju_amd.gamera import GamApp
    with GamApp(case) as app:
        app.set_grid(oc.grid)
        app.put_state(0, 0.0, oc.state.Gas, oc.state.magFlux, oc.state.Bxyz)
        app.hatch()
        oc.hatch_finish()
        e0 = compare_state(app, oc, 0)
        e1 = compare_state(app, oc, 1)
        assert abs(app.Model.dt - oc.m.dt) <= 1e-15 * oc.m.dt
    for e in (e0, e1):
        assert max(e["Gas"], e["magFlux"], e["Bxyz"]) == 0.0, e


def test_one_step(case):
    """One stepGamera from identical State/oState: every array within 1e-12."""
    say(f"[step1] {case.name}")
    oc = oracle.OracleCase(case)
    oc.step(1)   # leave the start-up trap behind: compare a generic step
    with app_from_oracle(oc) as app:
        st = app.stepGamera()
        assert st == 0
        oc.step(1)
        e = compare_state(app, oc, 0)
        eo = compare_state(app, oc, 1)
        dt_err = abs(app.Model.dt - oc.m.dt) / oc.m.dt
        say(f"   state {e}  ostate {eo}  dt {dt_err:.2e}")
        assert e["Gas"] <= TOL_STEP and e["magFlux"] <= TOL_STEP and e["Bxyz"] <= TOL_STEP, e
        assert eo["Gas"] == 0.0 and eo["magFlux"] == 0.0
        assert dt_err <= TOL_STEP
        assert app.Model.ts == oc.m.ts and abs(app.Model.t - oc.m.t) <= 1e-14 * max(oc.m.t, 1.0)


def test_first_step_trap(case):
    """The very first step after hatch (oState%time trap, mhdgroup.F90:25-28)."""
    say(f"[step0] {case.name}")
    oc = oracle.OracleCase(case)
    with app_from_oracle(oc) as app:
        st = app.stepGamera()
        oc.step(1)
        e = compare_state(app, oc, 0)
        say(f"   {e}")
        assert st == 0
        assert max(e["Gas"], e["magFlux"], e["Bxyz"]) <= TOL_STEP, e


def test_predictor_state(case):
    """The half-step state produced inside the step (Predictor, mhdgroup.F90:445-525)."""
    say(f"[predictor] {case.name}")
    oc = oracle.OracleCase(case)
    oc.step(1)
    with app_from_oracle(oc) as app:
        app.stepGamera()
        oc.step(1)
        hf = app.get_state(2)
        G, F = oc.G, oc.F
        ref_gas = oc.solver_array(0, G + (5,))
        ref_mag = oc.solver_array(1, F + (3,))
        ref_b = oc.solver_array(2, G + (3,))
        e = (field_err(hf["Gas"], ref_gas, GAS_FIELDS), field_err(hf["magFlux"], ref_mag, VEC_FIELD),
             field_err(hf["Bxyz"], ref_b, VEC_FIELD))
        assert max(e) <= TOL_STEP, e


def test_separate_calls_match_fused(case):
    """AdvanceMHD / CalcDT / EnforceBCs as separate C-ABI calls == fused gam_step."""
    say(f"[split] {case.name}")
    oc = oracle.OracleCase(case)
    oc.step(1)
    with app_from_oracle(oc) as a1, app_from_oracle(oc) as a2:
        a1.stepGamera(fused=True)
        a2.stepGamera(fused=False)
        s1, s2 = a1.get_state(0), a2.get_state(0)
        for k in ("Gas", "magFlux", "Bxyz"):
            assert np.array_equal(s1[k], s2[k]), k
        assert a1.Model.dt == a2.Model.dt and a1.Model.t == a2.Model.t


def test_multistep_drift_and_divb(case):
    """case.steps steps: drift stays small, div.B stays at rounding level."""
    say(f"[multi] {case.name}")
    oc = oracle.OracleCase(case)
    divb0 = oc.divb()
    with app_from_oracle(oc) as app:
        st = app.steps(case.steps)
        assert st == 0
        oc.step(case.steps)
        e = compare_state(app, oc, 0)
        say(f"   after {case.steps}: {e}")
        assert max(e["Gas"], e["magFlux"], e["Bxyz"]) <= 1e-10, e
        if case.model.doMHD:
            scale = float(np.abs(oc.state.magFlux).max()) * case.zones
            assert abs(app.divb() - divb0) <= 1e-13 * scale


TILE_SETS = [
    # every stencil stage on a row-shaped brick (the plain 1-D launch order)
    {st: (256, 1, 1, 0) for st in ("bflux_half", "bflux_post", "mom2face", "emf_i", "emf_j", "emf_k", "flux_i",
                                   "flux_j", "flux_k", "update")},
    # linear blocks (256 consecutive points) everywhere, hardware block order
    {st: (256, 1, 1, 7) for st in ("bflux_half", "bflux_post", "mom2face", "emf_i", "emf_j", "emf_k", "flux_i",
                                   "flux_j", "flux_k", "update")},
    # bricks that leave partial tiles on every edge of the small boxes
    {"bflux_half": (8, 8, 4, 1), "bflux_post": (4, 8, 8, 2), "mom2face": (32, 2, 4, 0), "emf_i": (8, 4, 8, 1),
     "emf_j": (4, 8, 8, 0), "emf_k": (2, 16, 8, 2), "flux_i": (2, 16, 8, 2), "flux_j": (4, 4, 16, 0),
     "flux_k": (1, 16, 16, 5), "update": (1, 16, 16, 6)},
]


@pytest.mark.parametrize("tset", range(len(TILE_SETS)))
def test_launch_tiles_change_no_bits(case, tset):
    """The launch brick of the stencil stages only reorders threads: two steps
    with other tiles give bit-identical State to the default tiles."""
    say(f"[tiles] {case.name} set {tset}")
    oc = oracle.OracleCase(case)
    with app_from_oracle(oc) as a1, app_from_oracle(oc) as a2:
        for st, t in TILE_SETS[tset].items():
            a2.set_tile(st, *t)
            assert a2.get_tile(st) == t
        assert a1.steps(2) == 0 and a2.steps(2) == 0
        s1, s2 = a1.get_state(0), a2.get_state(0)
        for k in ("Gas", "magFlux", "Bxyz"):
            assert np.array_equal(s1[k], s2[k]), k
        assert a1.Model.dt == a2.Model.dt


def test_set_tile_rejects_bad_shapes(case):
    say(f"[tiles-bad] {case.name}")
    from kaiju_amd.gamera import GamApp, HipStepError
    with GamApp(case) as app:
        with pytest.raises(HipStepError):
            app.set_tile("emf_i", 16, 4, 2, 0)    # 128 threads
        with pytest.raises(HipStepError):
            app.set_tile("flux_i", 64, 4, 1, 8)   # no mode 8
        with pytest.raises(HipStepError):
            app.set_tile("ct_update", 64, 4, 1, 0)   # not a tiled stage


def test_background_from_gpu(case):
    """gam_background_dipole (AddB0 on the GPU) vs orc_add_b0 on the same grid."""
    if case.background is None:
        pytest.skip("no background field in this case")
    say(f"[addb0] {case.name}")
    from kaiju_amd.gamera import GamApp
    oc = oracle.OracleCase(case)
    b = case.background
    with GamApp(case) as app:
        app.set_grid(oc.grid)
        app.background_dipole(b["M0"], b["xc"], b["bu"])
        bg = app.get_background()
    for k, ref in oc.bg.items():
        scale = max(float(np.abs(ref).max()), 1e-300)
        err = float(np.abs(bg[k] - ref).max()) / scale
        say(f"   {k}: {err:.2e}")
        # dpB0 is a difference of O(B0^2) face stresses divided by the volume
        tol = 1e-9 if k == "dpB0" else TOL_SETUP
        assert err <= tol, (k, err)
