"""Batched ensemble under one captured graph, and the brick autotuner.

CPU: the seeded member draw.  GPU (and CPU emulation): replaying the captured
graph gives, at every replay, bit-for-bit the state, clock and status of the
same blocks stepped eagerly with gam_step; a block that halts on a CalcDT
status inside the graph stops exactly where gam_step stops; the autotuner
restores the block it tuned and changes no result.  GPU: the ensemble workload
itself (32 seeded blocks, edges 64-160, FP32 bulk) steps under the graph with
every block conserving and keeping div.B, and a seeded draw of oracle-sized
FP64 blocks under the graph agrees with the oracle block by block.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from kaiju_amd import autotune, configs
from kaiju_amd.ensemble import Ensemble, draw_members
from tests.helpers import app_from_oracle, compare_state, say


def test_members_are_seeded_and_in_range():
    a = draw_members(32, 31337)
    b = draw_members(32, 31337)
    c = draw_members(32, 31338)
    assert [s.name for s in a] == [s.name for s in b]
    assert [s.N for s in a] != [s.N for s in c]
    kinds = {s.ic for s in a}
    assert kinds == {"ot", "bw"}
    for s in a:
        assert all(64 <= n <= 160 for n in s.N) and s.N[0] % 8 == 0
        assert s.model.precision == 1
        if s.ic == "bw":
            assert 1.0 <= s.ic_params["B0"] <= 20.0
            assert 10.0 <= s.ic_params["pRat"] <= 1000.0
    zones = sum(s.zones for s in a)
    assert 32 * 64 ** 3 < zones < 32 * 160 ** 3


def small_members(precision):
    """Four small uneven blocks (two of each kind), oracle-sized."""
    out = []
    for q, N in enumerate([(16, 20, 12), (24, 16, 18), (16, 18, 20), (8, 24, 1)]):
        if q % 2 == 0:
            s = configs.blast_wave(N, B0=3.0 + q, pin=20.0 + 10 * q, pout=0.1, precision=precision)
        else:
            s = configs.orszag_tang(N, precision=precision)
        out.append(dataclasses.replace(s, name=f"s{q}_{s.name}"))
    return out


def eager_twins(specs):
    from kaiju_amd.gamera import GamApp
    apps = []
    for s in specs:
        a = GamApp(s)
        a.InitModel()
        apps.append(a)
    return apps


def same(a, b):
    sa, sb = a.get_state(0), b.get_state(0)
    for k in ("Gas", "magFlux", "Bxyz"):
        assert np.array_equal(sa[k], sb[k]), k
    ca, cb = a.get_clock(), b.get_clock()
    assert (ca.t, ca.dt, ca.ts, ca.status, ca.dt0) == (cb.t, cb.dt, cb.ts, cb.status, cb.dt0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_graph_replays_match_eager_steps(precision):
    specs = small_members(precision)
    say(f"[ensemble] graph vs eager, precision {precision}, {[s.N for s in specs]}")
    twins = eager_twins(specs)
    with Ensemble(specs, tune=False, nsteps=2) as ens:
        for rep in range(3):
            ens.launch(1)
            assert ens.sync() == 0
            for t in twins:
                assert t.steps(2) == 0
            # eager work between replays: reads, a sync, a foreign launch
            for a, t in zip(ens.apps, twins):
                same(a, t)
                a.divb()
            say(f"   replay {rep} identical")
    for t in twins:
        t.close()


@pytest.mark.gpu
def test_graph_block_halts_where_gam_step_halts():
    specs = small_members(1)[:2]
    say("[ensemble] halt inside the graph")
    twins = eager_twins(specs)
    with Ensemble(specs, tune=False, nsteps=2) as ens:
        # dt0 far above dt: CalcDT's dt/dt0 <= limDT0 check fails after step 1
        for x in (ens.apps[0], twins[0]):
            c = x.get_clock()
            c.dt0 = c.dt * 1e6
            x.set_clock(c)
        st = ens.run(4)
        assert st == 1
        assert twins[0].steps(4) == 1 and twins[1].steps(4) == 0
        assert ens.apps[0].Model.ts == 1 and ens.apps[1].Model.ts == 4
        for a, t in zip(ens.apps, twins):
            same(a, t)
    for t in twins:
        t.close()


@pytest.mark.gpu
def test_autotune_restores_state_and_changes_no_result():
    spec = small_members(1)[0]
    say(f"[autotune] {spec.N}")
    a, b = eager_twins([spec, spec])
    s_before = a.get_state(0)
    cache = {}
    best = autotune.tune(a, cache=cache)
    s_after = a.get_state(0)
    for k in ("Gas", "magFlux", "Bxyz"):
        assert np.array_equal(s_before[k], s_after[k]), k
    assert a.Model.ts == b.Model.ts == 0
    for st, t in best.items():
        assert a.get_tile(st) == t and t in autotune.CANDIDATES[st]
    assert autotune.shape_key(a) in cache
    assert a.steps(3) == 0 and b.steps(3) == 0
    same(a, b)
    # a cached entry is applied without timing
    c = eager_twins([spec])[0]
    assert autotune.tune(c, cache=cache) == best
    for x in (a, b, c):
        x.close()


@pytest.mark.gpu
def test_graph_after_an_odd_eager_prestep():
    """A member stepped once eagerly before the capture (State in the second
    buffer set) replays in step with its twin, and a later odd eager step on it
    makes the next launch refuse instead of stepping swapped buffers."""
    from kaiju_amd.gamera import HipStepError
    specs = small_members(1)[:2]
    say("[ensemble] capture after an eager step")
    twins = eager_twins(specs)
    ens = Ensemble.__new__(Ensemble)
    ens.specs, ens.nsteps, ens.apps, ens.tiles, ens.graph = specs, 2, eager_twins(specs), {}, None
    try:
        for a, t in zip(ens.apps, twins):
            assert a.steps(1) == 0 and t.steps(1) == 0
        ens._capture()
        for rep in range(2):
            assert ens.run(2) == 0
            for t in twins:
                assert t.steps(2) == 0
            for a, t in zip(ens.apps, twins):
                same(a, t)
        assert ens.apps[0].steps(1) == 0
        with pytest.raises(HipStepError):
            ens.launch(1)
        assert ens.apps[0].steps(1) == 0   # even again: the graph runs
        assert ens.run(2) == 0
        assert twins[0].steps(4) == 0 and twins[1].steps(2) == 0
        for a, t in zip(ens.apps, twins):
            same(a, t)
    finally:
        ens.close()
        for t in twins:
            t.close()


@pytest.mark.gpu
def test_graph_run_after_a_halt():
    """A member that halts after an odd number of steps keeps its State through
    a second launch (its kernels return at once), the other member goes on."""
    specs = small_members(1)[:2]
    say("[ensemble] second launch after a halt")
    twins = eager_twins(specs)
    with Ensemble(specs, tune=False, nsteps=2) as ens:
        for x in (ens.apps[0], twins[0]):
            c = x.get_clock()
            c.dt0 = c.dt * 1e6
            x.set_clock(c)
        assert ens.run(2) == 1
        assert twins[0].steps(2) == 1 and twins[1].steps(2) == 0
        for a, t in zip(ens.apps, twins):
            same(a, t)
        assert ens.run(2) == 1
        assert twins[1].steps(2) == 0
        assert ens.apps[0].Model.ts == 1 and ens.apps[1].Model.ts == 4
        for a, t in zip(ens.apps, twins):
            same(a, t)
    for t in twins:
        t.close()


@pytest.mark.gpu
def test_graph_refuses_changes_to_held_handles():
    from kaiju_amd.gamera import HipStepError
    specs = small_members(1)[:1]
    with Ensemble(specs, tune=False, nsteps=2) as ens:
        a = ens.apps[0]
        with pytest.raises(HipStepError):
            a.set_tile("flux_j", 32, 8, 1, 0)
        with pytest.raises(HipStepError):
            a.set_bcs(("zerograd",) * 6)
        assert ens.run(2) == 0
    b = eager_twins(specs)[0]   # released: the same calls work
    b.set_tile("flux_j", 32, 8, 1, 0)
    b.close()


@pytest.mark.gpu
def test_graph_refuses_a_correction_wait():
    """A step that waits for another stream's event (gam_set_correction_wait)
    cannot be captured: the graph refuses the handle, takes it once the wait
    is cleared, and refuses a new wait while it holds the handle."""
    from kaiju_amd.gamera import HipStepError
    from kaiju_amd.net import GamNet
    specs = small_members(1)[:1]
    a = eager_twins(specs)[0]
    try:
        net = GamNet(*a.G)
        ev = net.infer_event()
    except HipStepError:   # the CPU emulation of the step: no GPU for the net library, so a
        net, ev = None, 1  # stand-in address; no step runs while it is set (the graph refuses first)
    try:
        a.set_correction_wait(ev)
        with pytest.raises(HipStepError, match="gam_set_correction_wait"):
            Ensemble(specs, tune=False, nsteps=2, apps=[a])
        a.set_correction_wait(None)
        ens = Ensemble(specs, tune=False, nsteps=2, apps=[a])
        try:
            with pytest.raises(HipStepError):
                a.set_correction_wait(ev)
            assert ens.run(2) == 0
        finally:
            ens.apps = []   # the app is closed below
            ens.close()
    finally:
        if net is not None:
            net.close()
        a.close()


def _totals(s: dict, vol: np.ndarray, nG: int):
    """Mass, momentum and the momentum scale sum |rho v| dV over the active cells."""
    A = tuple(slice(nG, -nG) for _ in range(3))
    g, v = s["Gas"][A], vol[A]
    mass = float((g[..., 0] * v).sum())
    mom = [float((g[..., 1 + d] * v).sum()) for d in range(3)]
    scale = float(sum((np.abs(g[..., 1 + d]) * v).sum() for d in range(3)))
    return mass, mom, scale


@pytest.mark.gpu
def test_ensemble_workload_32_seeded_blocks():
    """The ensemble configuration at its own size: 32 blocks drawn from seed
    31337 (edges 64-160 per axis, half Orszag-Tang, half blast waves with drawn
    field and pressure ratio, FP32 bulk), 4 steps under one captured graph
    with one launch per kernel for all 32 blocks.  Every block steps without a
    CalcDT halt, stays finite, keeps its mass and momentum (periodic blocks:
    FP32 storage rounding only) and its div.B."""
    specs = draw_members(32, 31337)
    say(f"[ensemble] C4 workload: {len(specs)} blocks, {sum(s.zones for s in specs)} zones")
    with Ensemble(specs, tune=False, nsteps=2, packed=True) as ens:
        say(f"   {ens.nodes() // 2} launches per step")
        before = []
        for a in ens.apps:
            vol = a.get_grid(only=("volume",))["volume"]
            before.append((_totals(a.get_state(0), vol, a.spec.nG), a.divb(), vol))
        assert ens.run(4) == 0
        for a, ((m0, p0, sc0), d0, vol) in zip(ens.apps, before):
            s = a.get_state(0)
            assert a.Model.ts == 4, a.spec.name
            assert np.isfinite(s["Gas"]).all() and np.isfinite(s["magFlux"]).all(), a.spec.name
            m1, p1, sc1 = _totals(s, vol, a.spec.nG)   # a blast wave starts at rest: scale after the steps
            assert abs(m1 - m0) <= 1e-6 * m0, (a.spec.name, m0, m1)
            assert max(abs(x - y) for x, y in zip(p0, p1)) <= 1e-6 * max(sc0, sc1), (a.spec.name, p0, p1)
            assert abs(a.divb() - d0) <= 1e-12 * np.abs(s["magFlux"]).max() * a.spec.zones, a.spec.name
    say("   32 blocks: 4 steps, status 0, mass, momentum and div.B kept")


@pytest.mark.gpu
def test_packed_graph_on_seeded_members_matches_the_oracle():
    """The oracle-sized seeded draw of the test below under the packed graph
    (FP64): every block agrees with the oracle's two steps of it."""
    _seeded_members_vs_oracle(packed=True)


@pytest.mark.gpu
def test_seeded_small_members_under_the_graph_match_the_oracle():
    """A seeded draw of oracle-sized blocks (edges 8-20, both kinds, FP64), each
    starting from the oracle's own hatched block, stepped twice under one
    captured graph: every block agrees with the oracle's two steps of it."""
    _seeded_members_vs_oracle(packed=False)


def _seeded_members_vs_oracle(packed):
    specs = draw_members(6, 4242, lo=8, hi=20, precision=0)
    say(f"[ensemble] seeded small members vs the oracle{' (packed)' if packed else ''}: {[(s.N, s.ic) for s in specs]}")
    ocs = [oracle.OracleCase(s) for s in specs]
    apps = [app_from_oracle(oc) for oc in ocs]
    with Ensemble(specs, tune=False, nsteps=2, apps=apps, packed=packed) as ens:
        assert ens.run(2) == 0
        for a, oc in zip(ens.apps, ocs):
            oc.step(2)
            e = compare_state(a, oc, 0)
            say(f"   {a.spec.name}: {e}")
            assert max(e["Gas"], e["magFlux"], e["Bxyz"]) <= 1e-11, (a.spec.name, e)
            assert a.Model.ts == oc.m.ts


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [1, 0])
def test_packed_graph_replays_match_eager_steps(precision):
    """gam_graph_create_packed: one launch per kernel for all blocks (each
    launch's blocks split among the members, their arguments read from a
    device table) gives at every replay bit for bit the eager steps' state,
    clock and status, with eager work between the replays."""
    specs = small_members(precision)
    say(f"[ensemble] packed graph vs eager, precision {precision}, {[s.N for s in specs]}")
    twins = eager_twins(specs)
    with Ensemble(specs, tune=False, nsteps=2, packed=True) as ens:
        # blocks of one kind (Boris blast waves, Orszag-Tang vortices) share
        # every launch, all blocks the ones without a Boris variant
        gb = ens.other_graph(packed=False)
        say(f"   {ens.nodes() // 2} launches per step for {len(specs)} blocks ({ens.nodes(gb) // 2} unpacked)")
        assert ens.nodes() <= 0.4 * ens.nodes(gb)
        for rep in range(3):
            ens.launch(1)
            assert ens.sync() == 0
            for t in twins:
                assert t.steps(2) == 0
            for a, t in zip(ens.apps, twins):
                same(a, t)
                a.divb()
            say(f"   replay {rep} identical")
    for t in twins:
        t.close()


@pytest.mark.gpu
def test_packed_graph_block_halts_where_gam_step_halts():
    """A block whose CalcDT halts inside the packed graph stops where gam_step
    stops; the other blocks run on."""
    say("[ensemble] halt inside the packed graph")
    specs = small_members(1)
    twins = eager_twins(specs)
    with Ensemble(specs, tune=False, nsteps=2, packed=True) as ens:
        # dt0 far above dt: CalcDT's dt/dt0 <= limDT0 check fails after step 1
        for x in (ens.apps[1], twins[1]):
            c = x.get_clock()
            c.dt0 = c.dt * 1e6
            x.set_clock(c)
        st = ens.run(4)
        assert st == 1
        assert [t.steps(4) for t in twins] == [0, 1, 0, 0]
        assert [a.Model.ts for a in ens.apps] == [4, 1, 4, 4]
        for a, t in zip(ens.apps, twins):
            same(a, t)
    for t in twins:
        t.close()


@pytest.mark.gpu
def test_packed_graph_refuses_what_it_cannot_pack():
    """Blocks the packed graph does not take (here a wall boundary) are refused
    with a message; the per-block graph takes them."""
    import dataclasses
    from kaiju_amd import gamera
    specs = small_members(1)[:2]
    specs[0] = dataclasses.replace(specs[0], bcs=("zerograd", "zerograd", "periodic", "periodic", "periodic", "periodic"))
    with pytest.raises(gamera.HipStepError, match="periodic MHD"):
        Ensemble(specs, tune=False, nsteps=2, packed=True)
