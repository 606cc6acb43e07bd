#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the GAMERA FP64 MHD step on one MI355X.

Workload: 3D MHD blast wave on a warped non-orthogonal 256^3 grid over the unit
cube (x = xi + 0.05 sin 2pi xi sin 2pi eta sin 2pi zeta, same for y, z),
rho = 1, p = 0.1 / 10 inside r < 0.1, v = 0, B = 10 (1,1,0)/sqrt 2, gamma = 5/3,
CFL 0.3, 7-point upwind reconstruction with the PDM limiter, FP64.  Geometry and
initial condition are built on the GPU before the timed region.

One step = one stepGamera: predictor, E field, stresses, CT/update, CalcDT,
periodic BCs.  The headline value is the plain FP64 step; the same workload in
FP32-bulk form (FP64 kept for magFlux/CT, the energy-to-pressure recovery and dt) is reported
beside it under "fp32_bulk" with its L1 error per variable against the FP64
run after the same steps, and the learned-coarse form (128^3 turbulence with the
trained correction, judged against the 256^3 FP64 solution) under
"learned_coarse" with its speed at equal L1 error.  value = steps * 256^3 /
wall time of the timed steps (the reference's kZCs definition x 1e3, reference
src/gamera/output.F90:64-85).  By default the timed steps are the workload's own
100-step run from t = 0, on a fresh handle after 5 warm-up steps on another.

Prints ONE JSON line.  ``--gpus N`` with N > 1 exits: this build runs one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E
ALGO_BYTES_PER_ZONE = 824.0      # FP64 step: 192 B state + 632 B geometry per zone-update
# FP32-bulk: Gas 20 B x 3 touches + magFlux (FP64) 24 B x 3 + FP32 bulk geometry
# (volume, face 3, Tf 27, Te 18, Teb 12, edge 3, di/dj/dk 3: 67 x 4 B) + FP64
# positions and volume (xfc 9, xyzcc 3, volume 1: 13 x 8 B)
ALGO_BYTES_PER_ZONE_FP32 = 60.0 + 72.0 + 67 * 4.0 + 13 * 8.0
METRIC = "zone-updates/s per MI355X: plain FP64 vs FP32/learned at fixed L1 error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gpu_run(n: int, steps: int, warmup: int, precision: int = 0):
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp
    spec = configs.blast_wave(n, steps=steps, precision=precision)
    if warmup > 0:   # warm up on a handle of its own; the timed steps start from t = 0
        with GamApp(spec) as w:
            w.InitModel()
            st = w.steps(warmup)
            w.sync()
        if st:
            raise RuntimeError(f"warmup status {st}")
    t0 = time.perf_counter()
    app = GamApp(spec)
    app.InitModel()
    log(f"[bench] setup {spec.name} on GPU: {time.perf_counter() - t0:.2f} s, dt0 = {app.Model.dt:.4e}")
    app.sync()
    t1 = time.perf_counter()
    st = app.steps(steps)
    app.sync()
    t2 = time.perf_counter()
    if st:
        raise RuntimeError(f"timed steps status {st}")
    # per-stage split from as many steps again with stage timers on (the timers
    # put every stage on one stream, so this pass is not the one timed above)
    app.set_timing(True)
    st = app.steps(steps)
    app.sync()
    if st:
        raise RuntimeError(f"stage-timer steps status {st}")
    tim = app.timings()
    app.set_timing(False)
    divb = app.divb()
    wall = t2 - t1
    log(f"[bench] {'FP32-bulk' if precision else 'FP64'}: {steps} steps in {wall:.3f} s, t = {app.Model.t:.4e}, divB = {divb:.3e}")
    return app, spec, wall, tim, divb


def l1_per_variable(a, b, nG):
    """L1 error of each Gas variable over the active cells, relative to the L1
    norm of its field (momentum components against the momentum magnitude)."""
    import numpy as np
    A = tuple(slice(nG, -nG) for _ in range(3))
    a, b = a[A], b[A]
    out = {}
    for v, name in enumerate(("rho", "momx", "momy", "momz", "energy")):
        grp = (1, 2, 3) if 1 <= v <= 3 else (v,)
        den = sum(float(np.abs(b[..., q]).sum()) for q in grp) / len(grp)
        out[name] = float(np.abs(a[..., v] - b[..., v]).sum()) / max(den, 1e-300)
    return out


def fp32_leg(n: int, steps: int, warmup: int, ms64: float, dt64: float):
    """The same workload in FP32-bulk form: throughput, and the L1 error per
    variable against FP64 at equal time: both precisions stepped `steps`
    times with one fixed dt (each precision's own CFL dt would end the runs at
    slightly different times, and the offset, not the precision, would
    dominate the difference in a moving blast)."""
    import numpy as np
    app, spec, wall, tim, divb = gpu_run(n, steps, warmup, precision=1)
    app.close()
    zups = steps * spec.zones / wall
    ms = wall / steps * 1e3
    s64, _, _ = _fixed_run(n, 0, dt64, steps)
    s32, _, _ = _fixed_run(n, 1, dt64, steps)
    l1 = l1_per_variable(s32["Gas"], s64["Gas"], spec.nG)
    mf = float(np.abs(s32["magFlux"] - s64["magFlux"]).sum() / np.abs(s64["magFlux"]).sum())
    bx = float(np.abs(s32["Bxyz"] - s64["Bxyz"]).sum() / np.abs(s64["Bxyz"]).sum())
    del s64, s32
    step_gbs = ALGO_BYTES_PER_ZONE_FP32 * zups / 1e9
    return dict(value=round(zups, 1), ms_per_step=round(ms, 4), speedup_vs_fp64=round(ms64 / ms, 3),
                l1_rel_vs_fp64={**{k: float(f"{v:.3e}") for k, v in l1.items()},
                                "magFlux": float(f"{mf:.3e}"), "Bxyz": float(f"{bx:.3e}")},
                l1_compared_after=f"{steps} steps of fixed dt = {dt64:.4e} in both precisions", divB_total_after=divb,
                algo_bytes_per_zone=ALGO_BYTES_PER_ZONE_FP32, step_algo_GBs=round(step_gbs, 1),
                step_frac=round(step_gbs / HBM_PEAK_GBS, 4),
                stage_avg_ms={k: round(v["ms"] / v["launches"], 4) for k, v in tim.items() if v["launches"] > 0})


def _fixed_run(n: int, precision: int, dt: float, nsteps: int):
    """Blast wave at n^3 stepped nsteps times with a fixed dt; (state, grid volume, wall s)."""
    import dataclasses
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp
    spec = configs.blast_wave(n, steps=nsteps, precision=precision)
    spec = dataclasses.replace(spec, model=dataclasses.replace(spec.model, fixedTimestep=True))
    with GamApp(spec) as app:
        app.InitModel()
        c = app.get_clock()
        c.dt, c.oStateTime = dt, c.stateTime - dt
        app.set_clock(c)
        app.sync()
        t0 = time.perf_counter()
        st = app.steps(nsteps)
        app.sync()
        wall = time.perf_counter() - t0
        if st:
            raise RuntimeError(f"fixed-dt run {n}^3 status {st}")
        s = app.get_state(0)
        vol = app.get_grid()["volume"]
    return s, vol, wall


def equal_error_leg(n: int = 128, steps: int = 20):
    """FP64 and FP32-bulk at n^3 against the FP64 solution at (2n)^3, same
    physical time (fixed dt, dt/2 on the fine grid): L1 error per conserved
    variable of each coarse run against the fine run downsampled by
    volume-weighted 2x2x2 averages (the warp maps the fine corners of even index
    onto the coarse corners)."""
    import numpy as np
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp
    with GamApp(configs.blast_wave(n)) as app:   # CFL dt of the initial state
        app.InitModel()
        dt = 0.8 * app.Model.dt
    s64, vol, w64 = _fixed_run(n, 0, dt, steps)
    s32, _, w32 = _fixed_run(n, 1, dt, steps)
    sf, volf, wf = _fixed_run(2 * n, 0, 0.5 * dt, 2 * steps)
    nG = 4
    A = slice(nG, nG + 2 * n)
    U = sf["Gas"][A, A, A] * volf[A, A, A][..., None]
    V = volf[A, A, A]
    sh = (n, 2, n, 2, n, 2)
    Vc = V.reshape(sh).sum(axis=(1, 3, 5))
    ref = np.stack([U[..., v].reshape(sh).sum(axis=(1, 3, 5)) / Vc for v in range(5)], axis=-1)
    del U, sf
    a = slice(nG, nG + n)
    names = ("rho", "momx", "momy", "momz", "energy")

    def l1(s):
        g = s["Gas"][a, a, a]
        out = {}
        for v, name in enumerate(names):
            grp = (1, 2, 3) if 1 <= v <= 3 else (v,)
            den = sum(float(np.abs(ref[..., q]).mean()) for q in grp) / len(grp)
            out[name] = float(f"{float(np.abs(g[..., v] - ref[..., v]).mean()) / den:.4e}")
        return out
    e64, e32 = l1(s64), l1(s32)
    ratio = max(e32[k] / e64[k] for k in names)
    log(f"[bench] equal-error check {n}^3 vs {2 * n}^3: FP64 {e64}  FP32 {e32}  worst ratio {ratio:.4f}")
    return dict(workload=f"blast wave, warped {n}^3 (FP64 and FP32 bulk) against FP64 at {2 * n}^3, {steps} steps "
                         f"of fixed dt = {dt:.4e} (fine: {2 * steps} of dt/2)",
                l1_vs_fine_fp64=e64, l1_vs_fine_fp32=e32, worst_error_ratio_fp32_over_fp64=round(ratio, 5),
                coarse_ms_per_step_fp64=round(w64 / steps * 1e3, 4), coarse_ms_per_step_fp32=round(w32 / steps * 1e3, 4))


def msph_leg():
    """The magnetosphere-like configuration (192x192x256 spherical-polar grid to
    the axis, LFM axis BCs with 12 averaged rings, dipole background, FP32 bulk,
    Boris): its 500 steps timed from the hatched state (setup excluded), then 20
    more with stage timers for the per-stage split."""
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp
    spec = configs.magnetosphere()
    t0 = time.perf_counter()
    app = GamApp(spec)
    app.InitModel()
    app.sync()
    setup = time.perf_counter() - t0
    divb0 = app.divb()
    t0 = time.perf_counter()
    st = app.steps(spec.steps)
    app.sync()
    wall = time.perf_counter() - t0
    divb = app.divb()
    t_end, dt = app.Model.t, app.Model.dt
    app.set_timing(True)
    app.steps(20)
    app.sync()
    tim = app.timings()
    app.close()
    log(f"[bench] msph {spec.N}: setup {setup:.1f} s, {spec.steps} steps in {wall:.3f} s, status {st}")
    return dict(workload="magnetosphere-like block: stretched spherical-polar 192x192x256 to the x axis (hole 1e-3 rad), "
                         "r 2-30 R_E, Earth dipole background, constant solar wind on the outer face (WindBC; the wind "
                         "starts tapered to rest inside ~12 R_E), inner radial face the Earth setup's ionospheric BC with "
                         "no ionosphere attached (IonInner with E = 0: mirrored gas at rest, the field fixed to the "
                         "dipole), LFM axis "
                         "BCs, ring average (12 rings), Boris 1000 km/s, FP32 bulk (FP64 CT/energy/dt)",
                value=round(spec.steps * spec.zones / wall, 1), unit="zone-updates/s",
                ms_per_step=round(wall / spec.steps * 1e3, 4), steps=spec.steps, status=st,
                t_end_code_units=t_end, dt_last=dt, divB_total_before=divb0, divB_total_after=divb,
                setup_s=round(setup, 2),
                stage_avg_ms={k: round(v["ms"] / v["launches"], 4) for k, v in tim.items() if v["launches"] > 0})


def roofline(tim: dict, ms_per_step: float, zups: float):
    # dominant kernel = largest share of the timed region
    stages = {k: v for k, v in tim.items() if v["launches"] > 0}
    dom = max(stages, key=lambda k: stages[k]["ms"])
    d = stages[dom]
    avg_ms = d["ms"] / d["launches"]
    if avg_ms <= 0:
        raise RuntimeError("stage timers recorded no time")
    achieved = d["bytes"] / (avg_ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "hbm_traffic_bw256_fp64.json")
    if os.path.exists(pmc):
        try:
            traffic = json.load(open(pmc)).get(dom, {}).get("bytes_per_launch")
        except Exception:
            traffic = None
    step_gbs = ALGO_BYTES_PER_ZONE * zups / 1e9
    return dict(kernel=dom, bound="hbm", achieved=round(achieved, 2), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic,
                avg_launch_ms=round(avg_ms, 4), algo_bytes_per_launch=d["bytes"],
                step_algo_GBs=round(step_gbs, 1), step_frac=round(step_gbs / HBM_PEAK_GBS, 4),
                stage_avg_ms={k: round(v["ms"] / v["launches"], 4) for k, v in stages.items()},
                # every stage's algorithmic bytes per launch over its average launch time, as a
                # fraction of the HBM peak (the same measure as frac, for the other kernels)
                stage_hbm_frac={k: round(v["bytes"] / (v["ms"] / v["launches"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                                for k, v in stages.items() if v["bytes"] > 0 and v["ms"] > 0})


def host_cpu() -> dict:
    """lscpu's model, sockets, cores and threads per core, and the logical CPUs
    this process may run on."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass

    def num(key):
        try:
            return int(info[key])
        except (KeyError, ValueError):
            return None
    sockets, cps, tpc = num("Socket(s)"), num("Core(s) per socket"), num("Thread(s) per core")
    return dict(model=info.get("Model name") or platform.processor() or platform.machine(),
                sockets=sockets, cores_per_socket=cps, threads_per_core=tpc,
                physical_cores=(sockets * cps) if sockets and cps else None,
                logical_cpus_visible=len(os.sched_getaffinity(0)))


def cpu_baseline(n: int, budget_s: float):
    """The oracle (CPU restatement, OpenMP) timed on this host's cores on a
    bounded sample: the same blast-wave problem at n^3, inputs built on the GPU
    and copied to the oracle, timed over whole steps.  One OpenMP thread per
    core of the process's CPU share (OMP_NUM_THREADS), bound close to cores."""
    import oracle
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp
    from tests.helpers import compare_state
    oracle.build()
    spec = configs.blast_wave(n, steps=1)
    app = GamApp(spec)
    app.InitModel()
    oc = oracle.OracleCase(spec, setup=False)
    g = app.get_grid()
    for k in oc.grid:
        oc.grid[k][...] = g[k]
    for which, st in ((0, oc.state), (1, oc.ostate)):
        s = app.get_state(which)
        st.Gas[...] = s["Gas"]; st.magFlux[...] = s["magFlux"]; st.Bxyz[...] = s["Bxyz"]
        st.s.time = s["time"]
    c = app.get_clock()
    oc.m.t, oc.m.dt, oc.m.dt0, oc.m.ts = c.t, c.dt, c.dt0, c.ts
    t_one = oc.timed_steps(1)
    nsteps = max(1, min(200, int(budget_s / max(t_one, 1e-6))))
    wall = oc.timed_steps(nsteps)
    # the same steps on the GPU: the sample doubles as a parity spot check
    app.steps(nsteps + 1)
    err = compare_state(app, oc, 0)
    app.close()
    cores = oracle.omp_threads()
    zups = nsteps * spec.zones / wall
    hc = host_cpu()
    return dict(value=round(zups, 1), unit="zone-updates/s", cores=cores, kind="port",
                sample=f"oracle (C/OpenMP restatement) blast wave warped {n}^3, {nsteps} steps, {wall:.1f} s on "
                       f"{cores} OpenMP threads (OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}, "
                       f"OMP_PLACES={os.environ.get('OMP_PLACES')}) of {hc['model']}",
                host=hc, gpu_vs_oracle_max_rel_err=max(err["Gas"], err["magFlux"], err["Bxyz"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (default: the 256^3 run's 100, from t = 0)")
    ap.add_argument("--warmup", type=int, default=5, help="untimed steps first, on a handle of their own")
    ap.add_argument("--n", type=int, default=256, help="cells per axis")
    ap.add_argument("--cpu-n", type=int, default=96, help="cells per axis of the CPU-baseline sample")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of oracle stepping")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-bulk leg")
    ap.add_argument("--no-msph", action="store_true", help="skip the magnetosphere leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-coarse leg")
    ap.add_argument("--msph-only", action="store_true",
                    help="run only the magnetosphere leg and print its JSON (profiling; not the bench line)")
    args = ap.parse_args()
    # the CPU baseline's OpenMP threads: one per core, packed (read when the
    # oracle library first starts its runtime, after the GPU legs)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.msph_only:
        print(json.dumps(msph_leg()))
        return

    app, spec, wall, tim, divb = gpu_run(args.n, args.steps, args.warmup)
    zones = spec.zones
    zups = args.steps * zones / wall
    ms = wall / args.steps * 1e3
    rf = roofline(tim, ms, zups)
    fp32 = None
    if not args.no_fp32:
        dt_late = app.Model.dt   # the CFL dt after the run: a fixed dt stable over all of it
        app.close()
        fp32 = fp32_leg(args.n, args.steps, args.warmup, ms, dt_late)
    else:
        app.close()
    eq = None
    if not args.no_fp32:
        try:
            eq = equal_error_leg()
        except Exception as e:
            log(f"[bench] equal-error leg failed: {e!r}")
            eq = {"error": repr(e)}
        if fp32 is not None:
            fp32["equal_error"] = eq
    msph = None
    if not args.no_msph:
        try:
            msph = msph_leg()
        except Exception as e:   # reported beside the headline, never in place of it
            log(f"[bench] msph leg failed: {e!r}")
            msph = {"value": None, "error": repr(e)}
    learned_leg = None
    if not args.no_learned:
        try:
            from kaiju_amd import learned
            learned_leg = learned.equal_error_leg(log=log)
        except Exception as e:   # reported beside the headline, never in place of it
            log(f"[bench] learned leg failed: {e!r}")
            learned_leg = {"value": None, "error": repr(e)}
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args.cpu_n, args.cpu_budget)
        except Exception as e:   # reported, never substituted for the GPU number
            log(f"[bench] cpu baseline failed: {e!r}")
            cpu = {"value": None, "error": repr(e)}
    out = {
        "metric": METRIC,
        "value": round(zups, 1),
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (blast-wave initial condition built on the GPU)",
        "config": {"workload": f"3D MHD blast wave, warped non-orthogonal {args.n}^3, plain FP64, 7UP+PDM, CFL 0.3",
                   "cells": [args.n] * 3, "ghosts": 4, "precision": "fp64", "divB_total_after": divb},
        "roofline": rf,
        "cpu_baseline": cpu,
        "fp32_bulk": fp32,
        "magnetosphere": msph,
        "learned_coarse": learned_leg,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
