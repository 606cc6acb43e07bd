"""Batched ensemble: independent GAMERA blocks of uneven sizes stepped together.

Members are drawn from a seed: per block, edge sizes uniform in [lo, hi] per
axis (the i edge a multiple of 8, the reference's own grid constraint,
init.F90:825-834), and either an Orszag-Tang vortex or a warped blast wave
whose field strength B0 ~ U[1, 20] and pressure ratio p_in/p_out ~ U[10, 1000]
come from the same generator.  Each block is one library handle with its own
grid, state and dt; after its launch bricks are tuned (``autotune``), all
blocks' fused steps are captured as independent branches of one hipGraph
(``gam_graph_create``), and the ensemble advances by replaying it: one graph
launch per ``nsteps`` steps of every block.  Packed (``packed=True``, the
default of the command line): the graph holds one launch per kernel of the
step for all blocks together (``gam_graph_create_packed``), each launch's
blocks split among the members, instead of each block's own launches.

    python -m kaiju_amd.ensemble [--blocks 32] [--seed 31337] [--steps 20] [--fp64] [--no-tune]

prints one JSON line: zone-updates/s of the ensemble under the packed graph,
under the graph of per-block branches, and of the same blocks stepped one
after another with gam_step, plus the tuning time.
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import sys
import time
from typing import List, Optional

import numpy as np

from kaiju_amd import autotune, configs
from kaiju_amd.gamera import GamApp, HipStepError, lib


def draw_members(n: int = 32, seed: int = 31337, lo: int = 64, hi: int = 160, precision: int = 1,
                 i_multiple: int = 8) -> List[configs.CaseSpec]:
    rng = np.random.default_rng(seed)
    iedges = np.arange(((lo + i_multiple - 1) // i_multiple) * i_multiple, hi + 1, i_multiple)
    out = []
    for q in range(n):
        ni = int(rng.choice(iedges))
        nj = int(rng.integers(lo, hi + 1))
        nk = int(rng.integers(lo, hi + 1))
        kind = "ot" if rng.random() < 0.5 else "bw"
        B0 = float(rng.uniform(1.0, 20.0))
        pr = float(rng.uniform(10.0, 1000.0))
        if kind == "ot":
            s = configs.orszag_tang((ni, nj, nk), precision=precision)
        else:
            pout = 0.1
            s = configs.blast_wave((ni, nj, nk), B0=B0, pin=pr * pout, pout=pout, precision=precision)
        out.append(dataclasses.replace(s, name=f"m{q:02d}_{s.name}"))
    return out


class Ensemble:
    """The members' apps and the graph that steps them."""

    packed = False
    _others = ()   # further graphs over the same blocks (other_graph)

    def __init__(self, specs: List[configs.CaseSpec], tune: bool = True, cache: Optional[dict] = None,
                 nsteps: int = 2, apps: Optional[List[GamApp]] = None, packed: bool = False):
        """apps: the members' apps, already set up (the ensemble then owns
        them); default: one per spec, grid and initial state from InitModel.
        packed: one launch per kernel for all blocks (gam_graph_create_packed)
        instead of one graph branch of launches per block."""
        if nsteps < 2 or nsteps % 2:
            raise ValueError("nsteps must be even")
        if apps is not None and len(apps) != len(specs):
            raise ValueError("one app per spec")
        self.specs = specs
        self.nsteps = nsteps
        self.packed = packed
        self.apps = []
        t0 = time.perf_counter()
        if apps is not None:
            self.apps = list(apps)
        else:
            for s in specs:
                a = GamApp(s)
                a.InitModel()
                self.apps.append(a)
        self.setup_s = time.perf_counter() - t0
        self.tiles = {}
        t0 = time.perf_counter()
        if tune:
            for a in self.apps:
                self.tiles[a.spec.name] = autotune.tune(a, cache=cache)
        self.tune_s = time.perf_counter() - t0
        self.graph = None
        self._capture()

    def _capture(self):
        L = lib()
        hs = (C.c_void_p * len(self.apps))(*[a.h.value for a in self.apps])
        g = C.c_void_p()
        create = L.gam_graph_create_packed if self.packed else L.gam_graph_create
        rc = create(hs, len(self.apps), self.nsteps, C.byref(g))
        if rc != 0 or not g.value:
            msgs = "; ".join(m for m in (L.gam_last_error(a.h).decode() for a in self.apps) if m)
            raise HipStepError(f"{create.__name__} rc={rc} {msgs}")
        self.graph = g

    def nodes(self, graph=None) -> int:
        """Launches (and copies) in a graph: per nsteps steps of every block."""
        n = C.c_int()
        rc = lib().gam_graph_nodes(self.graph if graph is None else graph, C.byref(n))
        if rc != 0:
            raise HipStepError(f"gam_graph_nodes rc={rc}")
        return n.value

    @property
    def zones(self) -> int:
        return sum(s.zones for s in self.specs)

    def other_graph(self, packed: bool):
        """A second graph over the same blocks, the other kind (packed or
        branches); the ensemble destroys it on close."""
        L = lib()
        hs = (C.c_void_p * len(self.apps))(*[a.h.value for a in self.apps])
        g = C.c_void_p()
        create = L.gam_graph_create_packed if packed else L.gam_graph_create
        rc = create(hs, len(self.apps), self.nsteps, C.byref(g))
        if rc != 0 or not g.value:
            msgs = "; ".join(m for m in (L.gam_last_error(a.h).decode() for a in self.apps) if m)
            raise HipStepError(f"{create.__name__} rc={rc} {msgs}")
        self._others = list(self._others) + [g]
        return g

    def launch(self, replays: int, graph=None) -> None:
        g = self.graph if graph is None else graph
        rc = lib().gam_graph_launch(g, replays)
        if rc < 0:
            raise HipStepError(f"gam_graph_launch rc={rc} {lib().gam_graph_error(g).decode()}")

    def sync(self, graph=None) -> int:
        g = self.graph if graph is None else graph
        rc = lib().gam_graph_sync(g)
        if rc < 0:
            raise HipStepError(f"gam_graph_sync rc={rc} {lib().gam_graph_error(g).decode()}")
        for a in self.apps:
            a._sync_model()
        return rc

    def run(self, steps: int, graph=None) -> int:
        """steps (a multiple of nsteps) steps of every block; largest status."""
        if steps % self.nsteps:
            raise ValueError(f"steps must be a multiple of {self.nsteps}")
        self.launch(steps // self.nsteps, graph)
        return self.sync(graph)

    def close(self):
        for g in self._others:
            lib().gam_graph_destroy(g)
        self._others = ()
        if self.graph is not None and self.graph.value:
            lib().gam_graph_destroy(self.graph)
            self.graph = None
        for a in self.apps:
            a.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=32)
    ap.add_argument("--seed", type=int, default=31337)
    ap.add_argument("--lo", type=int, default=64)
    ap.add_argument("--hi", type=int, default=160)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    a = ap.parse_args(argv)
    specs = draw_members(a.blocks, a.seed, a.lo, a.hi, precision=0 if a.fp64 else 1)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    with Ensemble(specs, tune=not a.no_tune, packed=True) as ens:
        log(f"[ensemble] {len(specs)} blocks, {ens.zones} zones; setup {ens.setup_s:.1f} s, tuning {ens.tune_s:.1f} s")
        ens.run(a.warmup)
        t0 = time.perf_counter()
        st = ens.run(a.steps)
        wall_g = time.perf_counter() - t0
        log(f"[ensemble] packed graph (one launch per kernel for all blocks): {a.steps} steps in {wall_g:.3f} s, "
            f"status {st}")
        gb = ens.other_graph(packed=False)
        ens.run(a.warmup, gb)
        t0 = time.perf_counter()
        ens.run(a.steps, gb)
        wall_b = time.perf_counter() - t0
        log(f"[ensemble] graph of per-block branches: {a.steps} steps in {wall_b:.3f} s")
        # the same blocks one after another, each with its own gam_step
        for x in ens.apps:
            x.sync()
        t0 = time.perf_counter()
        for x in ens.apps:
            lib().gam_step(x.h, a.steps)
        for x in ens.apps:
            x.sync()
        wall_e = time.perf_counter() - t0
        log(f"[ensemble] eager: {a.steps} steps in {wall_e:.3f} s")
        zones = ens.zones
        out = dict(metric="zone-updates/s, ensemble of uneven blocks", unit="zone-updates/s",
                   value=round(a.steps * zones / wall_g, 1), branch_graph_value=round(a.steps * zones / wall_b, 1),
                   eager_value=round(a.steps * zones / wall_e, 1),
                   graph_speedup=round(wall_e / wall_g, 3), packed_over_branches=round(wall_b / wall_g, 3),
                   launches_per_step=dict(packed=ens.nodes() / ens.nsteps, branches=ens.nodes(gb) / ens.nsteps),
                   blocks=len(specs), zones=zones,
                   seed=a.seed, edges=[a.lo, a.hi], precision="fp64" if a.fp64 else "fp32_bulk",
                   steps=a.steps, warmup=a.warmup, ms_per_step=round(wall_g / a.steps * 1e3, 3),
                   setup_s=round(ens.setup_s, 2), tune_s=round(ens.tune_s, 2), tuned=not a.no_tune,
                   status=st)
        print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
