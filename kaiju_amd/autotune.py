"""Launch-brick autotuner for the stencil stages of one block.

Every stencil stage (bFlux2Fld, Mom2Face, the three EMF and three flux
directions, the update) launches 256-thread blocks that cover a brick of its
index box (``gam_set_tile``).  The best brick depends on the block's shape and
on the precision, so a block is tuned where it runs: candidate bricks are timed
with the library's stage timers over a few real steps, each stage keeping its
fastest, and the block's State/oState/clock are restored afterwards -- tuning
changes no result, only speed (the launch brick never changes a bit of the
output; tests check this).  Results are cached per (shape, precision, scheme).
"""
from __future__ import annotations

import json
import os
from typing import Dict, Optional, Tuple

LINEAR, NOXCD = 3, 4

# candidate bricks per stage (ti, tj, tk, fast); index 0 is the library default
CANDIDATES: Dict[str, list] = {
    "bflux_half": [(256, 1, 1, LINEAR | NOXCD), (256, 1, 1, LINEAR), (64, 4, 1, NOXCD), (32, 4, 2, 2)],
    "bflux_post": [(256, 1, 1, LINEAR | NOXCD), (256, 1, 1, LINEAR), (64, 4, 1, NOXCD), (32, 4, 2, 2)],
    "mom2face": [(256, 1, 1, LINEAR), (256, 1, 1, LINEAR | NOXCD), (16, 4, 4, 0), (32, 4, 2, 2)],
    "emf_i": [(64, 1, 4, NOXCD), (64, 1, 4, 2), (32, 2, 4, 2), (256, 1, 1, LINEAR | NOXCD)],
    "emf_j": [(256, 1, 1, LINEAR | NOXCD), (32, 1, 8, 2), (64, 1, 4, 2), (32, 2, 4, NOXCD)],
    "emf_k": [(32, 8, 1, 1), (64, 4, 1, 1), (32, 8, 1, 1 | NOXCD), (256, 1, 1, LINEAR | NOXCD)],
    "flux_i": [(256, 1, 1, LINEAR | NOXCD), (256, 1, 1, LINEAR), (32, 8, 1, 0), (64, 4, 1, NOXCD)],
    "flux_j": [(32, 8, 1, 0), (32, 8, 1, NOXCD), (16, 16, 1, 1), (64, 4, 1, 1)],
    "flux_k": [(64, 1, 4, 2), (32, 1, 8, 2), (64, 1, 4, 2 | NOXCD), (16, 1, 16, 2)],
    "update": [(256, 1, 1, LINEAR | NOXCD), (256, 1, 1, LINEAR), (64, 4, 1, NOXCD), (32, 8, 1, NOXCD)],
}


def shape_key(app) -> str:
    s = app.spec
    m = s.model
    return (f"{s.N[0]}x{s.N[1]}x{s.N[2]}:p{m.precision}:{m.rmeth}:b{int(m.doBoris)}:m{int(m.doMHD)}"
            f":l{int(m.doLFMLim)}")


def _stages(app) -> list:
    st = list(CANDIDATES)
    if app.spec.do25D:
        st.remove("flux_k")
    if not app.spec.model.doMHD:
        st = [x for x in st if x not in ("mom2face", "emf_i", "emf_j", "emf_k")]
    return st


def tune(app, steps_per: int = 2, cache: Optional[dict] = None) -> Dict[str, Tuple[int, int, int, int]]:
    """Pick the fastest brick per stage for ``app`` (a GamApp) and apply it.

    Each candidate round sets the round's brick on every stage, runs one
    untimed and ``steps_per`` timed steps, and records each stage's mean launch
    time; rounds past a stage's list keep its best so far.  The block's state
    and clock are saved before and restored after.
    """
    key = shape_key(app)
    if cache is not None and key in cache:
        best = {k: tuple(v) for k, v in cache[key].items()}
        for st, t in best.items():
            app.set_tile(st, *t)
        return best
    stages = _stages(app)
    s0, s1, clk = app.get_state(0), app.get_state(1), app.get_clock()
    times: Dict[str, list] = {st: [] for st in stages}
    try:
        for r in range(max(len(CANDIDATES[st]) for st in stages)):
            for st in stages:
                if r < len(CANDIDATES[st]):
                    app.set_tile(st, *CANDIDATES[st][r])
            app.steps(1)
            app.set_timing(True)
            app.steps(steps_per)
            tim = app.timings()
            app.set_timing(False)
            for st in stages:
                if r < len(CANDIDATES[st]):
                    d = tim[st]
                    times[st].append(d["ms"] / max(d["launches"], 1))
    finally:
        app.set_timing(False)
        app.put_state(0, s0["time"], s0["Gas"], s0["magFlux"], s0["Bxyz"])
        app.put_state(1, s1["time"], s1["Gas"], s1["magFlux"], s1["Bxyz"])
        app.set_clock(clk)
        app._sync_model()
    best = {}
    for st in stages:
        q = min(range(len(times[st])), key=lambda i: times[st][i])
        best[st] = tuple(CANDIDATES[st][q])
        app.set_tile(st, *best[st])
    if cache is not None:
        cache[key] = {k: list(v) for k, v in best.items()}
    return best


def load_cache(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(path: str, cache: dict) -> None:
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
