"""Learned coarse-grid correction: training data, GPU training and rollouts.

The network of ``libgamera_net.so`` (three 3x3x3 convolutions, 8 -> 32 -> 32
-> 8 channels, BF16 MFMA with FP32 accumulation) maps the features of a coarse
state (rho, v, p, B) to a correction field o[G][8].  ``gam_set_correction``
makes every following step add ``scale * face * o[0:5]`` to the gas fluxes of
each active interface and ``scale * edge * o[5:8]`` to each active edge EMF,
before E2Flux and the update: a corrected step stays conservative and keeps
div.B at machine zero.

Training data are the solver's own fine trajectories.  Decaying MHD
turbulence from seeded Fourier modes (``TurbIC``) on the periodic unit cube
at (2N)^3 cells is stepped with a fixed dt/2 and every second state is
downsampled 2x: conserved variables averaged over 2x2x2 fine cells, face
fluxes summed over 2x2 fine faces (both exact on the Cartesian grid, so the
coarse magFlux keeps div.B = 0).  The target of the coarse step from sample n
is the downsampled fine state at n+1 minus the plain coarse step from the
downsampled states n (and n-1 as the AB2 predictor's old state).  Samples stay
resident on the GPU.  The loss is that of the linearised one-step change that
o makes (``gam_correction_grad``: exact for density, momentum and magFlux,
first order for the energy), and its gradient runs back through the network
(``gam_net_backward``) into Adam (``gam_net_adam``); every stage runs on the
GPU.  After the first phase, aggregation rounds add samples on the corrected
model's own rollout states (target: the downsampled fine state one step
later), so the network also learns to undo the drift of its rollouts.

    python -m kaiju_amd.learned [--N 32] [--train-seeds 0,1,2,3] [--test-seed 100] [--steps 24] [--iters 1500] [--dagger 0]

prints one JSON line with the one-step loss ratio and the plain and corrected
errors (one step from each sample, and a rollout from sample 0) against the
downsampled fine run of a held-out seed.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import math
import sys
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from kaiju_amd import configs
from kaiju_amd.gamera import DeviceArray, GamApp

GAMMA = 5.0 / 3.0
PRIM = ("rho", "vx", "vy", "vz", "p", "bx", "by", "bz")


def turb_spec(N: int, precision: int = 0, nG: int = 4) -> configs.CaseSpec:
    """Periodic unit cube, N^3 cells, fixed time step (set by the caller)."""
    m = configs.ModelSpec(gamma=GAMMA, CFL=0.3, fixedTimestep=True, precision=precision)
    return configs.CaseSpec(name=f"turb{N}", N=(N, N, N), bounds=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), grid="cart",
                            ic="turb", model=m, nG=nG)


# ----------------------------------------------------------------- initial state
@dataclasses.dataclass
class TurbIC:
    """rho = 1, p = 1, v and the vector potential A sums of `modes` seeded
    cosine modes with integer wave numbers |n_x|,|n_y|,|n_z| <= kmax, scaled to
    the rms v_rms and rms(curl A) = b_rms, plus a uniform field b_mean."""
    seed: int
    kmax: int = 4
    modes: int = 12
    v_rms: float = 1.0
    b_rms: float = 1.0
    b_mean: Sequence[float] = (0.0, 0.0, 0.0)

    def _draw(self, rng):
        out = []
        while len(out) < self.modes:
            n = rng.integers(-self.kmax, self.kmax + 1, size=3)
            if not n.any():
                continue
            out.append((2.0 * math.pi * n.astype(np.float64), rng.normal(size=3), rng.uniform(0, 2 * math.pi)))
        return out

    def fields(self, N: int, nG: int):
        """Gas (G,5) and magFlux (F,3), Fortran order, ghost cells holding the
        periodic images of the active ones."""
        rng = np.random.default_rng(self.seed)
        vmodes, amodes = self._draw(rng), self._draw(rng)
        dx = 1.0 / N
        Ni = N + 2 * nG
        xc = (np.arange(Ni) - nG + 0.5) * dx           # cell centres
        xe = (np.arange(Ni + 1) - nG) * dx             # corners

        def wave(k, ph, x):
            """exp(i (k.x + ph)) on the tensor grid x^3, as an outer product of 1-D factors."""
            ex, ey, ez = (np.exp(1j * k[d] * x) for d in range(3))
            return np.exp(1j * ph) * ex[:, None, None] * ey[None, :, None] * ez[None, None, :]

        v = np.zeros((3, Ni, Ni, Ni))
        for k, a, ph in vmodes:
            c = wave(k, ph, xc).real
            for d in range(3):
                v[d] += a[d] * c
        v *= self.v_rms / math.sqrt(np.mean(np.sum(v[:, nG:-nG, nG:-nG, nG:-nG] ** 2, axis=0)))
        # edge integrals of A: L[d][i,j,k] = integral of A_d along the d-edge from corner (i,j,k)
        L = np.zeros((3, Ni + 1, Ni + 1, Ni + 1))
        for k, a, ph in amodes:
            w = wave(k, ph, xe)
            for d in range(3):
                if abs(k[d]) > 0:   # integral of cos(k.x + ph) over [x_d, x_d + dx]
                    L[d] += a[d] * (w * ((np.exp(1j * k[d] * dx) - 1.0) / (1j * k[d]))).real
                else:
                    L[d] += a[d] * w.real * dx
        F = L.shape[1]
        mag = np.zeros((F, F, F, 3), order="F")
        s = lambda a, di, dj, dk: np.roll(a, (-di, -dj, -dk), axis=(0, 1, 2))   # a[i+di, j+dj, k+dk]
        mag[..., 0] = L[1] + s(L[2], 0, 1, 0) - s(L[1], 0, 0, 1) - L[2]
        mag[..., 1] = s(L[0], 0, 0, 1) - L[0] - s(L[2], 1, 0, 0) + L[2]
        mag[..., 2] = s(L[1], 1, 0, 0) - L[1] - s(L[0], 0, 1, 0) + L[0]
        a0, a1 = nG, nG + N
        brms = math.sqrt(np.mean(mag[a0:a1, a0:a1, a0:a1] ** 2)) / dx ** 2 * math.sqrt(3.0)
        mag *= self.b_rms / brms
        for d in range(3):
            mag[..., d] += self.b_mean[d] * dx * dx
        mag = wrap_faces(mag[a0:a1, a0:a1, a0:a1], nG)   # the rolled edge layer is invalid: rebuild ghosts
        v = np.moveaxis(v[:, a0:a1, a0:a1, a0:a1], 0, -1)
        gas = np.empty((N, N, N, 5))
        gas[..., 0] = 1.0
        gas[..., 1:4] = v
        gas[..., 4] = 1.0 / (GAMMA - 1.0) + 0.5 * np.sum(v ** 2, axis=-1)
        return wrap_cells(gas, nG), mag

    def max_speed(self) -> float:
        """Upper estimate of |v| + fast speed for the time step."""
        b = self.b_rms * 2.5 + float(np.linalg.norm(self.b_mean))
        return self.v_rms * 2.5 + math.sqrt(GAMMA + b * b)


# ----------------------------------------------------------------- 2x downsampling
def wrap_cells(act: np.ndarray, nG: int) -> np.ndarray:
    """Active cells (N,N,N,...) -> ghost-inclusive periodic array (Fortran order)."""
    N = act.shape[0]
    m = (np.arange(N + 2 * nG) - nG) % N
    return np.asfortranarray(act[np.ix_(m, m, m)])


def wrap_faces(act: np.ndarray, nG: int) -> np.ndarray:
    """Active faces (N,N,N,3) (face index = lower face of the cell) -> (N+2nG+1)^3 x 3."""
    N = act.shape[0]
    m = (np.arange(N + 2 * nG + 1) - nG) % N
    return np.asfortranarray(act[np.ix_(m, m, m)])


def downsample(gas: np.ndarray, mag: np.ndarray, nG: int):
    """Fine (2N)^3 state -> coarse N^3 state, ghosts by periodic wrap."""
    Nf = gas.shape[0] - 2 * nG
    N = Nf // 2
    g = gas[nG:nG + Nf, nG:nG + Nf, nG:nG + Nf]
    gc = g.reshape(N, 2, N, 2, N, 2, 5).mean(axis=(1, 3, 5))
    f = mag[nG:nG + Nf, nG:nG + Nf, nG:nG + Nf]
    mc = np.empty((N, N, N, 3))
    mc[..., 0] = f[0::2, :, :, 0].reshape(N, N, 2, N, 2).sum(axis=(2, 4))
    mc[..., 1] = f[:, 0::2, :, 1].reshape(N, 2, N, N, 2).sum(axis=(1, 4))
    mc[..., 2] = f[:, :, 0::2, 2].reshape(N, 2, N, 2, N).sum(axis=(1, 3))
    return wrap_cells(gc, nG), wrap_faces(mc, nG)


def primitives(gas: np.ndarray, B: np.ndarray, nG: int) -> Dict[str, np.ndarray]:
    """Active-cell rho, v, p, B."""
    a = slice(nG, gas.shape[0] - nG)
    g = gas[a, a, a]
    rho = g[..., 0]
    v = g[..., 1:4] / rho[..., None]
    p = (GAMMA - 1.0) * (g[..., 4] - 0.5 * rho * np.sum(v * v, axis=-1))
    b = B[a, a, a]
    return dict(rho=rho, vx=v[..., 0], vy=v[..., 1], vz=v[..., 2], p=p, bx=b[..., 0], by=b[..., 1], bz=b[..., 2])


def cell_b(mag: np.ndarray, nG: int, dx: float) -> np.ndarray:
    """Cell-centred B of a Cartesian-grid magFlux (mean of the two faces / area)."""
    N = mag.shape[0] - 1
    B = np.empty((N, N, N, 3))
    B[..., 0] = 0.5 * (mag[:-1, :-1, :-1, 0] + mag[1:, :-1, :-1, 0])
    B[..., 1] = 0.5 * (mag[:-1, :-1, :-1, 1] + mag[:-1, 1:, :-1, 1])
    B[..., 2] = 0.5 * (mag[:-1, :-1, :-1, 2] + mag[:-1, :-1, 1:, 2])
    return B / (dx * dx)


def l1_errors(gas, mag, ref_gas, ref_mag, nG: int, dx: float) -> Dict[str, float]:
    """Mean |x - x_ref| over the active cells, per primitive variable."""
    a = primitives(gas, cell_b(mag, nG, dx), nG)
    b = primitives(ref_gas, cell_b(ref_mag, nG, dx), nG)
    return {k: float(np.mean(np.abs(a[k] - b[k]))) for k in PRIM}


def l1_total(e: Dict[str, float], norm: Dict[str, float]) -> float:
    """Sum over variables of the L1 error relative to the variable's scale."""
    return float(sum(e[k] / norm[k] for k in PRIM))


# ----------------------------------------------------------------- trajectories
@dataclasses.dataclass
class Trajectory:
    """Downsampled fine states at t_n = n * dt (coarse step dt), host arrays."""
    N: int
    nG: int
    dt: float
    gas: List[np.ndarray]
    mag: List[np.ndarray]
    fine_s: float = 0.0


def choose_dt(ic: TurbIC, N: int, cfl: float = 0.3) -> float:
    return cfl * (1.0 / N) / ic.max_speed()


def set_time(app: GamApp, t: float, dt: float, ts: int):
    c = app.get_clock()
    c.t, c.dt, c.stateTime, c.oStateTime, c.ts, c.status = t, dt, t, t - dt, ts, 0
    app.set_clock(c)


def fine_trajectory(ic: TurbIC, N: int, steps: int, dt: Optional[float] = None, nG: int = 4, log=None) -> Trajectory:
    """Run the solver at (2N)^3 with dt/2 for `steps` coarse intervals and
    downsample every second state."""
    from kaiju_amd.gamera import gen_corners
    dt = choose_dt(ic, N) if dt is None else dt
    gas, mag = ic.fields(2 * N, nG)
    t0 = time.perf_counter()
    with GamApp(turb_spec(2 * N, 0, nG)) as app:
        app.set_corners(gen_corners(app.spec))
        app.put_state(0, 0.0, gas, mag, None)
        del gas, mag
        set_time(app, 0.0, 0.5 * dt, 0)
        app.hatch()
        out_g, out_m = [], []
        for n in range(steps + 1):
            if n:
                st = app.steps(2)
                if st != 0:
                    raise RuntimeError(f"fine run halted with status {st} at coarse step {n}")
            s = app.get_state(0)
            g, m = downsample(s["Gas"], s["magFlux"], nG)
            del s
            out_g.append(g)
            out_m.append(m)
            if log and n % 6 == 0:
                log(f"[learned] seed {ic.seed}: fine sample {n}/{steps}")
    return Trajectory(N=N, nG=nG, dt=dt, gas=out_g, mag=out_m, fine_s=time.perf_counter() - t0)


def coarse_app(N: int, precision: int = 0, nG: int = 4) -> GamApp:
    from kaiju_amd.gamera import gen_corners
    app = GamApp(turb_spec(N, precision, nG))
    app.set_corners(gen_corners(app.spec))
    return app


def load_pair(app: GamApp, tr: Trajectory, n: int):
    """State = sample n, oState = sample n-1 (sample 0 for n = 0), clock at t_n."""
    t, dt = n * tr.dt, tr.dt
    m = max(n - 1, 0)
    app.put_state(0, t, tr.gas[n], tr.mag[n], None)
    app.put_state(1, t - dt, tr.gas[m], tr.mag[m], None)
    set_time(app, t, dt, n)


def plain_step(app: GamApp, gas, mag, ogas, omag, t: float, dt: float, ts: int):
    """One uncorrected coarse step from (State, oState); the new Gas and magFlux."""
    app.put_state(0, t, gas, mag, None)
    app.put_state(1, t - dt, ogas, omag, None)
    set_time(app, t, dt, ts)
    app.set_correction(None)
    if app.steps(1) != 0:
        raise RuntimeError(f"coarse step at t = {t} halted")
    s = app.get_state(0)
    return s["Gas"], s["magFlux"]


# ----------------------------------------------------------------- training samples
def _flat(a: np.ndarray) -> np.ndarray:
    return np.ravel(a, order="F")      # the library's memory order


class Sample:
    """One training sample resident on the GPU: the state the network sees
    (Gas, magFlux, FP64) and the target change of its step."""

    def __init__(self, t: float, gas, mag, tg, tm):
        self.t = t
        self.gas, self.mag = DeviceArray.of(_flat(gas), np.float64), DeviceArray.of(_flat(mag), np.float64)
        self.tg, self.tm = DeviceArray.of(_flat(tg), np.float64), DeviceArray.of(_flat(tm), np.float64)

    def free(self):
        for d in (self.gas, self.mag, self.tg, self.tm):
            d.free()


def trajectory_targets(app: GamApp, tr: Trajectory):
    """Host targets (Gas, magFlux) of the plain coarse step from each sample n < last."""
    out = []
    for n in range(len(tr.gas) - 1):
        m = max(n - 1, 0)
        g1, m1 = plain_step(app, tr.gas[n], tr.mag[n], tr.gas[m], tr.mag[m], n * tr.dt, tr.dt, n)
        out.append((tr.gas[n + 1] - g1, tr.mag[n + 1] - m1))
    return out


def trajectory_samples(tr: Trajectory, tgts) -> List[Sample]:
    return [Sample(n * tr.dt, tr.gas[n], tr.mag[n], tg, tm) for n, (tg, tm) in enumerate(tgts)]


def rollout_samples(app: GamApp, tr: Trajectory, corr: "Corrector", steps: int) -> List[Sample]:
    """Samples on the corrected model's own states: roll out from sample 0 with
    the correction, then target each visited state's plain step at the
    downsampled fine state one step later (the correction learns to undo the
    drift its rollouts accumulate)."""
    app.set_correction(None)
    app.put_state(0, 0.0, tr.gas[0], tr.mag[0], None)
    app.put_state(1, -tr.dt, tr.gas[0], tr.mag[0], None)
    set_time(app, 0.0, tr.dt, 0)
    visited = []
    try:
        for n in range(steps):
            s0, s1 = app.get_state(0), app.get_state(1)
            if not (np.isfinite(s0["Gas"]).all() and np.isfinite(s0["magFlux"]).all()):
                break                                  # the rollout left the physical states: stop there
            visited.append((s0["Gas"], s0["magFlux"], s1["Gas"], s1["magFlux"]))
            app.set_correction(corr.infer(), corr.norm.scale)
            if app.steps(1) != 0:
                break
    finally:
        app.set_correction(None)
    out = []
    for n, (g, m, og, om) in enumerate(visited):
        g1, m1 = plain_step(app, g, m, og, om, n * tr.dt, tr.dt, n)
        tg, tm = tr.gas[n + 1] - g1, tr.mag[n + 1] - m1
        if not (np.isfinite(tg).all() and np.isfinite(tm).all()):
            break
        out.append(Sample(n * tr.dt, g, m, tg, tm))
    return out


# ----------------------------------------------------------------- training
@dataclasses.dataclass
class Normalisation:
    scale: float                       # correction scale s
    weights: np.ndarray                # w_rho, w_mom, w_energy, w_magFlux
    feat_scales: tuple                 # rho, v, p, B
    feat_offsets: tuple                # rho0, p0


def normalisation(trs: List[Trajectory], tgts) -> Normalisation:
    N, nG, dt = trs[0].N, trs[0].nG, trs[0].dt
    dx = 1.0 / N
    a = slice(nG, nG + N)
    acc = np.zeros(4)
    cnt = np.zeros(4)
    for tl in tgts:
        for tg, tm in tl:
            g = tg[a, a, a]
            acc += [np.sum(g[..., 0] ** 2), np.sum(g[..., 1:4] ** 2), np.sum(g[..., 4] ** 2), np.sum(tm[a, a, a] ** 2)]
            cnt += [g[..., 0].size, g[..., 1:4].size, g[..., 4].size, tm[a, a, a].size]
    rho_r, mom_r, en_r, mag_r = np.sqrt(acc / cnt) + 1e-300
    ncell = N ** 3
    w = np.array([1.0 / rho_r ** 2, 1.0 / mom_r ** 2, 1.0 / en_r ** 2, 1.0 / mag_r ** 2]) / ncell
    scale = float(max(rho_r, mom_r, en_r, mag_r / dx ** 2) * dx / dt)
    stats = {k: [] for k in PRIM}
    for tr in trs:
        for g, m in zip(tr.gas, tr.mag):
            p = primitives(g, cell_b(m, nG, dx), nG)
            for k in PRIM:
                stats[k].append((float(p[k].mean()), float(np.mean(p[k] ** 2))))
    mean = lambda k: float(np.mean([s[0] for s in stats[k]]))
    msq = lambda ks: float(np.mean([s[1] for k in ks for s in stats[k]]))
    rho0, p0 = mean("rho"), mean("p")
    fs = (math.sqrt(max(msq(("rho",)) - rho0 ** 2, 0.0)) + 1e-6, math.sqrt(msq(("vx", "vy", "vz"))),
          math.sqrt(max(msq(("p",)) - p0 ** 2, 0.0)) + 1e-6, math.sqrt(msq(("bx", "by", "bz"))))
    return Normalisation(scale=scale, weights=w, feat_scales=fs, feat_offsets=(rho0, p0))


class Corrector:
    """A network sized for one coarse block, with its normalisation."""

    def __init__(self, app: GamApp, norm: Normalisation, seed: int = 0, params: Optional[np.ndarray] = None):
        from kaiju_amd.net import GamNet, init_params
        self.app, self.norm = app, norm
        self.net = GamNet(*app.G)
        # the net keeps its own stream; a step waits for its inference only
        # before the kernels that read the correction, so the inference
        # overlaps the step's predictor, bFlux2Fld and Mom2Face
        app.set_correction_wait(self.net.infer_event())
        self.net.set_params(init_params(seed, out_scale=0.02) if params is None else params)
        self.net.set_scales(*norm.feat_scales, gamma=GAMMA)
        self.net.set_offsets(*norm.feat_offsets)
        self.precision = app.spec.model.precision
        self.zero = DeviceArray.of(np.zeros(int(np.prod(app.G)) * 8, np.float32))

    def close(self):
        if getattr(self.app, "h", None) is not None:   # the block forgets the event before the net frees it
            self.app.set_correction_wait(None)
        self.net.close()
        self.zero.free()

    def infer(self):
        """Correction field for the app's current State (device, FP32 [G][8]),
        stored on the active cells, the only ones the step reads
        (gam_net_infer: LDS-staged layers, no activations kept for training).
        Queued on the net's stream after the block's queued work; the step
        waits for it where it reads the correction."""
        self.app.sync()
        g, _, b = self.app.state_device(0)
        self.net.infer(g, b, self.precision, self.app.nG)
        return self.net.output_ptr()

    def forward(self):
        """The same field through the training forward, which keeps every
        layer's activations for gam_net_backward."""
        self.app.sync()
        g, _, b = self.app.state_device(0)
        self.net.features_from_device(g, b, self.precision)
        self.net.forward()
        self.net.sync()
        return self.net.output_ptr()

    def _loss(self, smp: Sample, dt: float, o_ptr=None, backward=False) -> float:
        self.app.put_state_device(0, smp.t, smp.gas.ptr, smp.mag.ptr)
        o = self.forward() if o_ptr is None else o_ptr
        self.app.correction_target_device(smp.tg.ptr, smp.tm.ptr)
        dbuf = self.net.grad_buffer_ptr()
        loss = self.app.correction_grad(o, self.norm.scale, dt, self.norm.weights, dbuf)
        if backward and math.isfinite(loss):
            self.net.backward(dbuf)
        return loss

    def train(self, pool: List[Sample], dt: float, iters: int, lr: float = 1e-3, batch: int = 4, seed: int = 0,
              log=None) -> List[float]:
        rng = np.random.default_rng(seed)
        hist = []
        for it in range(1, iters + 1):
            self.net.zero_grad()
            losses = [self._loss(pool[rng.integers(len(pool))], dt, backward=True) for _ in range(batch)]
            good = [l for l in losses if math.isfinite(l)]
            if not good:
                raise RuntimeError("every sample of a minibatch has a non-finite loss")
            self.net.adam(lr * (0.1 ** (it / iters)), self.step + it)   # lr decays 10x over a phase
            hist.append(float(np.mean(good)))
            if it % 100 == 0 and not np.isfinite(self.net.get_params()[0]).all():
                raise RuntimeError(f"network parameters became non-finite at iteration {it}")
            if log and (it % max(1, iters // 10) == 0 or it == 1):
                log(f"[learned] iter {it} loss {hist[-1]:.4f}")
        self.step += iters
        return hist

    step = 0

    def mean_loss(self, samples: List[Sample], dt: float, zero: bool = False) -> float:
        """Mean loss over samples at the current parameters (zero: of o = 0)."""
        return float(np.mean([self._loss(s, dt, self.zero.ptr if zero else None) for s in samples]))


def face_sq(tm: np.ndarray, N: int, nG: int) -> float:
    """Sum of squares over the active faces (normal index in [nG, nG+N], others in [nG, nG+N))."""
    a, b = slice(nG, nG + N), slice(nG, nG + N + 1)
    return float(np.sum(tm[b, a, a, 0] ** 2) + np.sum(tm[a, b, a, 1] ** 2) + np.sum(tm[a, a, b, 2] ** 2))


def rollout(app: GamApp, tr: Trajectory, steps: int, corr: Optional[Corrector] = None) -> List[Dict[str, float]]:
    """Coarse run from sample 0 (oState = State), errors against the samples."""
    N, nG = tr.N, tr.nG
    dx = 1.0 / N
    app.set_correction(None)
    app.put_state(0, 0.0, tr.gas[0], tr.mag[0], None)
    app.put_state(1, -tr.dt, tr.gas[0], tr.mag[0], None)
    set_time(app, 0.0, tr.dt, 0)
    errs = []
    try:
        if corr is not None:   # the field's address is fixed; each step's inference is queued before it
            app.set_correction(corr.net.output_ptr(), corr.norm.scale)
        for n in range(steps):
            if corr is not None:
                corr.infer()
            st = app.steps(1)
            if st != 0:
                raise RuntimeError(f"coarse rollout halted with status {st} at step {n}")
            s = app.get_state(0)
            errs.append(l1_errors(s["Gas"], s["magFlux"], tr.gas[n + 1], tr.mag[n + 1], nG, dx))
    finally:
        app.set_correction(None)
    return errs


def one_step_errors(app: GamApp, tr: Trajectory, corr: Optional[Corrector] = None) -> List[Dict[str, float]]:
    """Per sample n: L1 errors of one coarse step from the downsampled states
    against sample n+1, with or without the correction."""
    out = []
    try:
        for n in range(len(tr.gas) - 1):
            load_pair(app, tr, n)
            app.set_correction(None if corr is None else corr.infer(), 0.0 if corr is None else corr.norm.scale)
            if app.steps(1) != 0:
                raise RuntimeError(f"coarse step from sample {n} halted")
            s = app.get_state(0)
            out.append(l1_errors(s["Gas"], s["magFlux"], tr.gas[n + 1], tr.mag[n + 1], tr.nG, 1.0 / tr.N))
    finally:
        app.set_correction(None)
    return out


def var_norm(tr: Trajectory) -> Dict[str, float]:
    """Per-variable scale for error sums: mean |x - mean x| of the reference,
    averaged over its samples (the density starts uniform)."""
    P = [primitives(g, cell_b(m, tr.nG, 1.0 / tr.N), tr.nG) for g, m in zip(tr.gas, tr.mag)]
    return {k: float(np.mean([np.mean(np.abs(p[k] - p[k].mean())) for p in P])) + 1e-12 for k in PRIM}


def experiment(N: int = 32, train_seeds=(0, 1, 2, 3), test_seed: int = 100, steps: int = 24, iters: int = 1500,
               lr: float = 1e-3, batch: int = 4, precision: int = 0, dagger: int = 0, log=None,
               ic: Optional[dict] = None) -> dict:
    """Train on the train seeds' trajectories (then `dagger` rounds on the
    corrected model's own rollouts), evaluate on a held-out seed.  `ic`
    overrides TurbIC's amplitudes and wave numbers."""
    log = log or (lambda *a: None)
    ic = ic or {}
    ics = [TurbIC(seed=s, **ic) for s in train_seeds]
    ic_test = TurbIC(seed=test_seed, **ic)
    dt = min(choose_dt(x, N) for x in ics + [ic_test])
    t0 = time.perf_counter()
    trs = [fine_trajectory(x, N, steps, dt, log=log) for x in ics]
    test = fine_trajectory(ic_test, N, steps, dt, log=log)
    t_fine = time.perf_counter() - t0
    log(f"[learned] fine trajectories {t_fine:.1f} s")
    app = coarse_app(N, precision)
    pool, test_samples, corr = [], [], None
    try:
        tgts = []
        for tr in trs:
            tgts.append(trajectory_targets(app, tr))
            log(f"[learned] targets of {len(tgts)}/{len(trs)} training trajectories")
        norm = normalisation(trs, tgts)
        for tr, tl in zip(trs, tgts):
            pool += trajectory_samples(tr, tl)
        del tgts
        test_samples = trajectory_samples(test, trajectory_targets(app, test))
        corr = Corrector(app, norm)
        L0 = corr.mean_loss(test_samples, dt, zero=True)
        t0 = time.perf_counter()
        hist = corr.train(pool, dt, iters, lr=lr, batch=batch, log=log)
        n_traj = len(pool)
        cap = 50.0 * corr.mean_loss(pool, dt, zero=True)
        for r in range(dagger):
            for tr in trs:
                for smp in rollout_samples(app, tr, corr, steps):
                    if corr.mean_loss([smp], dt, zero=True) <= cap:   # drop states far off the data
                        pool.append(smp)
                    else:
                        smp.free()
            log(f"[learned] aggregation round {r + 1}: {len(pool)} samples")
            hist += corr.train(pool, dt, max(1, iters // 2), lr=lr, batch=batch, seed=r + 1, log=log)
        t_train = time.perf_counter() - t0
        n_pool = len(pool)
        L1 = corr.mean_loss(test_samples, dt)
        os_plain = one_step_errors(app, test)
        os_corr = one_step_errors(app, test, corr)
        plain = rollout(app, test, steps)
        fixed = rollout(app, test, steps, corr)
    finally:
        for smp in pool + test_samples:
            smp.free()
        if corr is not None:
            corr.close()
        app.close()
    vn = var_norm(test)
    op = [l1_total(e, vn) for e in os_plain]
    oc = [l1_total(e, vn) for e in os_corr]
    ep = [l1_total(e, vn) for e in plain]
    ec = [l1_total(e, vn) for e in fixed]
    n_iter = len(hist)
    return dict(N=N, steps=steps, iters=iters, dagger=dagger, batch=batch, precision=precision, dt=dt,
                train_seeds=list(train_seeds), test_seed=test_seed,
                loss_start=hist[0], loss_end=float(np.mean(hist[-max(1, n_iter // 20):])),
                samples_trajectory=n_traj, samples_total=n_pool,
                test_loss_zero=L0, test_loss=L1, test_loss_ratio=L1 / L0,
                one_step_gain=float(np.mean(op) / np.mean(oc)),
                one_step_gain_per_var={k: float(np.mean([e[k] for e in os_plain]) / np.mean([e[k] for e in os_corr]))
                                       for k in PRIM},
                err_plain_final=ep[-1], err_corrected_final=ec[-1],
                err_plain_mean=float(np.mean(ep)), err_corrected_mean=float(np.mean(ec)),
                gain_final=ep[-1] / ec[-1], gain_mean=float(np.mean(ep) / np.mean(ec)),
                rollout_gain_per_step=[round(a / b, 4) for a, b in zip(ep, ec)],
                plain_per_var=plain[-1], corrected_per_var=fixed[-1],
                fine_s=round(t_fine, 2), train_s=round(t_train, 2), ms_per_iter=round(t_train / n_iter * 1e3, 3))


def _timed_steps(app: GamApp, tr: Trajectory, n: int, corr: Optional[Corrector] = None) -> float:
    """Mean wall time of n coarse steps from sample 0 (with the correction
    inferred and applied before each step when corr is given)."""
    load_pair(app, tr, 0)
    app.sync()
    if corr is not None:
        app.set_correction(corr.net.output_ptr(), corr.norm.scale)
    t0 = time.perf_counter()
    try:
        for _ in range(n):
            if corr is not None:
                corr.infer()
            if app.steps(1) != 0:
                raise RuntimeError("timed coarse step halted")
        app.sync()
    finally:
        app.set_correction(None)
    return (time.perf_counter() - t0) / n


def equal_error_leg(N: int = 128, train_seeds=(0, 1), test_seed: int = 100, steps: int = 8, train_steps: int = 8,
                    iters: int = 600, batch: int = 4, log=None) -> dict:
    """The learned-coarse form at equal L1 error.  A network is trained on the
    GPU on the solver's (2N)^3 FP64 trajectories of the training seeds,
    downsampled 2x; on a held-out seed, N^3 FP64 runs with and without the
    correction are compared with the (2N)^3 solution over `steps` coarse steps.
    The plain scheme's convergence order p comes from the same comparison at
    (N/2)^3 (reference downsampled 4x, steps of 2 dt), so the plain resolution
    whose error equals the corrected run's is N_eq = N (E_plain/E_corr)^(1/p),
    and the plain FP64 cost at that error is the measured N^3 step time times
    (N_eq/N)^4 (zones x steps for the same physical time).  Errors: sum over
    rho, v, p, B of the active-cell L1 error relative to each variable's mean
    fluctuation, at the last step."""
    log = log or (lambda *a: None)
    ics = [TurbIC(seed=s) for s in train_seeds]
    ic_test = TurbIC(seed=test_seed)
    dt = min(choose_dt(x, N) for x in ics + [ic_test])
    t0 = time.perf_counter()
    trs = [fine_trajectory(x, N, train_steps, dt) for x in ics]
    test = fine_trajectory(ic_test, N, steps, dt)
    t_fine = time.perf_counter() - t0
    log(f"[learned] fine {2 * N}^3 trajectories: {t_fine:.1f} s")
    app = coarse_app(N)
    pool, corr = [], None
    try:
        tgts = [trajectory_targets(app, tr) for tr in trs]
        norm = normalisation(trs, tgts)
        for tr, tl in zip(trs, tgts):
            pool += trajectory_samples(tr, tl)
        del tgts
        corr = Corrector(app, norm)
        t0 = time.perf_counter()
        corr.train(pool, dt, iters, batch=batch)
        app.sync()
        t_train = time.perf_counter() - t0
        log(f"[learned] training {iters} minibatches of {batch}: {t_train:.1f} s")
        vn = var_norm(test)
        e_plain = l1_total(rollout(app, test, steps)[-1], vn)
        e_corr = l1_total(rollout(app, test, steps, corr)[-1], vn)
        t_plain = _timed_steps(app, test, 20)
        t_corr = _timed_steps(app, test, 20, corr)
    finally:
        for smp in pool:
            smp.free()
        if corr is not None:
            corr.close()
        app.close()
    # plain at (N/2)^3, steps of 2 dt, against the reference downsampled once more
    half = [downsample(g, m, test.nG) for g, m in zip(test.gas[::2], test.mag[::2])]
    tr2 = Trajectory(N=N // 2, nG=test.nG, dt=2 * dt, gas=[g for g, _ in half], mag=[m for _, m in half])
    with coarse_app(N // 2) as app2:
        e_half = l1_total(rollout(app2, tr2, steps // 2)[-1], vn)
    p = math.log2(e_half / e_plain) if e_half > e_plain > 0 else float("nan")
    n_eq = N * (e_plain / e_corr) ** (1.0 / p) if p > 0 and e_corr > 0 else float("nan")
    cost_eq = t_plain * (n_eq / N) ** 4
    zones = N ** 3
    log(f"[learned] L1 at {N}^3: plain {e_plain:.4f}, corrected {e_corr:.4f}; at {N // 2}^3 plain {e_half:.4f}; "
        f"order {p:.2f}, N_eq {n_eq:.1f}; step {t_plain * 1e3:.3f} ms plain, {t_corr * 1e3:.3f} ms corrected")
    return dict(workload=f"decaying MHD turbulence (Fourier modes k <= 4 in v and A, rms 1, rho = p = 1), {N}^3 FP64 "
                         f"corrected by the 3-layer BF16 conv net, trained on {len(train_seeds)} seeds' {2 * N}^3 FP64 "
                         f"trajectories ({train_steps} coarse steps each, {iters} minibatches of {batch}), judged on "
                         f"held-out seed {test_seed} against its {2 * N}^3 FP64 solution over {steps} steps",
                value=round(zones / t_corr, 1), unit="zone-updates/s",
                ms_per_step_corrected=round(t_corr * 1e3, 4), ms_per_step_plain=round(t_plain * 1e3, 4),
                l1_vs_fine_plain=round(e_plain, 5), l1_vs_fine_corrected=round(e_corr, 5),
                l1_vs_fine_plain_half=round(e_half, 5), plain_order=round(p, 3),
                plain_cells_at_equal_error=round(n_eq, 2),
                plain_fp64_ms_at_equal_error=round(cost_eq * 1e3, 4),
                speedup_at_equal_error=round(cost_eq / t_corr, 4),
                fine_s=round(t_fine, 1), train_s=round(t_train, 1))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=32)
    ap.add_argument("--train-seeds", default="0,1,2,3")
    ap.add_argument("--test-seed", type=int, default=100)
    ap.add_argument("--steps", type=int, default=24)
    ap.add_argument("--iters", type=int, default=1500)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--dagger", type=int, default=0)
    ap.add_argument("--fp32", action="store_true")
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)
    r = experiment(a.N, tuple(int(s) for s in a.train_seeds.split(",")), a.test_seed, a.steps, a.iters, a.lr,
                   a.batch, 1 if a.fp32 else 0, a.dagger, log)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
