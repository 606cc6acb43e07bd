"""CPU reference of the learned-correction network (kaiju_amd/csrc/gam_net.hip).

TEST INFRASTRUCTURE ONLY: the checker for the GPU network's forward and
backward kernels.  The network has no counterpart in the reference; what pins
this restatement is (a) its backward against central finite differences of its
own forward in float64 (tests/test_net.py) and (b) the layouts and
rounding points documented in include/gamera_net.h.

Float64 arithmetic, with BF16 rounding (round to nearest even) at the points
where the GPU stores BF16: input features, weights, hidden activations, and the
upstream gradients of each backward layer.  Hidden activation: leaky ReLU with
slope 1/8 for negative inputs.
"""
from __future__ import annotations

import numpy as np

CO = (32, 32, 8)
CI = (8, 32, 32)
TAPS = 27
LEAK = 0.125


def num_params() -> int:
    return sum(TAPS * co * ci + co for co, ci in zip(CO, CI))


def unpack(w: np.ndarray):
    """FP32 parameter vector -> [(W[27][co][ci], b[co]) for the 3 layers]."""
    out, o = [], 0
    for co, ci in zip(CO, CI):
        n = TAPS * co * ci
        out.append((w[o:o + n].reshape(TAPS, co, ci).astype(np.float64), w[o + n:o + n + co].astype(np.float64)))
        o += n + co
    return out


def act(z: np.ndarray) -> np.ndarray:
    """Hidden activation: leaky ReLU, slope 1/8 below zero."""
    return np.where(z > 0, z, LEAK * z)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def geometry(Ni: int, Nj: int, Nk: int):
    """(tap offsets, lo, ntile) of the GPU kernels for a ghost-inclusive box."""
    nc = Ni * Nj * Nk
    O = Ni * Nj + Ni + 1
    ntile = (nc - 2 * O) // 16
    off = []
    for t in range(TAPS):
        dx, dy, dz = t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1
        off.append(dx + Ni * (dy + Nj * dz))
    return np.array(off), O, ntile


def conv(a: np.ndarray, W: np.ndarray, off, rng):
    """out[p, co] = sum_tap sum_ci a[p + off[tap], ci] W[tap, co, ci] for p in rng."""
    out = np.zeros((len(rng), W.shape[1]))
    for t in range(TAPS):
        out += a[rng + off[t]] @ W[t].T
    return out


class NetRef:
    def __init__(self, Ni, Nj, Nk, w: np.ndarray, bf: bool = True):
        self.dims = (Ni, Nj, Nk)
        self.nc = Ni * Nj * Nk
        self.off, self.lo, self.ntile = geometry(Ni, Nj, Nk)
        self.rng = np.arange(self.lo, self.lo + 16 * self.ntile)
        self.bf = bf
        self.layers = [((bf16(W) if bf else W), b) for W, b in unpack(w)]

    def r(self, x):
        return bf16(x) if self.bf else np.asarray(x, dtype=np.float64)

    def forward(self, feat: np.ndarray) -> np.ndarray:
        """feat [nc, 8] -> out [nc, 8] (zero outside the computed range)."""
        nc, rng = self.nc, self.rng
        a = self.r(feat)
        self.acts = [a]
        for l, (W, b) in enumerate(self.layers):
            z = conv(a, W, self.off, rng) + b
            if l < 2:
                nxt = np.zeros((nc, W.shape[1]))
                nxt[rng] = self.r(act(z))
                a = nxt
                self.acts.append(a)
            else:
                out = np.zeros((nc, W.shape[1]))
                out[rng] = z
        return out

    def backward(self, dout: np.ndarray):
        """dout [nc, 8] -> [(dW, db)] per layer, using the last forward."""
        nc, rng, off = self.nc, self.rng, self.off
        g = np.zeros_like(dout, dtype=np.float64)   # upstream gradient, on the computed cells only
        g[rng] = self.r(dout[rng])
        grads = [None] * 3
        for l in (2, 1, 0):
            W, b = self.layers[l]
            a = self.acts[l]
            dW = np.zeros_like(W)
            for t in range(TAPS):
                dW[t] = g[rng].T @ a[rng + off[t]]
            grads[l] = (dW, g[rng].sum(axis=0))
            if l == 0:
                break
            # gradient of the layer input: transposed convolution, then the activation's slope
            gi = np.zeros((nc, W.shape[2]))
            for t in range(TAPS):
                gi[rng] += g[rng - off[t]] @ W[t]
            gi[rng] *= np.where(self.acts[l][rng] > 0, 1.0, LEAK)
            g = np.zeros((nc, W.shape[2]))
            g[rng] = self.r(gi[rng])
        return grads


def flat_grads(grads) -> np.ndarray:
    return np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])
