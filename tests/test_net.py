"""The learned-correction network's kernels against the CPU reference.

CPU: the reference's own backward equals central finite differences of its
forward (float64, no BF16), which pins the checker.  GPU: each layer of the
MFMA forward equals the reference layer applied to the GPU's own BF16 input
activations; the whole forward and the parameter gradients equal the
reference's (same BF16 rounding points) to BF16-level tolerances; Adam on the
GPU fits a small regression.
"""
import numpy as np
import pytest

from oracle import net_ref
from tests.helpers import say

DIMS = (12, 10, 9)     # ghost-inclusive box of a small block


def _rand_problem(seed=0, dims=DIMS):
    rng = np.random.default_rng(seed)
    nc = int(np.prod(dims))
    w = np.concatenate([np.concatenate([rng.normal(0, np.sqrt(2.0 / (27 * ci)), 27 * co * ci), rng.normal(0, 0.1, co)])
                        for co, ci in zip(net_ref.CO, net_ref.CI)]).astype(np.float32)
    feat = rng.normal(size=(nc, 8)).astype(np.float32)
    dout = rng.normal(size=(nc, 8)).astype(np.float32)
    return w, feat, dout


def test_reference_backward_matches_finite_differences():
    dims = (8, 8, 8)
    w, feat, dout = _rand_problem(1, dims)
    ref = net_ref.NetRef(*dims, w, bf=False)
    ref.forward(feat)
    g = net_ref.flat_grads(ref.backward(dout))
    rng = np.random.default_rng(2)
    w64 = w.astype(np.float64)
    loss = lambda wv: float((net_ref.NetRef(*dims, wv, bf=False).forward(feat) * dout).sum())
    for q in rng.choice(w.size, 24, replace=False):
        h = 1e-5 * max(1.0, abs(w64[q]))
        wp, wm = w64.copy(), w64.copy()
        wp[q] += h
        wm[q] -= h
        fd = (loss(wp) - loss(wm)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 * (1 + abs(fd)), (q, fd, g[q])


@pytest.mark.gpu
def test_gpu_forward_layers_and_output():
    from kaiju_amd.net import GamNet
    w, feat, _ = _rand_problem(3)
    say(f"[net] forward {DIMS}")
    with GamNet(*DIMS) as net:
        net.set_params(w)
        net.set_features(feat)
        net.forward()
        acts = [net.get_act(l) for l in range(3)]
        out = net.get_output().astype(np.float64)
    ref = net_ref.NetRef(*DIMS, w)
    rng = ref.rng
    assert np.array_equal(acts[0][:, :8], net_ref.bf16(feat))
    # each hidden layer from the GPU's own BF16 input: equal up to one BF16 rounding step
    for l in (0, 1):
        W, b = ref.layers[l]
        z = net_ref.conv(acts[l][:, :W.shape[2]], W, ref.off, rng) + b
        want = net_ref.bf16(net_ref.act(z))
        got = acts[l + 1][rng, :W.shape[1]]
        ulp = np.maximum(np.abs(want), 1e-30) * 2.0 ** -7
        assert np.all(np.abs(got - want) <= ulp), (l, float(np.abs(got - want).max()))
    W, b = ref.layers[2]
    z = net_ref.conv(acts[2], W, ref.off, rng) + b
    err = np.abs(out[rng] - z).max() / np.abs(z).max()
    say(f"   output layer from GPU activations: {err:.2e}")
    assert err <= 1e-5
    full = ref.forward(feat)
    errf = np.linalg.norm(out[rng] - full[rng]) / np.linalg.norm(full[rng])
    say(f"   whole forward vs reference: {errf:.2e}")
    assert errf <= 2e-2


@pytest.mark.gpu
@pytest.mark.parametrize("dims,halo,precision", [((21, 18, 17), 4, 0), ((21, 18, 17), 3, 1), ((40, 13, 11), 4, 1),
                                                 ((24, 24, 24), 4, 0)])
def test_gpu_infer_equals_forward_on_its_box(dims, halo, precision):
    """gam_net_infer (features and the three layers, LDS-staged bricks) stores
    exactly gam_net_forward's output on the cells at least `halo` from the
    faces, bit for bit, and nothing elsewhere; grids that are no multiple of
    the brick, FP64 and FP32 states."""
    from kaiju_amd.gamera import DeviceArray
    from kaiju_amd.net import GamNet, HipStepError
    w, _, _ = _rand_problem(5, dims)
    rng = np.random.default_rng(11)
    nc = int(np.prod(dims))
    rho, pr = rng.uniform(0.5, 1.5, nc), rng.uniform(0.5, 2.0, nc)
    v, B = rng.normal(0, 0.5, (3, nc)), rng.normal(0, 1.0, (3, nc))
    gas = np.concatenate([rho, rho * v[0], rho * v[1], rho * v[2], pr / (5.0 / 3.0 - 1) + 0.5 * rho * (v ** 2).sum(0)])
    dt = np.float64 if precision == 0 else np.float32
    say(f"[net] infer vs forward {dims} halo {halo} precision {precision}")
    G, Bd = DeviceArray.of(gas, dt), DeviceArray.of(B.ravel(), dt)
    try:
        with GamNet(*dims) as net:
            net.set_params(w)
            net.set_scales(1.1, 0.7, 1.3, 0.9)
            net.set_offsets(1.0, 1.2)
            for bad in (2, (min(dims) + 1) // 2):   # three stencil layers; a box left on every axis
                with pytest.raises(HipStepError):
                    net.infer(G.ptr, Bd.ptr, precision, bad)
            net.infer(G.ptr, Bd.ptr, precision, halo)
            out_i = net.get_output()
            net.features_from_device(G.ptr, Bd.ptr, precision)
            net.forward()
            out_f = net.get_output()
    finally:
        G.free()
        Bd.free()
    Ni, Nj, Nk = dims
    oi, of = out_i.reshape(Nk, Nj, Ni, 8), out_f.reshape(Nk, Nj, Ni, 8)
    box = (slice(halo, Nk - halo), slice(halo, Nj - halo), slice(halo, Ni - halo))
    assert np.abs(of[box]).max() > 0
    assert np.array_equal(oi[box].view(np.uint32), of[box].view(np.uint32))
    outside = np.ones((Nk, Nj, Ni), bool)
    outside[box] = False
    assert not oi[outside].any()   # the output buffer starts at zero; infer writes the box only


@pytest.mark.gpu
def test_gpu_net_on_a_block_stream():
    """gam_net_set_stream queues the net on a block's stream (gam_stream) and
    back on its own: the same output either way, and the infer event exists."""
    from kaiju_amd import configs
    from kaiju_amd.gamera import GamApp, lib
    from kaiju_amd.net import GamNet
    spec = configs.orszag_tang((16, 16, 8), precision=0)
    with GamApp(spec) as app:
        app.InitModel()
        g, _, b = app.state_device(0)
        w, _, _ = _rand_problem(6, app.G)
        with GamNet(*app.G) as net:
            net.set_params(w)
            assert net.infer_event()
            net.infer(g, b, 0, app.nG)
            own = net.get_output()
            net.set_stream(lib().gam_stream(app.h))
            net.infer(g, b, 0, app.nG)
            shared = net.get_output()
            net.set_stream(None)
            net.infer(g, b, 0, app.nG)
            back = net.get_output()
    assert np.abs(own).max() > 0
    assert np.array_equal(own.view(np.uint32), shared.view(np.uint32))
    assert np.array_equal(own.view(np.uint32), back.view(np.uint32))


@pytest.mark.gpu
def test_gpu_gradients_match_reference():
    from kaiju_amd.net import GamNet
    w, feat, dout = _rand_problem(4)
    say(f"[net] backward {DIMS}")
    with GamNet(*DIMS) as net:
        net.set_params(w)
        net.set_features(feat)
        net.forward()
        net.zero_grad()
        net.backward_host(dout)
        _, g = net.get_params()
    ref = net_ref.NetRef(*DIMS, w)
    ref.forward(feat)
    gr = net_ref.flat_grads(ref.backward(dout))
    o = 0
    for l, (co, ci) in enumerate(zip(net_ref.CO, net_ref.CI)):
        n = 27 * co * ci + co
        e = np.linalg.norm(g[o:o + n] - gr[o:o + n]) / np.linalg.norm(gr[o:o + n])
        say(f"   layer {l} grad rel err {e:.2e}")
        assert e <= 3e-2, (l, e)
        o += n


@pytest.mark.gpu
def test_gpu_adam_fits_a_target_field():
    """Training on the GPU with its own kernels: fit a fixed target output."""
    from kaiju_amd.net import GamNet, init_params
    rng = np.random.default_rng(5)
    nc = int(np.prod(DIMS))
    feat = rng.normal(size=(nc, 8)).astype(np.float32)
    target = np.zeros((nc, 8), np.float32)
    target[:, :5] = 0.5 * np.tanh(feat[:, :5])
    ref = net_ref.NetRef(*DIMS, init_params(0))
    rows = ref.rng
    say("[net] adam regression")
    with GamNet(*DIMS) as net:
        net.set_params(init_params(0))
        net.set_features(feat)
        losses = []
        for t in range(1, 151):
            net.forward()
            out = net.get_output()
            d = np.zeros_like(out)
            d[rows] = 2 * (out[rows] - target[rows]) / rows.size
            losses.append(float(((out[rows] - target[rows]) ** 2).sum() / rows.size))
            net.zero_grad()
            net.backward_host(d)
            if t == 1:
                w0, g0 = net.get_params()
            net.adam(3e-3, t)
            if t == 1:   # first bias-corrected step: w - lr g / (|g| + eps)
                w1, _ = net.get_params()
                want = w0 - np.float32(3e-3) * g0 / (np.abs(g0) + np.float32(1e-8))
                assert np.allclose(w1, want, rtol=1e-5, atol=1e-8)
        say(f"   loss {losses[0]:.3e} -> {losses[-1]:.3e}")
        assert losses[-1] < 0.3 * losses[0]
