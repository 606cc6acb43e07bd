"""Checkpoint/restart of a block (GamApp.save_restart / load_restart): a run
stopped, written, read back into a fresh handle and continued matches the
uninterrupted run bit for bit, in FP64 and FP32 bulk (the CPU emulation runs
the same test)."""
import numpy as np
import pytest

from kaiju_amd import configs
from kaiju_amd.gamera import GamApp
from tests.helpers import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [0, 1])
def test_restart_is_bit_exact(tmp_path, precision):
    spec = configs.blast_wave(16, steps=6, precision=precision)
    say(f"[restart] {spec.name} precision {precision}")
    with GamApp(spec) as a:
        a.InitModel()
        assert a.steps(6) == 0
        sa = a.get_state(0)
        ca = a.get_clock()
    with GamApp(spec) as b:
        b.InitModel()
        assert b.steps(3) == 0
        b.save_restart(str(tmp_path / "r.npz"))
    with GamApp(spec) as c:
        c.load_restart(str(tmp_path / "r.npz"))
        assert c.Model.ts == 3
        assert c.steps(3) == 0
        sc = c.get_state(0)
        cc = c.get_clock()
    for k in ("Gas", "magFlux", "Bxyz"):
        assert np.array_equal(sa[k], sc[k]), k
    assert (ca.t, ca.dt, ca.dt0, ca.ts) == (cc.t, cc.dt, cc.dt0, cc.ts)


def test_restart_rejects_another_shape(tmp_path):
    with GamApp(configs.blast_wave(16, steps=1)) as a:
        a.InitModel()
        a.save_restart(str(tmp_path / "r.npz"))
    with GamApp(configs.blast_wave(8, steps=1)) as b:
        with pytest.raises(ValueError):
            b.load_restart(str(tmp_path / "r.npz"))
