"""Checkpoint/restart of a block: a run stopped, written, read back into a
fresh handle and continued matches the uninterrupted run bit for bit, in FP64
and FP32 bulk, both through GamApp.save_restart / load_restart (.npz) and
through the reference's own restart file (write_restart_h5 /
read_restart_h5: RunID.gam.Res.#####.h5, which carries the grid too).  The
CPU emulation runs the same tests."""
import os

import numpy as np
import pytest

from kaiju_amd import configs
from kaiju_amd import gamio
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


@pytest.mark.skipif(not os.path.exists(gamio.IO_LIB), reason="libgamera_io.so not built (no HDF5 C library)")
@pytest.mark.parametrize("precision", [0, 1])
def test_reference_format_restart_is_bit_exact(tmp_path, precision):
    spec = configs.blast_wave(16, steps=6, precision=precision)
    say(f"[restart h5] {spec.name} precision {precision}")
    with GamApp(spec) as a:
        a.InitModel()
        assert a.steps(6) == 0
        sa = a.get_state(0)
        ca = a.get_clock()
    path = str(tmp_path / gamio.restart_name("bw16", 0))
    with GamApp(spec) as b:
        b.InitModel()
        assert b.steps(3) == 0
        b.Model.nOut, b.Model.nRes = 2, 0
        b.write_restart_h5(path)
    with GamApp(spec) as c:
        ioc = c.read_restart_h5(path)   # corners from the file, not from the spec
        assert (c.Model.ts, c.Model.nOut, c.Model.nRes) == (3, 2, 1)
        assert ioc.ts == 3
        assert c.steps(3) == 0
        sc = c.get_state(0)
        cc = c.get_clock()
    for k in ("Gas", "magFlux", "Bxyz"):
        assert np.array_equal(sa[k], sc[k]), k
    assert (ca.t, ca.dt, ca.dt0, ca.ts) == (cc.t, cc.dt, cc.dt0, cc.ts)


@pytest.mark.skipif(not os.path.exists(gamio.IO_LIB), reason="libgamera_io.so not built (no HDF5 C library)")
def test_grid_file_feeds_the_metric(tmp_path):
    """A grid file (readH5Grid's X/Y/Z) gives the same metric as the corners
    it was written from."""
    spec = configs.blast_wave(12, steps=1)
    path = str(tmp_path / "bw12_grid.h5")
    with GamApp(spec) as a:
        a.InitModel()
        gamio.write_grid(path, a._xyz, a.nG)
        ga = a.get_grid()
    assert gamio.grid_dims(path) == spec.N
    with GamApp(spec) as b:
        b.InitModel(corners=gamio.read_grid(path))
        gb = b.get_grid()
    for k in ga:
        assert np.array_equal(ga[k], gb[k]), k
