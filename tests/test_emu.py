"""The HIP kernel sources, compiled for the CPU (tests/emu), checked before any
GPU run: the GPU parity tests pass against the oracle on the emulation, and an
AddressSanitizer build of the same sources runs every kernel in bounds; the
same for the learned-correction kernels of the step library."""
import os
import subprocess
import sys

import pytest

from tests.emu import build_emu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(lib, extra_env, kexpr, timeout, test_file="test_gpu_parity.py"):
    env = dict(os.environ)
    env.update(extra_env)
    env["KAIJU_EMU_LIB"] = lib
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "gpu",
           os.path.join(ROOT, "tests", test_file), "-k", kexpr]
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return r.stdout


def test_kernels_match_oracle_on_cpu_emulation():
    lib = build_emu.build(asan=False)
    out = _run(lib, {}, "bw16_warp_boris or ot32x32x1_25d or bw16_hydro or 8cent or cyl", 600)
    assert " passed" in out


def test_kernels_in_bounds_under_asan():
    lib = build_emu.build(asan=True)
    env = {"LD_PRELOAD": build_emu.asan_runtime(), "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1"}
    out = _run(lib, env, "(test_multistep or test_hatch or test_metric or test_initial) and (bw16_warp_boris or ot32x32x4_8cent or 25d or nolfm or ot24_bc or cyl or odd)", 900)
    assert " passed" in out


def test_correction_kernels_on_cpu_emulation_and_in_bounds():
    lib = build_emu.build(asan=False)
    out = _run(lib, {}, "correction or corrected", 600, "test_learned.py")
    assert " passed" in out
    lib = build_emu.build(asan=True)
    env = {"LD_PRELOAD": build_emu.asan_runtime(), "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1"}
    out = _run(lib, env, "correction or corrected", 900, "test_learned.py")
    assert " passed" in out


def test_ring_kernels_on_cpu_emulation_and_in_bounds():
    lib = build_emu.build(asan=False)
    out = _run(lib, {}, "axis", 600, "test_ring.py")
    assert " passed" in out
    lib = build_emu.build(asan=True)
    env = {"LD_PRELOAD": build_emu.asan_runtime(), "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1"}
    out = _run(lib, env, "axis", 900, "test_ring.py")
    assert " passed" in out


def test_graph_replays_on_cpu_emulation_and_in_bounds():
    """Captured-graph replays: eager pre-steps, a halt then a second launch,
    and the setters a held handle refuses (the emulation records and replays
    the captured launches)."""
    lib = build_emu.build(asan=False)
    # the FP64 replays and the packed graph against the oracle run on the GPU only
    out = _run(lib, {}, "graph and not eager_steps[0] and not on_seeded", 600, "test_ensemble.py")
    assert " passed" in out
    lib = build_emu.build(asan=True)
    env = {"LD_PRELOAD": build_emu.asan_runtime(), "ASAN_OPTIONS": "detect_leaks=0:abort_on_error=1"}
    out = _run(lib, env, "graph_after or graph_run_after or packed_graph_replays_match_eager_steps[1]", 900,
               "test_ensemble.py")
    assert " passed" in out


def test_calcdt_failure_branches_on_cpu_emulation():
    """CalcDT's 10x-drop and dt <= TINY branches against the oracle."""
    lib = build_emu.build(asan=False)
    out = _run(lib, {}, "drop or disaster", 300, "test_calcdt.py")
    assert " passed" in out


def test_restart_files_on_cpu_emulation():
    """Bit-exact resume through the .npz and the reference's HDF5 restart."""
    lib = build_emu.build(asan=False)
    out = _run(lib, {}, "restart or grid_file", 600, "test_restart.py")
    assert " passed" in out
