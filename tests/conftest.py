"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on CPU."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP step)")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle
    oracle.build()
    return oracle.lib()


def pytest_sessionstart(session):
    # KAIJU_EMU_LIB points the binding at the CPU emulation build of the same
    # sources (tests/emu); used to run the GPU tests' logic on a CPU.
    path = os.environ.get("KAIJU_EMU_LIB")
    if path:
        from kaiju_amd import gamera
        gamera.use_library(path)
