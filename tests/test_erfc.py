"""The flux kernels' complementary error function (kaiju_amd/csrc/gam_erfc.h,
made by kaiju_amd/csrc/erfc_fit.py): erfc(x) = exp(-x^2) P((u-2)/(u+2)),
u = min(|x|, 6), erfc(-x) = 2 - erfc(x).  The coefficients are evaluated as
the kernels do (Horner; the kernels' fused multiply-adds differ by an ulp)
and compared with Python's math.erfc on a dense grid.  The kernels need erfc
to an absolute accuracy: it weights O(1) flux terms."""
import math
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "kaiju_amd", "csrc", "gam_erfc.h")


def coefs(name):
    src = open(HDR).read()
    m = re.search(name + r"\[(\d+)\] = \{([^}]*)\}", src)
    vals = [float(v.rstrip("f")) for v in m.group(2).split(",")]
    assert len(vals) == int(m.group(1))
    return vals


def erfc_model(x, c, dtype):
    u = dtype(min(abs(x), 6.0))
    t = dtype((u - dtype(2)) / (u + dtype(2)))
    p = dtype(c[0])
    for q in c[1:]:
        p = dtype(p * t + dtype(q))
    r = dtype(np.exp(dtype(-x * x))) * p
    return float(dtype(2) - r) if x < 0 else float(r)


def test_fp64_erfc_absolute_error():
    c = coefs("kErfcD")
    xs = np.concatenate([np.linspace(-9, 9, 7201), [0.0, 6.0, -6.0, 1e-300, 5.9999, 26.0, -26.0, 40.0]])
    err = max(abs(erfc_model(float(x), c, np.float64) - math.erfc(float(x))) for x in xs)
    assert err <= 5e-16, err
    # the two branches meet: erfc(x) + erfc(-x) = 2
    for x in (0.3, 1.7, 4.2):
        assert abs(erfc_model(x, c, np.float64) + erfc_model(-x, c, np.float64) - 2.0) <= 5e-16


def test_fp32_erfc_absolute_error():
    c = coefs("kErfcF")
    xs = np.linspace(-9, 9, 3601)
    err = max(abs(erfc_model(float(np.float32(x)), c, np.float32) - math.erfc(float(np.float32(x)))) for x in xs)
    assert err <= 6e-7, err
