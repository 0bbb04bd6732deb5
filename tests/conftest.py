import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "os2-xe_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI library)")
    config.addinivalue_line("markers", "slow: larger CPU cases")


def make_mat(rows, cols, ld, dist, rng, phi=1.0):
    """Column-major test matrix with ld padding filled with 1e300 (tests/test_ozaki2.cpp:29-52)."""
    out = np.full((ld, cols), 1e300, order="F")
    if dist == "uniform":
        v = rng.uniform(-0.5, 0.5, (rows, cols))
    elif dist == "phi":
        v = rng.uniform(-0.5, 0.5, (rows, cols)) * np.exp(phi * rng.standard_normal((rows, cols)))
    elif dist == "ints":
        v = rng.integers(-1000, 1001, (rows, cols)).astype(np.float64)
    elif dist == "maxmag":
        v = np.full((rows, cols), 1.0 - 2.0 ** -53)
    elif dist == "sincos":
        r = np.arange(rows)[:, None]
        c = np.arange(cols)[None, :]
        v = np.sin((r + 3 * c).astype(np.float64))
    elif dist == "u11":
        v = rng.uniform(-1.0, 1.0, (rows, cols))
    elif dist == "wide":
        e = rng.integers(-20, 21, (rows, cols))
        sgn = rng.choice([-1.0, 1.0], (rows, cols))
        v = (1.0 + rng.random((rows, cols))) * sgn * np.exp2(e.astype(np.float64))
    else:
        raise ValueError(dist)
    out[:rows, :] = v
    return out


def make_cmat(rows, cols, ld, dist, rng, phi=1.0):
    re = make_mat(rows, cols, ld, dist, rng, phi)
    im = make_mat(rows, cols, ld, dist, rng, phi)
    out = np.empty((ld, cols), dtype=np.complex128, order="F")
    out.real = re
    out.imag = im
    return out


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
