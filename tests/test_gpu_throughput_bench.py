"""The throughput benchmark (oz2/bench_ozaki2.py, reference benchmarks/bench_ozaki2.cpp)
on the GPU: its device inputs follow their definition (sin / cos, or the splitmix64 +
Box-Muller phiX entries, whose uniforms are exact and whose transcendental functions
agree to a few ulp), and a small sweep prints one line per size and s with the
native-DGEMM check, under a tuning override."""
import io

import numpy as np
import pytest
import torch

from bench_inputs import phi_matrix
from oz2 import bench_ozaki2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rows,cols", [(37, 29), (5, 1)])
def test_fill_matches_definition(rows, cols):
    r = np.arange(rows)[:, None]
    c = np.arange(cols)[None, :]
    X = torch.empty(rows * cols, dtype=torch.float64, device="cuda")
    bench_ozaki2.fill(X, rows, cols, True, -1.0, 1)
    np.testing.assert_allclose(X.cpu().numpy(), np.sin(r + c).reshape(-1, order="F"), rtol=0, atol=4e-16)
    bench_ozaki2.fill(X, rows, cols, False, -1.0, 12345)
    np.testing.assert_allclose(X.cpu().numpy(), np.cos(r * c * 1.0).reshape(-1, order="F"), rtol=0, atol=4e-16)
    for phi, seed in ((0.0, 1), (1.5, 12345)):
        bench_ozaki2.fill(X, rows, cols, True, phi, seed)
        np.testing.assert_allclose(X.cpu().numpy(), phi_matrix(rows, cols, phi, seed), rtol=1e-13, atol=0)


def test_throughput_rows():
    buf = io.StringIO()
    rows = bench_ozaki2.main(["--sizes", "256,300x500x200", "--s", "8,16", "--reps", "2", "--dgemm", "--check",
                              "--dist", "phi1", "--tuning", "batched_gemm=0,pad_lda=0"], out=buf)
    text = buf.getvalue()
    print(text)
    assert "'batched_gemm': False" in text and "'pad_lda': False" in text
    lines = [ln for ln in text.split("\n") if " s=" in ln and "TFLOPS" in ln]
    assert len(lines) == 4 and len(rows) == 4
    for ln, r in zip(lines, rows):
        err = float(ln.split("max|C-C_dgemm|/(|A||B|)")[1])
        assert err < (1e-5 if r["s"] == 8 else 1e-14), ln
        assert " dgemm " in ln and r["dgemm"] > 0 and r["best"] <= r["median"]
    assert "vs GPU DGEMM" in text


def test_mfma_rate_harness():
    # bench.py --mfma-rate: the register-only MFMA loop runs on every CU and reports a rate
    # below the nominal peak (a loop that did less work than counted would exceed it)
    import importlib.util
    import os
    spec = importlib.util.spec_from_file_location("bench_main", os.path.join(os.path.dirname(__file__), "..", "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    r = bench.mfma_rate(torch, reps=1)
    assert 1000.0 < r["tops"] <= 1.02 * bench.INT8_PEAK_TOPS, r
