"""Ozaki Scheme I baseline of the accuracy benchmark (oz2/ozaki1.py, reference
benchmarks/ozaki1_gpu.hpp) on the GPU: bit-identical to the exact CPU oracle
(oracle/ozaki1.py) on small inputs, including zero rows, padded leading dimensions,
wide exponent ranges and non-finite rows."""
import numpy as np
import pytest
import torch

from conftest import make_mat
from gpu_cases import bits, log, to_dev
from oracle import ozaki1 as O1
from oz2 import ozaki1
from test_ozaki1_host import extreme_scales

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("L", [2, 3, 5, 7, 9])
@pytest.mark.parametrize("dist", ["uniform", "phi", "extreme"])
def test_matches_oracle(L, dist, rng):
    m, n, k = 19, 13, 37
    lda, ldb, ldc = m + 3, k + 1, m + 2
    A = make_mat(m, k, lda, "uniform" if dist == "uniform" else "phi", rng, 2.0).reshape(-1, order="F")
    B = make_mat(k, n, ldb, "uniform" if dist == "uniform" else "phi", rng, 2.0).reshape(-1, order="F")
    if dist == "extreme":  # subnormal, underflowing and near-overflow results
        extreme_scales(A, B, m, n, k, lda, ldb)
    A[6 + 0 * lda: 7 + (k - 1) * lda: lda] = 0.0  # a zero row of A
    log(f"ozaki1 L={L} {dist}")
    dC = to_dev(np.full(ldc * n, 7.0))
    ozaki1.gemm(m, n, k, to_dev(A), lda, to_dev(B), ldb, dC, ldc, L)
    torch.cuda.synchronize()
    C = dC.cpu().numpy().reshape(n, ldc)
    want = O1.gemm(m, n, k, A, lda, B, ldb, L).reshape(n, m)
    assert np.array_equal(bits(C[:, :m]), bits(want))
    assert (C[:, m:] == 7.0).all()  # rows m..ldc-1 untouched


def test_non_finite_row_and_errors():
    m, n, k = 20, 16, 16
    A = np.ones(m * k)
    A[3] = np.inf  # row 3 of A
    B = np.ones(k * n)
    dC = to_dev(np.zeros(m * n))
    ozaki1.gemm(m, n, k, to_dev(A), m, to_dev(B), k, dC, m, 4)
    C = dC.cpu().numpy().reshape(n, m).T
    assert np.isnan(C[3]).all() and np.all(C[np.arange(m) != 3] == k)
    with pytest.raises(ValueError):
        ozaki1.gemm(m, n, k, to_dev(A), m, to_dev(B), k, dC, m, 10)
