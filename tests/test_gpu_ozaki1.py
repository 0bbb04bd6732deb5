"""Ozaki Scheme I baseline of the accuracy benchmark on the library's kernels
(oz2_ozaki1_dgemm through oz2/ozaki1.py; reference benchmarks/ozaki1_gpu.hpp):
bit-identical to the exact CPU oracle (oracle/ozaki1.py) on small inputs, including
zero rows, padded leading dimensions, wide exponent ranges and non-finite rows; at the
accuracy benchmark's sizes bit-identical to the independent torch restatement
(tests/ozaki1_torch.py, hipBLASLt GEMMs) in every entry and to the oracle on sampled
rows and columns, and across the int32 chunking of long anti-diagonals."""
import numpy as np
import pytest
import torch

from conftest import make_mat
from gpu_cases import bits, log, to_dev
from oracle import ozaki1 as O1
import ozaki1_torch
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


def _sampled_oracle(A, B, m, n, k, L, C, rng, count=6):
    # the exponent of row i depends on row i of A only (column j on column j of B), so the
    # oracle on the sampled rows and columns gives those entries of the full product
    rows = np.unique(np.concatenate([[0, m - 1], rng.integers(0, m, count)]))
    cols = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, count)]))
    As = A.reshape(k, m)[:, rows].reshape(-1)
    Bs = B.reshape(n, k)[cols].reshape(-1)
    want = O1.gemm(rows.size, cols.size, k, As, rows.size, Bs, k, L).reshape(cols.size, rows.size)
    got = C.reshape(n, m)[np.ix_(cols, rows)]
    assert np.array_equal(bits(got), bits(want))


@pytest.mark.parametrize("L", [4, 7, 9])
@pytest.mark.parametrize("dist", ["urand", "phi2"])
def test_benchmark_size_matches_torch_restatement_and_oracle(L, dist):
    # the accuracy benchmark's shape (m = n = 512) at k = 2048
    rng = np.random.default_rng(L * 10 + len(dist))
    m = n = 512
    k = 2048
    if dist == "urand":
        A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    else:
        A = rng.uniform(-0.5, 0.5, m * k) * np.exp(2.0 * rng.standard_normal(m * k))
        B = rng.uniform(-0.5, 0.5, k * n) * np.exp(2.0 * rng.standard_normal(k * n))
    log(f"ozaki1 L={L} {dist} {m}x{n}x{k}")
    dA, dB = to_dev(A), to_dev(B)
    dC, dT = to_dev(np.zeros(m * n)), to_dev(np.zeros(m * n))
    ozaki1.gemm(m, n, k, dA, m, dB, k, dC, m, L)
    ozaki1_torch.gemm(m, n, k, dA, m, dB, k, dT, m, L)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    assert np.array_equal(bits(C), bits(dT.cpu().numpy()))
    _sampled_oracle(A, B, m, n, k, L, C, rng, count=3)


def test_long_anti_diagonals_are_chunked():
    # L = 9 at k = 60000: the longest anti-diagonal spans 9 x 60032 k, more than one int32 chunk
    # of digit products (523264 k), so its sums meet in int64 in the Horner step
    rng = np.random.default_rng(5)
    m, n, k, L = 40, 24, 60000, 9
    A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    A[:m] = -1.0  # full-magnitude digits in the first column of A
    log(f"ozaki1 L={L} {m}x{n}x{k}")
    dA, dB = to_dev(A), to_dev(B)
    dC, dT = to_dev(np.zeros(m * n)), to_dev(np.zeros(m * n))
    ozaki1.gemm(m, n, k, dA, m, dB, k, dC, m, L)
    ozaki1_torch.gemm(m, n, k, dA, m, dB, k, dT, m, L)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    assert np.array_equal(bits(C), bits(dT.cpu().numpy()))
    _sampled_oracle(A, B, m, n, k, L, C, rng, count=0)  # the corners: the Python oracle is slow at this k


def test_limits_and_messages():
    dA = to_dev(np.ones(64))
    with pytest.raises(ValueError, match="oz1: L must be"):
        ozaki1.gemm(8, 8, 8, dA, 8, dA, 8, to_dev(np.zeros(64)), 8, 1)
    with pytest.raises(ValueError, match="oz1: k too large"):
        ozaki1.gemm(1, 1, 1 << 19, dA, 1, dA, 1 << 19, to_dev(np.zeros(1)), 1, 4)
    C = to_dev(np.full(64, 3.0))
    ozaki1.gemm(8, 8, 0, dA, 8, dA, 1, C, 8, 4)  # k = 0: C = 0
    torch.cuda.synchronize()
    assert not C.cpu().numpy().any()
