"""Host-side pieces of the torch restatement of Ozaki Scheme I (tests/ozaki1_torch.py) on CPU tensors: the
exact power-of-two scaling against math.ldexp (one correctly rounded step, including
subnormal results, underflow to zero and overflow), and the whole scheme on CPU torch
bit-identical to the exact oracle (oracle/ozaki1.py)."""
import math

import numpy as np
import pytest
import torch

from conftest import make_mat
from oracle import ozaki1 as O1
import ozaki1_torch as ozaki1


def test_ldexp_matches_math(rng):
    vals = [1.0, -1.0, 3.0, 2.0 ** 106 - 1, -(2.0 ** 80 + 12345), 0.0, 5.0, 7.0 * 2 ** 60 + 1]
    vals += [float(int(v)) for v in rng.integers(-2 ** 62, 2 ** 62, 64)]
    exps = [-1250, -1200, -1180, -1130, -1110, -1076, -1075, -1074, -1073, -1060, -1023, -1022,
            -1021, -500, -1, 0, 1, 700, 917, 960, 1000, 1023, 1024, 1100, 1200, 1300]
    v = torch.tensor([a for a in vals for _ in exps], dtype=torch.float64)
    e = torch.tensor([b for _ in vals for b in exps], dtype=torch.int64)
    got = ozaki1._ldexp(v, e).numpy()

    def ref(a, b):
        try:
            return math.ldexp(a, b)
        except OverflowError:
            return math.copysign(math.inf, a)
    want = np.array([ref(a, b) for a in vals for b in exps])
    assert np.array_equal(got.view(np.int64), want.view(np.int64))


def test_pow2_every_normal_exponent():
    e = torch.arange(-1022, 1024, dtype=torch.int64)
    want = np.array([math.ldexp(1.0, int(b)) for b in range(-1022, 1024)])
    assert np.array_equal(ozaki1._pow2(e).numpy(), want)


def extreme_scales(A, B, m, n, k, lda, ldb):
    """rows of A and columns of B scaled so that some products are subnormal, some
    underflow to zero and some are close to the overflow threshold"""
    ra = [-520, 500, -1000, 0, 3]
    cb = [-540, 520, 0, -80]
    for i, e in enumerate(ra[:m]):
        A[i: i + (k - 1) * lda + 1: lda] *= 2.0 ** e
    for j, e in enumerate(cb[:n]):
        B[j * ldb: j * ldb + k] *= 2.0 ** e


@pytest.mark.parametrize("L", [2, 4, 9])
@pytest.mark.parametrize("extreme", [False, True])
def test_cpu_path_matches_oracle(L, extreme, rng):
    m, n, k = 9, 7, 21
    lda, ldb, ldc = m + 2, k + 3, m + 1
    A = make_mat(m, k, lda, "phi", rng, 30.0).reshape(-1, order="F")
    B = make_mat(k, n, ldb, "phi", rng, 30.0).reshape(-1, order="F")
    if extreme:
        extreme_scales(A, B, m, n, k, lda, ldb)
    A[2 + 0 * lda: 3 + (k - 1) * lda: lda] = 0.0
    C = torch.full((ldc * n,), 7.0, dtype=torch.float64)
    ozaki1.gemm(m, n, k, torch.from_numpy(A), lda, torch.from_numpy(B), ldb, C, ldc, L)
    got = C.numpy().reshape(n, ldc)
    want = O1.gemm(m, n, k, A, lda, B, ldb, L).reshape(n, m)
    assert np.array_equal(got[:, :m].view(np.int64), want.view(np.int64))
    assert (got[:, m:] == 7.0).all()
