"""The GPU double-double GEMM of the measurement harness (liboz2_check.so,
oz2chk_dd_gemm) against the oracle's double-double GEMM, bit for bit in hi and
lo, for every op and padded leading dimensions.  Every full-size ulp and
correctly-rounded figure, and every training label of the moduli selector,
comes from the GPU form; the oracle's form is pinned by exact Fraction sums
(tests/test_oracle.py) and follows the reference's accumulation
(tests/test_ozaki2.cpp:56-84)."""
import itertools

import numpy as np
import pytest
import torch

from conftest import make_cmat, make_mat
from gpu_cases import bits, flat, log, to_dev
from oracle import oracle as O
from oz2 import check

pytestmark = pytest.mark.gpu
CODE = {"N": 0, "T": 1, "C": 2}


def run_dd(ta, tb, m, n, k, A, lda, B, ldb, cplx):
    w = 2 if cplx else 1
    hi = torch.empty(w * m * n, dtype=torch.float64, device="cuda")
    lo = torch.empty_like(hi)
    dA, dB = to_dev(A), to_dev(B)  # held until the kernel has run
    rc = check.lib().oz2chk_dd_gemm(None, CODE[ta], CODE[tb], m, n, k, dA.data_ptr(), lda, dB.data_ptr(), ldb,
                                    hi.data_ptr(), lo.data_ptr(), int(cplx))
    assert rc == 0
    torch.cuda.synchronize()
    return hi.cpu().numpy(), lo.cpu().numpy()


@pytest.mark.parametrize("ta,tb", list(itertools.product("NT", "NT")))
@pytest.mark.parametrize("dist", ["u11", "phi", "wide"])
def test_dd_gemm_real_matches_oracle(ta, tb, dist):
    rng = np.random.default_rng([CODE[ta], CODE[tb], len(dist), ord(dist[0])])
    for m, n, k in ((37, 29, 300), (130, 65, 1000), (1, 1, 1)):
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb = ar + 3, br + 5
        A, B = flat(make_mat(ar, ac, lda, dist, rng, phi=2.0)), flat(make_mat(br, bc, ldb, dist, rng, phi=2.0))
        A[:lda] = 0.0  # a zero column of A
        log(f"dd real {ta}{tb} {m}x{n}x{k} {dist}")
        hi, lo = run_dd(ta, tb, m, n, k, A, lda, B, ldb, False)
        whi, wlo = O.dd_gemm(ta, tb, m, n, k, A, lda, B, ldb)
        assert np.array_equal(bits(hi), bits(whi)) and np.array_equal(bits(lo), bits(wlo))
        # the harness's wrapper (used by the benchmark and the selector's training) is the same call
        dA, dB = to_dev(A), to_dev(B)
        dhi, dlo = check.dd_gemm(ta, tb, m, n, k, dA, lda, dB, ldb)
        torch.cuda.synchronize()
        assert np.array_equal(bits(dhi.cpu().numpy()), bits(whi)) and np.array_equal(bits(dlo.cpu().numpy()), bits(wlo))


@pytest.mark.parametrize("ta,tb", list(itertools.product("NTC", "NTC")))
def test_dd_gemm_complex_matches_oracle(ta, tb):
    rng = np.random.default_rng(CODE[ta] * 3 + CODE[tb])
    m, n, k = 23, 31, 150
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb = ar + 7, br + 2
    A = np.ascontiguousarray(flat(make_cmat(ar, ac, lda, "phi", rng))).view(np.float64)
    B = np.ascontiguousarray(flat(make_cmat(br, bc, ldb, "u11", rng))).view(np.float64)
    log(f"dd complex {ta}{tb} {m}x{n}x{k}")
    hi, lo = run_dd(ta, tb, m, n, k, A, lda, B, ldb, True)
    whi, wlo = O.dd_zgemm(ta, tb, m, n, k, A, lda, B, ldb)
    assert np.array_equal(bits(hi), bits(whi)) and np.array_equal(bits(lo), bits(wlo))


def test_dd_gemm_large_sampled():
    # a benchmark-sized product on sampled entries (oracle dd_entries, the same accumulation)
    rng = np.random.default_rng(3)
    m, n, k = 1500, 1300, 4096
    A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    hi, lo = run_dd("N", "N", m, n, k, A, m, B, k, False)
    rows, cols = rng.integers(0, m, 300), rng.integers(0, n, 300)
    rows[:2], cols[:2] = [0, m - 1], [n - 1, 0]
    whi, wlo = O.dd_entries("N", "N", k, A, m, B, k, rows, cols)
    at = rows + cols * m
    assert np.array_equal(bits(hi[at]), bits(whi)) and np.array_equal(bits(lo[at]), bits(wlo))


@pytest.fixture(scope="module")
def oz_handle():
    import oz2
    h = oz2.Ozaki2()
    yield h
    h.close()


@pytest.mark.parametrize("dist", ["u11", "phi", "wide"])
def test_bound_ratio_matches_oracle(oz_handle, dist):
    """check.bound_ratio (the bench's form of the reference's element-wise bound, on the GPU)
    against the oracle's oz2o_bound_ratio on the same product, including a row and a column of
    zeros (exponent 0) and a result with a planted error past the bound"""
    rng = np.random.default_rng([7, len(dist)])
    for (m, n, k, s) in ((37, 29, 300, 12), (130, 65, 1000, 16), (64, 40, 70, 20)):
        log(f"bound ratio {dist} m={m} n={n} k={k} s={s}")
        A = make_mat(m, k, m + 3, dist, rng)
        B = make_mat(k, n, k, dist, rng)
        A[5, :] = 0.0
        B[:, 3] = 0.0
        A, B = flat(A), flat(B)
        lda, ldb, ldc = m + 3, k, m + 1
        dA, dB = to_dev(A), to_dev(B)
        dC = torch.zeros(ldc * n, dtype=torch.float64, device="cuda")
        oz_handle.gemm("N", "N", m, n, k, 1.0, dA, lda, dB, ldb, 0.0, dC, ldc, s)
        torch.cuda.synchronize()
        hi, lo = check.dd_gemm("N", "N", m, n, k, dA, lda, dB, ldb)
        tk = O.precision_bits(k, s)
        for planted in (False, True):
            if planted:  # one entry moved far past its bound
                dC[ldc + 2] += 1e-3 * max(1.0, abs(float(dC[ldc + 2])))
            got = check.bound_ratio(m, n, k, dA, lda, dB, ldb, dC, ldc, hi, lo, tk)
            want = O.bound_ratio("N", "N", m, n, k, A, lda, B, ldb, 1.0, 0.0, np.zeros(ldc * n), dC.cpu().numpy(),
                                 ldc, hi.cpu().numpy(), lo.cpu().numpy(), s)
            assert got == pytest.approx(want, rel=1e-12), (planted, got, want)
            assert (got > 1.0) == planted


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("T", "C")])
def test_zbound_ratio_matches_oracle(oz_handle, ta, tb):
    """check.zbound_ratio against the oracle's oz2o_zbound_ratio (alpha = 1, beta = 0), with a
    planted error past the bound"""
    rng = np.random.default_rng([11, CODE[ta], CODE[tb]])
    for (m, n, k, s) in ((23, 31, 200, 14), (70, 40, 513, 18)):
        log(f"zbound ratio {ta}{tb} m={m} n={n} k={k} s={s}")
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = ar + 2, br + 1, m + 3
        A = np.ascontiguousarray(flat(make_cmat(ar, ac, lda, "phi", rng))).view(np.float64)
        B = np.ascontiguousarray(flat(make_cmat(br, bc, ldb, "phi", rng))).view(np.float64)
        dA, dB = to_dev(A), to_dev(B)
        dC = torch.zeros(2 * ldc * n, dtype=torch.float64, device="cuda")
        oz_handle.zgemm(ta, tb, m, n, k, 1.0, dA, lda, dB, ldb, 0.0, dC, ldc, s)
        torch.cuda.synchronize()
        hi_h, lo_h = run_dd(ta, tb, m, n, k, A, lda, B, ldb, True)
        hi, lo = to_dev(hi_h), to_dev(lo_h)
        tk = O.precision_bits(2 * k, s)
        for planted in (False, True):
            if planted:
                dC[2 * (ldc + 1) + 1] += 1e-3 * max(1.0, abs(float(dC[2 * (ldc + 1) + 1])))
            got = check.zbound_ratio(ta, tb, m, n, k, dA, lda, dB, ldb, dC, ldc, hi, lo, tk)
            want = O.zbound_ratio(ta, tb, m, n, k, A, lda, B, ldb, 1 + 0j, 0j, np.zeros(2 * ldc * n),
                                  dC.cpu().numpy(), ldc, hi_h, lo_h, s)
            assert got == pytest.approx(want, rel=1e-12), (planted, got, want)
            assert (got > 1.0) == planted
