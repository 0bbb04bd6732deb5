"""Shared drivers for the GPU parity tests: run one DGEMM/ZGEMM case through
the C ABI (liboz2.so) and through the oracle on the same inputs, then check
  * bit-identical C against the oracle (the reference's algorithm restated),
  * the rigorous element-wise bound against a double-double product
    (tests/test_ozaki2.cpp:156-178, 412-433),
  * ldc padding rows untouched (test:179-182).
"""
import sys

import numpy as np
import torch

import oz2
from conftest import make_cmat, make_mat
from oracle import oracle as O


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def flat(X):
    return np.asfortranarray(X).reshape(-1, order="F")


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def real_case(oz, ta, tb, m, n, k, s, dist, alpha=1.0, beta=0.0, lda_extra=0, ldc_extra=0, phi=1.0, seed=1,
              check_oracle=True, check_bound=True):
    rng = np.random.default_rng(seed)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + lda_extra, br + lda_extra, m + ldc_extra
    A = flat(make_mat(ar, ac, lda, dist, rng, phi))
    B = flat(make_mat(br, bc, ldb, dist, rng, phi))
    C0 = rng.uniform(-1, 1, ldc * n)
    log(f"case dgemm {ta}{tb} m={m} n={n} k={k} s={s} {dist} alpha={alpha} beta={beta}")
    dC = to_dev(C0.copy())
    oz.gemm(ta, tb, m, n, k, alpha, to_dev(A), lda, to_dev(B), ldb, beta, dC, ldc, s)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    res = {"C": C, "A": A, "B": B, "C0": C0, "lda": lda, "ldb": ldb, "ldc": ldc}
    pad_ok = all(np.array_equal(C[j * ldc + m:(j + 1) * ldc], C0[j * ldc + m:(j + 1) * ldc]) for j in range(n))
    assert pad_ok, "ldc padding modified"
    if check_oracle:
        ref = C0.copy()
        O.dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, ref, ldc, s)
        nbad = int(np.count_nonzero(bits(C) != bits(ref)))
        res["mismatch"] = nbad
        assert nbad == 0, f"{nbad} entries differ from the oracle"
    if check_bound:
        hi, lo = O.dd_gemm(ta, tb, m, n, k, A, lda, B, ldb)
        r = O.bound_ratio(ta, tb, m, n, k, A, lda, B, ldb, alpha, beta, C0, C, ldc, hi, lo, s)
        res["bound"] = r
        res["hi"], res["lo"] = hi, lo
        assert r <= 1.0, f"err/bound = {r}"
    return res


def complex_case(oz, ta, tb, m, n, k, s, dist, alpha=1 + 0j, beta=0j, ld_extra=0, phi=1.0, seed=3,
                 check_oracle=True, check_bound=True, c_shift=0):
    """c_shift: C starts that many doubles into its device buffer (1 = only 8-byte aligned)"""
    rng = np.random.default_rng(seed)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ar + ld_extra, br + ld_extra, m + ld_extra
    A = make_cmat(ar, ac, lda, dist, rng, phi).reshape(-1, order="F").view(np.float64).copy()
    B = make_cmat(br, bc, ldb, dist, rng, phi).reshape(-1, order="F").view(np.float64).copy()
    C0 = rng.uniform(-1, 1, 2 * ldc * n)
    log(f"case zgemm {ta}{tb} m={m} n={n} k={k} s={s} {dist} alpha={alpha} beta={beta}")
    dbuf = to_dev(np.concatenate([np.full(c_shift, 7.0), C0]))
    dC = dbuf[c_shift:]
    oz.zgemm(ta, tb, m, n, k, alpha, to_dev(A), lda, to_dev(B), ldb, beta, dC, ldc, s)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    assert np.all(dbuf[:c_shift].cpu().numpy() == 7.0)
    res = {"C": C}
    for j in range(n):
        assert np.array_equal(C[2 * (j * ldc + m):2 * (j + 1) * ldc], C0[2 * (j * ldc + m):2 * (j + 1) * ldc])
    if check_oracle:
        ref = C0.copy()
        O.zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, ref, ldc, s)
        nbad = int(np.count_nonzero(bits(C) != bits(ref)))
        assert nbad == 0, f"{nbad} values differ from the oracle"
    if check_bound:
        hi, lo = O.dd_zgemm(ta, tb, m, n, k, A, lda, B, ldb)
        r = O.zbound_ratio(ta, tb, m, n, k, A, lda, B, ldb, alpha, beta, C0, C, ldc, hi, lo, s)
        res["bound"] = r
        assert r <= 1.0, f"err/bound = {r}"
    return res


def fma(a, b, c):
    """correctly rounded a*b + c in double precision (Python's Fraction -> float rounds to nearest even)"""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))
