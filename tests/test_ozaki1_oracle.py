"""Pins of the Ozaki Scheme I oracle (oracle/ozaki1.py), the accuracy benchmark's
comparison baseline (reference benchmarks/ozaki1_gpu.hpp): digits reconstruct their
integer, products are exact when the scheme drops nothing, and the result stays within
the truncation bound of the dropped digit products."""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import ozaki1 as O1


@pytest.mark.parametrize("L", [2, 3, 5, 9])
def test_digits_reconstruct(L, rng):
    lim = 2 ** (7 * L - 2)
    xs = [0, 1, -1, 63, 64, -64, -65, lim - 1, -(lim - 1)] + [int(v) for v in rng.integers(-lim + 1, lim, 200)]
    for x in xs:
        d = O1.digits(x, L)
        assert all(-64 <= v <= 63 for v in d)
        assert sum(v * 128 ** (L - 1 - i) for i, v in enumerate(d)) == x


def test_rne_scaled_ties():
    assert O1.rne_scaled(2.5, 0) == 2 and O1.rne_scaled(3.5, 0) == 4 and O1.rne_scaled(-2.5, 0) == -2
    assert O1.rne_scaled(0.75, 1) == 2 and O1.rne_scaled(5e-324, 1074) == 1


def test_small_integers_exact(rng):
    # |a|, |b| < 64 fit one digit at the top of tk bits only if the row max leaves room; with
    # L = 9 (tk = 61) each integer entry is x = v 2^D exactly and keeps all of its bits in
    # the two leading digits, so every kept product is the whole product: the result is A B
    m, n, k = 5, 4, 7
    A = rng.integers(-50, 51, (m, k)).astype(np.float64)
    B = rng.integers(-50, 51, (k, n)).astype(np.float64)
    A[2] = 0.0
    C = O1.gemm(m, n, k, A.reshape(-1, order="F"), m, B.reshape(-1, order="F"), k, 9)
    assert np.array_equal(C, (A @ B).reshape(-1, order="F"))


@pytest.mark.parametrize("L", [2, 3, 4])
def test_truncation_bound(L, rng):
    # the dropped digit products (i + j >= L) bound the error: |C - AB| <= sum_p |a||b| 2^(-7(L-1)+1)
    # relative to the row / column scale, plus the final rounding
    m, n, k = 4, 3, 6
    A = rng.uniform(-1, 1, (m, k))
    B = rng.uniform(-1, 1, (k, n))
    C = O1.gemm(m, n, k, A.reshape(-1, order="F"), m, B.reshape(-1, order="F"), k, L).reshape(n, m).T
    tk = O1.precision_bits(L)
    for i in range(m):
        for j in range(n):
            exact = sum(Fraction(A[i, p]) * Fraction(B[p, j]) for p in range(k))
            ra = max(abs(A[i])) * 2.0 ** -tk  # size of one unit of x for row i (within 2x)
            rb = max(abs(B[:, j])) * 2.0 ** -tk
            # dropped part: digits i + j >= L weigh at most 64 * 64 * 128^(2(L-1) - L) units^2, L terms per pair
            drop = k * L * 64 * 64 * 128.0 ** (L - 2) * (2 * ra) * (2 * rb) * 128.0 ** 0
            rnd = (abs(A[i]).max() * 2.0 ** (1 - tk)) * abs(B[:, j]).sum() + abs(B[:, j]).max() * 2.0 ** (1 - tk) * abs(A[i]).sum()
            assert abs(Fraction(C[i, j]) - exact) <= Fraction(drop + rnd) + abs(exact) * Fraction(2) ** -52
