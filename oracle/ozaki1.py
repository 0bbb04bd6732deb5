"""Ozaki Scheme I oracle: test infrastructure only.  Exact CPU restatement, with Python
integers, of the reference's comparison baseline (benchmarks/ozaki1_gpu.hpp:10-155)
for small matrices:

  tk = 7L - 2; D_i = tk - frexp_exp(max_p |A_ip|) (0 for a zero row), likewise E_j over
  the columns of B (ozaki2_gpu.hpp:360-377); x = RNE(v 2^D) exactly (hpp:251-289);
  x = sum_i s_i 128^(L-1-i) with balanced digits s_i in [-64, 63] (launch_slices:
  d = ((x + 64) & 127) - 64, x = (x - d) >> 7, least significant digit first);
  T = sum_{d=0}^{L-1} 128^(L-1-d) sum_{i+j=d} A_i B_j (the digit products kept by the
  scheme); C = RN(T) 2^(7(L-1) - D_i - E_j) (launch_horner: the double-double Horner is
  exact for |T| < 2^106, then one rounding and ldexp).

Pinned by tests/test_ozaki1_oracle.py: digit reconstruction, exact integer products
when no digit product is dropped, and the truncation bound.
"""
import math
from fractions import Fraction

import numpy as np


def precision_bits(L):
    return 7 * L - 2


def num_gemms(L):
    return L * (L + 1) // 2


def exponent(mx, tk):
    """tk - frexp_exp(mx) for a finite max magnitude, 0 for zero"""
    if mx == 0:
        return 0
    return tk - math.frexp(mx)[1]


def rne_scaled(v, e):
    """RNE(v 2^e) as an int, exactly"""
    q = Fraction(v) * (Fraction(2) ** e)
    f = math.floor(q)
    r = q - f
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and f % 2 == 1):
        f += 1
    return int(f)


def digits(x, L):
    """balanced base-128 digits, most significant first"""
    out = [0] * L
    for i in range(L - 1, -1, -1):
        d = ((x + 64) & 127) - 64
        x = (x - d) >> 7
        out[i] = d
    return out


def gemm(m, n, k, A, lda, B, ldb, L):
    """column-major m x n result (float64 numpy) of Ozaki Scheme I with L slices"""
    tk = precision_bits(L)
    Am = [[float(A[i + p * lda]) for p in range(k)] for i in range(m)]
    Bm = [[float(B[p + j * ldb]) for j in range(n)] for p in range(k)]
    D = [exponent(max(abs(v) for v in row), tk) for row in Am]
    E = [exponent(max(abs(Bm[p][j]) for p in range(k)), tk) for j in range(n)]
    da = [[digits(rne_scaled(Am[i][p], D[i]), L) for p in range(k)] for i in range(m)]
    db = [[digits(rne_scaled(Bm[p][j], E[j]), L) for j in range(n)] for p in range(k)]
    C = np.zeros(m * n)
    for i in range(m):
        for j in range(n):
            T = 0
            for d in range(L):
                S = 0
                for a in range(d + 1):
                    S += sum(da[i][p][a] * db[p][j][d - a] for p in range(k))
                T = T * 128 + S
            C[i + j * m] = round_to_double(Fraction(T) * Fraction(2) ** (7 * (L - 1) - D[i] - E[j]))
    return C


def round_to_double(q):
    """RN(q) as a double; +-inf when the rounded value overflows (Python raises there)"""
    try:
        return float(q)
    except OverflowError:
        return math.inf if q > 0 else -math.inf
