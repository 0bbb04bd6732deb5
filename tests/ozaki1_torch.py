"""Ozaki Scheme I restated on torch operations and hipBLASLt INT8 GEMMs (torch._int_mm):
test infrastructure, the independent cross-check of the library's Ozaki Scheme I path
(oz2_ozaki1_dgemm, os2-xe_amd/csrc/ozaki1.hip) at sizes the exact Python oracle
(oracle/ozaki1.py) cannot reach.  It follows the reference's baseline
(benchmarks/ozaki1_gpu.hpp:10-155):

For L slices (2 <= L <= 9) each row of A and column of B is scaled to tk = 7L - 2 bits
with the Scheme II exponent rule (D = tk - frexp_exp(max |x|); 0 for a zero row or
column, non-finite rows give NaN), x = RNE(v 2^D) is cut into L balanced base-128
digits in [-64, 63] (launch_slices), the L(L+1)/2 digit products with i + j <= L - 1
are int8 GEMMs with int32 sums (k 2^12 < 2^31), and a double-double Horner
T = 128 T + S_d over the anti-diagonals d = 0 .. L-1 (S_d summed in int64) gives
C = RN(T) 2^(7(L-1) - D_i - E_j) (launch_horner).
"""
import torch

INT32_MIN = -2 ** 31


def precision_bits(L):
    return 7 * L - 2


def num_gemms(L):
    return L * (L + 1) // 2


def _exponents(M, dim, tk):
    """per row (dim=1) or column (dim=0) of M: tk - frexp_exp(max |x|), 0 if all zero, INT32_MIN if non-finite"""
    mx = M.abs().amax(dim=dim)
    _, e = torch.frexp(mx)
    D = (tk - e.to(torch.int64))
    D = torch.where(mx == 0, torch.zeros_like(D), D)
    return torch.where(torch.isfinite(mx), D, torch.full_like(D, INT32_MIN))


def _pow2(e):
    """exact 2^e as float64 for an int64 tensor e in [-1022, 1023], from the exponent bits"""
    return ((e + 1023) << 52).view(torch.float64)


def _ldexp(v, e):
    """v 2^e with one rounding, for float64 v and int64 e (torch.ldexp goes through pow,
    which is not exact for every exponent on the GPU).  e = e1 + e2 with e2 in
    [-1022, 1023]: v 2^e1 is exact while it stays normal (|e1| <= 178 here: |v| < 2^107
    and beyond |e| > 1200 the result is 0 or infinite either way), then v 2^e2 rounds once."""
    e = e.clamp(-1200, 1200)
    e2 = e.clamp(-1022, 1023)
    return v * _pow2(e - e2) * _pow2(e2)


def _scaled_ints(M, D):
    """RNE(M 2^D) as int64 (D broadcast against M; rows with D = INT32_MIN give 0)"""
    bad = D == INT32_MIN
    Dz = torch.where(bad, torch.zeros_like(D), D)
    h = torch.div(Dz, 2, rounding_mode="floor")
    # two exact power-of-two steps: each factor and each product stays in the double range
    v = torch.where(bad, torch.zeros_like(M), M) * _pow2(h) * _pow2(Dz - h)
    return torch.round(v).to(torch.int64)


def _slices(x, L):
    """L int8 digits of x in [-64, 63], most significant first: x = sum_i s_i 128^(L-1-i)"""
    out = [None] * L
    for i in range(L - 1, -1, -1):
        d = ((x + 64) & 127) - 64
        x = (x - d) >> 7
        out[i] = d.to(torch.int8)
    return out


def _int_mm(a, b):
    """int8 (m x k) @ int8 (k x n) -> int32, padded to the library's size rules"""
    m, k = a.shape
    n = b.shape[1]
    mp, kp, np_ = max(32, -(-m // 8) * 8), max(16, -(-k // 8) * 8), max(16, -(-n // 8) * 8)
    if (mp, kp, np_) != (m, k, n):
        a2 = torch.zeros((mp, kp), dtype=torch.int8, device=a.device)
        b2 = torch.zeros((kp, np_), dtype=torch.int8, device=b.device)
        a2[:m, :k] = a
        b2[:k, :n] = b
        return torch._int_mm(a2, b2)[:m, :n]
    return torch._int_mm(a.contiguous(), b.contiguous())


def gemm(m, n, k, A, lda, B, ldb, C, ldc, L):
    """C = A B (alpha = 1, beta = 0), column-major float64 device tensors, op = N"""
    if L < 2 or L > 9:
        raise ValueError("oz1: L must be in [2, 9]")
    if k * 4096 >= 2 ** 31:
        raise ValueError("oz1: k too large (int32 GEMM)")
    if m == 0 or n == 0:
        return
    tk = precision_bits(L)
    Am = A[: lda * k].view(k, lda)[:, :m].t()  # m x k
    Bm = B[: ldb * n].view(n, ldb)[:, :k].t()  # k x n
    D = _exponents(Am, 1, tk)
    E = _exponents(Bm, 0, tk)
    sa = _slices(_scaled_ints(Am, D[:, None]), L)
    sb = _slices(_scaled_ints(Bm, E[None, :]), L)
    hi = torch.zeros((m, n), dtype=torch.float64, device=A.device)
    lo = torch.zeros_like(hi)
    for d in range(L):
        S = torch.zeros((m, n), dtype=torch.int64, device=A.device)
        for i in range(d + 1):
            S += _int_mm(sa[i], sb[d - i]).to(torch.int64)
        sd = S.to(torch.float64)
        h2, l2 = hi * 128.0, lo * 128.0
        s = h2 + sd
        bb = s - h2
        err = (h2 - (s - bb)) + (sd - bb) + l2
        hi = s + err
        lo = err - (hi - s)
    shift = 7 * (L - 1) - D[:, None] - E[None, :]
    bad = (D[:, None] == INT32_MIN) | (E[None, :] == INT32_MIN)
    out = _ldexp(hi + lo, shift)
    out = torch.where(bad, torch.full_like(out, float("nan")), out)
    C[: ldc * n].view(n, ldc)[:, :m] = out.t()
