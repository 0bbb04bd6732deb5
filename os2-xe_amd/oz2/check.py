"""Accuracy measurement against a double-double product (the reference's
bench_accuracy.cpp:71-94 metrics): a double-double GEMM on the GPU
(liboz2_check.so, oz2chk_dd_gemm) gives hi + lo; the error of an entry C is
|(C - hi) - lo| in units of ulp(hi) = nextafter(|hi|, inf) - |hi|.  Used by the
benchmark and to label the selector's training data."""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_PATH = os.path.join(os.path.dirname(_HERE), "lib", "liboz2_check.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_PATH):
            raise RuntimeError(f"oz2: {_PATH} is not built (make -C os2-xe_amd)")
        L = ctypes.CDLL(_PATH)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.oz2chk_dd_gemm.argtypes = [vp, i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, vp, i32]
        L.oz2chk_gen.argtypes = [ctypes.c_char_p, i64, i64, ctypes.c_uint64, i32, vp]
        _lib = L
    return _lib


_T = {"N": 0, "T": 1, "C": 1}


def dd_gemm(ta, tb, m, n, k, A, lda, B, ldb):
    """(hi, lo) float64 device tensors of length m*n (column-major) for op(A) op(B)"""
    hi = torch.empty(m * n, dtype=torch.float64, device="cuda")
    lo = torch.empty_like(hi)
    rc = lib().oz2chk_dd_gemm(None, _T[ta], _T[tb], m, n, k, A.data_ptr(), lda, B.data_ptr(), ldb, hi.data_ptr(),
                              lo.data_ptr(), 0)
    if rc != 0:
        raise RuntimeError(f"oz2: dd reference failed ({rc})")
    return hi, lo


def ulp_errors(C, hi, lo, m, ldc):
    """ulp error of every entry of the m x n result stored with leading dimension ldc"""
    n = hi.numel() // max(1, m)
    Cv = C.reshape(-1)[: ldc * n].reshape(n, ldc)[:, :m].reshape(-1)
    err = ((Cv - hi) - lo).abs()
    a = hi.abs()
    ulp = torch.where(hi == 0, torch.full_like(a, 5e-324), torch.nextafter(a, torch.full_like(a, float("inf"))) - a)
    return err / ulp


def metrics(C, hi, lo, m, ldc):
    """ulp_max, ulp_mean, cr_frac of C against hi + lo"""
    e = ulp_errors(C, hi, lo, m, ldc)
    n = hi.numel() // max(1, m)
    Cv = C.reshape(-1)[: ldc * n].reshape(n, ldc)[:, :m].reshape(-1)
    return {"ulp_max": float(e.max()), "ulp_mean": float(e.mean()), "cr_frac": float((Cv == hi + lo).double().mean())}


def bound_ratio(m, n, k, A, lda, B, ldb, C, ldc, hi, lo, tk):
    """Largest |C - (hi + lo)| / bound over the entries of C = A B (op N/N, alpha = 1, beta = 0,
    column-major device tensors) against the reference's rigorous element-wise bound at scaling
    width tk (tests/test_ozaki2.cpp:162-173): with D_i = tk - e(max_p |A_ip|), E_j = tk -
    e(max_p |B_pj|) (0 for a zero row or column; e the frexp exponent), r_i = sum_p |A_ip| and
    c_j = sum_p |B_pj|, bound = (c_j 2^-D_i / 2 + r_i 2^-E_j / 2 + k 2^-(D_i+E_j) / 4)(1 + 2u)
    + u |C_ij|, u = 2^-53.  At most 1 when every entry is within it."""
    Am = A.reshape(-1)[: lda * k].reshape(k, lda)[:, :m].abs()  # [p][i]
    Bm = B.reshape(-1)[: ldb * n].reshape(n, ldb)[:, :k].abs()  # [j][p]
    ra, cb = Am.sum(0), Bm.sum(1)

    def expo(mx):
        e = torch.frexp(mx).exponent.to(torch.float64)
        return torch.where(mx == 0, torch.zeros_like(e), tk - e)

    D, E = expo(Am.amax(0)), expo(Bm.amax(1))
    u = 2.0 ** -53
    got = C.reshape(-1)[: ldc * n].reshape(n, ldc)[:, :m]  # [j][i]
    h, l = hi.reshape(n, m), lo.reshape(n, m)
    tb = (0.5 * (cb[:, None] * torch.exp2(-D)[None, :]) + 0.5 * (ra[None, :] * torch.exp2(-E)[:, None]) +
          0.25 * k * torch.exp2(-(D[None, :] + E[:, None])))
    err = ((got - h) - l).abs()
    bound = tb * (1 + 2 * u) + u * got.abs() + 1e-300
    return float((err / bound).max())


def zbound_ratio(ta, tb, m, n, k, A, lda, B, ldb, C, ldc, hi, lo, tk):
    """The complex form of bound_ratio for C = op(A) op(B) with alpha = 1, beta = 0 (interleaved
    complex128 device tensors, ops N, T or C; hi, lo interleaved m x n), tk the width of the 2k
    plan: per real component, bound = (s_j 2^-D_i / 2 + r_i 2^-E_j / 2 + k 2^-(D_i+E_j) / 2)(1 + 4u)
    + u |C_ij component|, where r_i, s_j sum |Re| + |Im| over the row of op(A) / column of op(B)
    and D, E come from the largest |Re| or |Im| of each (tests/test_ozaki2.cpp:162-173 for the 3M
    form)."""
    def parts(X, ld, rows, op):  # [row of op(X)][p] of |Re| and |Im|
        X = X.view(torch.float64) if X.is_complex() else X
        if op == "N":  # column-major rows x k: element (r, p) at r + p ld
            Xv = X.view(-1, 2)[: ld * k].reshape(k, ld, 2)[:, :rows, :].transpose(0, 1)
        else:          # stored k x rows: element (r, p) of op(X) at p + r ld
            Xv = X.view(-1, 2)[: ld * rows].reshape(rows, ld, 2)[:, :k, :]
        return Xv.abs()

    Ap = parts(A, lda, m, ta)                      # [i][p][2]
    Bp = parts(B, ldb, n, "N" if tb == "T" or tb == "C" else "T")  # [j][p][2]: column j of op(B)
    sa, sb = Ap.sum((1, 2)), Bp.sum((1, 2))

    def expo(mx):
        e = torch.frexp(mx).exponent.to(torch.float64)
        return torch.where(mx == 0, torch.zeros_like(e), tk - e)

    D, E = expo(Ap.amax((1, 2))), expo(Bp.amax((1, 2)))
    u = 2.0 ** -53
    tb_ = (0.5 * (sb[:, None] * torch.exp2(-D)[None, :]) + 0.5 * (sa[None, :] * torch.exp2(-E)[:, None]) +
           0.5 * k * torch.exp2(-(D[None, :] + E[:, None])))  # [j][i]
    got = C.view(-1, 2)[: ldc * n].reshape(n, ldc, 2)[:, :m, :]
    h, l = hi.view(n, m, 2), lo.view(n, m, 2)
    err = ((got - h) - l).abs()
    bound = tb_[:, :, None] * (1 + 4 * u) + u * got.abs() + 1e-300
    return float((err / bound).max())


def gen(dist, rows, cols, seed, is_a):
    """rows x cols column-major float64 (numpy) input of the accuracy benchmark: the reference's
    generator (bench_accuracy.cpp:41-69) for dist in urand, phiX, sincos, spike, cancel"""
    import numpy as np
    X = np.empty(rows * cols, dtype=np.float64)
    rc = lib().oz2chk_gen(dist.encode(), rows, cols, seed, 1 if is_a else 0, X.ctypes.data)
    if rc != 0:
        raise ValueError(f"oz2: unknown distribution {dist!r}")
    return X


def full_metrics(C, hi, lo, absab):
    """the reference's accuracy metrics (bench_accuracy.cpp:71-94) of C (flat, same layout as
    hi, lo, absab): ulp_max, ulp_mean, cr_frac, ab_max = max err / sum |a||b|, fro_rel =
    ||err||_F / ||hi||_F"""
    err = ((C - hi) - lo).abs()
    a = hi.abs()
    ulp = torch.where(hi == 0, torch.full_like(a, 5e-324), torch.nextafter(a, torch.full_like(a, float("inf"))) - a)
    eu = err / ulp
    pos = absab > 0
    ab = (err[pos] / absab[pos]).max() if bool(pos.any()) else torch.zeros((), dtype=torch.float64)
    return {"ulp_max": float(eu.max()), "ulp_mean": float(eu.mean()), "cr_frac": float((C == hi + lo).double().mean()),
            "ab_max": float(ab), "fro_rel": float(torch.sqrt((err * err).sum() / (hi * hi).sum()))}
