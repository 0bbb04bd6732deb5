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
