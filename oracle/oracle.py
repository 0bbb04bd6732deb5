"""ctypes binding of the CPU oracle (oracle/oz2_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product library never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboz2_oracle.so")
_lib = None

TRANS = {"N": 0, "T": 1, "C": 2}
MAX_MODULI = 24


class Plan(ctypes.Structure):
    _fields_ = [
        ("s", ctypes.c_int), ("NL", ctypes.c_int), ("tk", ctypes.c_int), ("log2M", ctypes.c_double),
        ("m", ctypes.c_uint32 * MAX_MODULI), ("y", ctypes.c_uint32 * MAX_MODULI),
        ("W", (ctypes.c_uint32 * 8) * MAX_MODULI), ("M", ctypes.c_uint32 * 8),
    ]

    def big(self, words):
        return sum(int(w) << (32 * i) for i, w in enumerate(words))

    @property
    def M_int(self):
        return self.big(self.M)

    def W_int(self, t):
        return self.big(self.W[t])


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER
        d, i64, i32, vp = ctypes.c_double, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p
        L.oz2o_moduli.restype = P(ctypes.c_uint16)
        L.oz2o_moduli.argtypes = [i32]
        L.oz2o_make_plan.argtypes = [i32, i64, P(Plan)]
        L.oz2o_precision_bits.argtypes = [i64, i32]
        L.oz2o_moduli_for_bits.argtypes = [i64, i32]
        L.oz2o_exponent_from_maxbits.argtypes = [ctypes.c_uint64, i32]
        L.oz2o_exponent_from_maxbits.restype = ctypes.c_int32
        L.oz2o_exponents.argtypes = [i32, i32, i64, i64, vp, i64, i32, i32, vp]
        L.oz2o_scaled_int.argtypes = [d, i32, P(i32), P(ctypes.c_uint64), P(ctypes.c_uint64)]
        L.oz2o_split_values.argtypes = [vp, vp, i64, i32, vp]
        L.oz2o_crt_value.argtypes = [P(Plan), vp, i32]
        L.oz2o_crt_value.restype = d
        L.oz2o_dgemm.argtypes = [i32, i32, i64, i64, i64, d, vp, i64, vp, i64, d, vp, i64, i32, i32]
        L.oz2o_zgemm.argtypes = [i32, i32, i64, i64, i64, vp, vp, i64, vp, i64, vp, vp, i64, i32, i32]
        L.oz2o_dgemm_closed.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, i64, i32, i32]
        L.oz2o_zgemm_closed.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, i64, i32, i32]
        L.oz2o_dgemm_entries.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, i32, i64, vp, vp, vp]
        L.oz2o_zgemm_entries.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, i32, i64, vp, vp, vp]
        L.oz2o_dd_entries.argtypes = [i32, i32, i64, vp, i64, vp, i64, i64, vp, vp, vp, vp, i32]
        L.oz2o_dd_gemm.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, vp, i32]
        L.oz2o_dd_zgemm.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, vp, i32]
        L.oz2o_bound_ratio.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, d, d, vp, vp, i64, vp, vp, i32]
        L.oz2o_bound_ratio.restype = d
        L.oz2o_zbound_ratio.argtypes = [i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, vp, vp, vp, i64, vp, vp, i32]
        L.oz2o_zbound_ratio.restype = d
        L.oz2o_metrics.argtypes = [i64, i64, vp, i64, vp, vp, P(d), P(d), P(d)]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def moduli(s):
    p = lib().oz2o_moduli(s)
    return [int(p[i]) for i in range(s)]


def make_plan(s, k):
    p = Plan()
    rc = lib().oz2o_make_plan(s, k, ctypes.byref(p))
    if rc:
        raise ValueError(f"make_plan(s={s}, k={k}) rejected (rc={rc})")
    return p


def precision_bits(k, s):
    return lib().oz2o_precision_bits(k, s)


def moduli_for_bits(k, bits):
    return lib().oz2o_moduli_for_bits(k, bits)


def exponent_from_maxbits(b, tk):
    return lib().oz2o_exponent_from_maxbits(b, tk)


def exponents(trans, is_b, length, k, X, ldx, tk, cplx=False):
    ex = np.zeros(length, dtype=np.int32)
    lib().oz2o_exponents(TRANS[trans], int(is_b), length, k, _p(X), ldx, tk, int(cplx), _p(ex))
    return ex


def scaled_int(v, e):
    neg = ctypes.c_int()
    hi = ctypes.c_uint64()
    lo = ctypes.c_uint64()
    lib().oz2o_scaled_int(float(v), int(e), ctypes.byref(neg), ctypes.byref(hi), ctypes.byref(lo))
    mag = (hi.value << 64) | lo.value
    return -mag if neg.value else mag


def split_values(v, e, s):
    v = np.ascontiguousarray(v, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.int32)
    out = np.zeros((s, v.size), dtype=np.int8)
    lib().oz2o_split_values(_p(v), _p(e), v.size, s, _p(out))
    return out


def crt_value(plan, residues, shift):
    r = np.ascontiguousarray(residues, dtype=np.uint32)
    return lib().oz2o_crt_value(ctypes.byref(plan), _p(r), int(shift))


def _f64(a):
    return np.asfortranarray(a, dtype=np.float64)


def dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, nthreads=0):
    """Residue-route Ozaki-II DGEMM on flat column-major buffers; C is updated in place."""
    rc = lib().oz2o_dgemm(TRANS[ta], TRANS[tb], m, n, k, alpha, _p(A), lda, _p(B), ldb, beta, _p(C), ldc, s,
                          nthreads)
    if rc:
        raise ValueError(f"oracle dgemm rejected arguments (rc={rc})")


def zgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s, nthreads=0):
    al = np.array([alpha.real, alpha.imag], dtype=np.float64)
    be = np.array([beta.real, beta.imag], dtype=np.float64)
    rc = lib().oz2o_zgemm(TRANS[ta], TRANS[tb], m, n, k, _p(al), _p(A), lda, _p(B), ldb, _p(be), _p(C), ldc, s,
                          nthreads)
    if rc:
        raise ValueError(f"oracle zgemm rejected arguments (rc={rc})")


def dgemm_closed(ta, tb, m, n, k, A, lda, B, ldb, C, ldc, s, nthreads=0):
    rc = lib().oz2o_dgemm_closed(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, _p(C), ldc, s, nthreads)
    if rc:
        raise ValueError(f"oracle closed form rejected arguments (rc={rc})")


def zgemm_closed(ta, tb, m, n, k, A, lda, B, ldb, C, ldc, s, nthreads=0):
    rc = lib().oz2o_zgemm_closed(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, _p(C), ldc, s, nthreads)
    if rc:
        raise ValueError(f"oracle closed form rejected arguments (rc={rc})")


def gemm_entries(ta, tb, m, n, k, A, lda, B, ldb, s, rows, cols, cplx=False):
    """Closed-form values (alpha=1, beta=0) at entries (rows[e], cols[e]); complex gives (re, im) pairs."""
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    out = np.zeros(rows.size * (2 if cplx else 1))
    f = lib().oz2o_zgemm_entries if cplx else lib().oz2o_dgemm_entries
    rc = f(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, s, rows.size, _p(rows), _p(cols), _p(out))
    if rc:
        raise ValueError(f"oracle entries rejected arguments (rc={rc})")
    return out


def dd_entries(ta, tb, k, A, lda, B, ldb, rows, cols, cplx=False):
    rows = np.ascontiguousarray(rows, dtype=np.int64)
    cols = np.ascontiguousarray(cols, dtype=np.int64)
    hi = np.zeros(rows.size * (2 if cplx else 1))
    lo = np.zeros_like(hi)
    lib().oz2o_dd_entries(TRANS[ta], TRANS[tb], k, _p(A), lda, _p(B), ldb, rows.size, _p(rows), _p(cols), _p(hi),
                          _p(lo), int(cplx))
    return hi, lo


def dd_gemm(ta, tb, m, n, k, A, lda, B, ldb, nthreads=0):
    hi = np.zeros(m * n)
    lo = np.zeros(m * n)
    lib().oz2o_dd_gemm(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, _p(hi), _p(lo), nthreads)
    return hi, lo


def dd_zgemm(ta, tb, m, n, k, A, lda, B, ldb, nthreads=0):
    hi = np.zeros(2 * m * n)
    lo = np.zeros(2 * m * n)
    lib().oz2o_dd_zgemm(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, _p(hi), _p(lo), nthreads)
    return hi, lo


def bound_ratio(ta, tb, m, n, k, A, lda, B, ldb, alpha, beta, C0, C, ldc, hi, lo, s):
    return lib().oz2o_bound_ratio(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, alpha, beta, _p(C0), _p(C),
                                  ldc, _p(hi), _p(lo), s)


def zbound_ratio(ta, tb, m, n, k, A, lda, B, ldb, alpha, beta, C0, C, ldc, hi, lo, s):
    al = np.array([alpha.real, alpha.imag], dtype=np.float64)
    be = np.array([beta.real, beta.imag], dtype=np.float64)
    return lib().oz2o_zbound_ratio(TRANS[ta], TRANS[tb], m, n, k, _p(A), lda, _p(B), ldb, _p(al), _p(be), _p(C0),
                                   _p(C), ldc, _p(hi), _p(lo), s)


def metrics(rows, cols, C, ldc, hi, lo):
    a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    lib().oz2o_metrics(rows, cols, _p(C), ldc, _p(hi), _p(lo), ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
    return {"ulp_max": a.value, "ulp_mean": b.value, "cr_frac": c.value}
