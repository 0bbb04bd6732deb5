"""Ozaki Scheme I (slicing), the comparison baseline of the accuracy benchmark: the
counterpart of the reference's oz1::Ozaki1 (benchmarks/ozaki1_gpu.hpp:86-155), run by
liboz2.so's oz2_ozaki1_dgemm on this library's kernels - the Scheme II scale kernels
for the exponents (tk = 7L - 2), a slicing kernel for the L balanced base-128 digits,
one int32-output INT8 GEMM per anti-diagonal of digit products and a double-double
Horner kernel (os2-xe_amd/csrc/ozaki1.hip).  No CPU fallback.
"""
import oz2

_handle = None


def precision_bits(L):
    return 7 * L - 2


def num_gemms(L):
    return L * (L + 1) // 2


def gemm(m, n, k, A, lda, B, ldb, C, ldc, L, oz=None):
    """C = A B (alpha = 1, beta = 0), column-major float64 device tensors, op = N; on the
    handle `oz` (default: one handle of this module on the default stream).  Raises
    oz2.OZ2InvalidArgument (a ValueError) for L outside [2, 9] or k 2^12 >= 2^31."""
    global _handle
    if oz is None:
        if _handle is None:
            _handle = oz2.Ozaki2()
        oz = _handle
    oz.ozaki1(m, n, k, A, lda, B, ldb, C, ldc, L)
