// ddref.hip - double-double reference GEMM on the GPU (liboz2_check.so).
//
// Measurement harness only, not part of the product library: bench.py and the
// tests use it to compute ulp errors and correctly-rounded fractions of
// full-size results.  Each entry accumulates sum_p a_ip b_pj in double-double
// with the reference's update order (tests/test_ozaki2.cpp:62-76): TwoProd via
// fma, then TwoSum, sequentially over p -- the same operations in the same
// order as the CPU oracle, so both give identical (hi, lo).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int TB = 16;

__device__ __forceinline__ void dd_add(double& sh, double& sl, double a, double b) {
    const double ph = a * b;
    const double pl = __builtin_fma(a, b, -ph);
    const double s = sh + ph;
    const double bb = s - sh;
    const double e = (sh - (s - bb)) + (ph - bb);
    const double t = e + sl + pl;
    sh = s + t;
    sl = t - (sh - s);
}

// real: op(X)(r, c) of a column-major X
__device__ __forceinline__ double opv(const double* X, int64_t ld, int trans, int64_t r, int64_t c) {
    return trans == 0 ? X[r + c * ld] : X[c + r * ld];
}

__global__ void __launch_bounds__(256) dd_real_kernel(int ta, int tb, int64_t m, int64_t n, int64_t k,
                                                      const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ hi, double* __restrict__ lo) {
    __shared__ double sA[TB][TB + 1], sB[TB][TB + 1];
    const int tx = threadIdx.x % TB, ty = threadIdx.x / TB;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * TB, j0 = static_cast<int64_t>(blockIdx.y) * TB;
    double sh = 0, sl = 0;
    for (int64_t p0 = 0; p0 < k; p0 += TB) {
        {
            const int64_t i = i0 + tx, p = p0 + ty;
            sA[ty][tx] = (i < m && p < k) ? opv(A, lda, ta, i, p) : 0.0;
            const int64_t j = j0 + tx, q = p0 + ty;
            sB[ty][tx] = (j < n && q < k) ? opv(B, ldb, tb, q, j) : 0.0;
        }
        __syncthreads();
        const int kk_end = static_cast<int>(k - p0 < TB ? k - p0 : TB);
        for (int kk = 0; kk < kk_end; ++kk) dd_add(sh, sl, sA[kk][tx], sB[kk][ty]);
        __syncthreads();
    }
    const int64_t i = i0 + tx, j = j0 + ty;
    if (i < m && j < n) {
        hi[i + j * m] = sh;
        lo[i + j * m] = sl;
    }
}

// complex (interleaved); conj for trans == 2; output hi/lo interleaved (re, im)
__global__ void __launch_bounds__(256) dd_cplx_kernel(int ta, int tb, int64_t m, int64_t n, int64_t k,
                                                      const double* __restrict__ A, int64_t lda,
                                                      const double* __restrict__ B, int64_t ldb,
                                                      double* __restrict__ hi, double* __restrict__ lo) {
    __shared__ double2 sA[TB][TB + 1], sB[TB][TB + 1];
    const int tx = threadIdx.x % TB, ty = threadIdx.x / TB;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * TB, j0 = static_cast<int64_t>(blockIdx.y) * TB;
    double h0 = 0, l0 = 0, h1 = 0, l1 = 0;
    for (int64_t p0 = 0; p0 < k; p0 += TB) {
        {
            const int64_t i = i0 + tx, p = p0 + ty;
            double2 va = make_double2(0.0, 0.0);
            if (i < m && p < k) {
                const int64_t idx = ta == 0 ? i + p * lda : p + i * lda;
                va = make_double2(A[2 * idx], ta == 2 ? -A[2 * idx + 1] : A[2 * idx + 1]);
            }
            sA[ty][tx] = va;
            const int64_t j = j0 + tx, q = p0 + ty;
            double2 vb = make_double2(0.0, 0.0);
            if (j < n && q < k) {
                const int64_t idx = tb == 0 ? q + j * ldb : j + q * ldb;
                vb = make_double2(B[2 * idx], tb == 2 ? -B[2 * idx + 1] : B[2 * idx + 1]);
            }
            sB[ty][tx] = vb;
        }
        __syncthreads();
        const int kk_end = static_cast<int>(k - p0 < TB ? k - p0 : TB);
        for (int kk = 0; kk < kk_end; ++kk) {
            const double2 a = sA[kk][tx], b = sB[kk][ty];
            dd_add(h0, l0, a.x, b.x);
            dd_add(h0, l0, -a.y, b.y);
            dd_add(h1, l1, a.x, b.y);
            dd_add(h1, l1, a.y, b.x);
        }
        __syncthreads();
    }
    const int64_t i = i0 + tx, j = j0 + ty;
    if (i < m && j < n) {
        hi[2 * (i + j * m)] = h0;
        hi[2 * (i + j * m) + 1] = h1;
        lo[2 * (i + j * m)] = l0;
        lo[2 * (i + j * m) + 1] = l1;
    }
}

}  // namespace

extern "C" int oz2chk_dd_gemm(void* stream, int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A,
                              int64_t lda, const double* B, int64_t ldb, double* hi, double* lo, int cplx) {
    if (m <= 0 || n <= 0) return 0;
    const dim3 grid(static_cast<unsigned>((m + TB - 1) / TB), static_cast<unsigned>((n + TB - 1) / TB));
    if (cplx)
        hipLaunchKernelGGL(dd_cplx_kernel, grid, dim3(TB * TB), 0, static_cast<hipStream_t>(stream), ta, tb, m, n, k,
                           A, lda, B, ldb, hi, lo);
    else
        hipLaunchKernelGGL(dd_real_kernel, grid, dim3(TB * TB), 0, static_cast<hipStream_t>(stream), ta, tb, m, n, k,
                           A, lda, B, ldb, hi, lo);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
