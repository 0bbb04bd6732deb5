/*
 * oz2_oracle.h - CPU oracle for the Ozaki Scheme II DGEMM/ZGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke check and the CPU-baseline leg of bench.py call it; the product
 * library never links, loads or calls it.
 *
 * It restates, in plain C, the algorithm of the reference header
 * src/ozaki2_gpu.hpp:
 *   plan (moduli, tk, y_t, W_t)        hpp:55-200
 *   exponents, exact scaled split      hpp:251-303, 360-377, 407-520
 *   exact INT8 GEMM per modulus        hpp:1023-1042 (oneMKL gemm_batch, int8 -> int32)
 *   CRT + alpha/beta epilogue          hpp:526-664
 *   double-double reference, bound     tests/test_ozaki2.cpp:56-84, 156-173, 356-426
 *   accuracy metrics                   benchmarks/bench_accuracy.cpp:71-94
 * and adds a second, independent route (exact products summed in 256-bit
 * integers, "closed form") that must agree bit-for-bit with the residue route.
 *
 * Pinning: every function here is checked by tests/test_oracle.py against
 * Python big-integer golden vectors (tests/golden), brute force, and the
 * closed form; see the header of that test file.
 *
 * Matrices are column-major. Trans codes: 0 = N, 1 = T, 2 = C.
 * Return codes: 0 ok, 1 invalid argument, 2 runtime error.
 */
#ifndef OZ2_ORACLE_H
#define OZ2_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OZ2O_MAX_MODULI 24
#define OZ2O_WORDS 8 /* 256-bit little-endian u32 words */

typedef struct {
    int s, NL, tk;
    double log2M;
    uint32_t m[OZ2O_MAX_MODULI];
    uint32_t y[OZ2O_MAX_MODULI];
    uint32_t W[OZ2O_MAX_MODULI][OZ2O_WORDS];
    uint32_t M[OZ2O_WORDS];
} oz2o_plan;

/* moduli set of size s (2..24); returns pointer to s entries or NULL */
const uint16_t* oz2o_moduli(int s);
int oz2o_make_plan(int s, int64_t k, oz2o_plan* p);
int oz2o_precision_bits(int64_t k, int s); /* -1 if make_plan rejects */
int oz2o_moduli_for_bits(int64_t k, int bits);

/* exponent of one row/column from the max |x| bit pattern (hpp:360-377) */
int32_t oz2o_exponent_from_maxbits(uint64_t maxbits, int tk);
/* exponents of rows of op(A) (is_b=0, len m) or columns of op(B) (is_b=1, len n) */
void oz2o_exponents(int trans, int is_b, int64_t rows_or_cols, int64_t k, const double* X, int64_t ldx, int tk,
                    int cplx, int32_t* ex);

/* exact RNE(v * 2^e) as sign + 128-bit magnitude (hpp:251-289); e != INT_MIN */
void oz2o_scaled_int(double v, int e, int* neg, uint64_t* hi, uint64_t* lo);
/* symmetric int8 residues of RNE(v_i*2^e_i) for n values: planes[t*n + i] */
void oz2o_split_values(const double* v, const int32_t* e, int64_t n, int s, int8_t* planes);

/* exact CRT of s residues r_t in [0, m_t) -> double RNE(C')*2^-shift (hpp:526-588) */
double oz2o_crt_value(const oz2o_plan* p, const uint32_t* r, int shift);

/* full residue-route Ozaki-II DGEMM / ZGEMM (interleaved complex) */
int oz2o_dgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s, int nthreads);
int oz2o_zgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* alpha, const double* A, int64_t lda,
               const double* B, int64_t ldb, const double* beta, double* C, int64_t ldc, int s, int nthreads);

/* independent closed form for alpha=1, beta=0 (C written with ldc) */
int oz2o_dgemm_closed(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int s, int nthreads);
int oz2o_zgemm_closed(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int s, int nthreads);

/* closed form at selected entries (rows[e], cols[e]) only: O((count + #rows + #cols) k) */
int oz2o_dgemm_entries(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                       int64_t ldb, int s, int64_t count, const int64_t* rows, const int64_t* cols, double* out);
int oz2o_zgemm_entries(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                       int64_t ldb, int s, int64_t count, const int64_t* rows, const int64_t* cols, double* out);
/* double-double products at selected entries (hi, lo) */
void oz2o_dd_entries(int ta, int tb, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                     int64_t count, const int64_t* rows, const int64_t* cols, double* hi, double* lo, int cplx);

/* double-double references; hi/lo are m x n dense (ld m); complex: re/im interleaved pairs */
void oz2o_dd_gemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                  int64_t ldb, double* hi, double* lo, int nthreads);
void oz2o_dd_zgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                   int64_t ldb, double* hi, double* lo, int nthreads);

/* rigorous element-wise bound ratio max(err/bound) (test:156-178), real */
double oz2o_bound_ratio(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double alpha, double beta, const double* C0, const double* C,
                        int64_t ldc, const double* hi, const double* lo, int s);
/* complex bound ratio (test:384-433) */
double oz2o_zbound_ratio(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                         const double* B, int64_t ldb, const double* alpha, const double* beta, const double* C0,
                         const double* C, int64_t ldc, const double* hi, const double* lo, int s);

/* accuracy metrics (bench_accuracy.cpp:71-94) over count values with stride-1 hi/lo and C(ld) */
void oz2o_metrics(int64_t rows, int64_t cols, const double* C, int64_t ldc, const double* hi, const double* lo,
                  double* ulp_max, double* ulp_mean, double* cr_frac);

#ifdef __cplusplus
}
#endif
#endif
