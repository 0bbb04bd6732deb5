/*
 * oz2_oracle.c - CPU oracle for the Ozaki Scheme II DGEMM/ZGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Never linked into, loaded by or
 * called from the product library.  Restates the reference header
 * src/ozaki2_gpu.hpp in plain C; each function cites the lines it follows.
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "oz2_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;
typedef __int128 i128;

/* ------------------------------------------------------------------------ */
/* Moduli sets for s = 2..24: the pairwise-coprime sets of the reference's   */
/* src/coprime_tables.hpp:10-111 (COPRIME_DATA at COPRIME_OFFSET[s]).        */
/* ------------------------------------------------------------------------ */
static const uint16_t kSets[] = {
    /* s=2  */ 256, 255,
    /* s=3  */ 256, 255, 253,
    /* s=4  */ 256, 255, 253, 251,
    /* s=5  */ 256, 255, 253, 251, 247,
    /* s=6  */ 256, 255, 253, 251, 247, 241,
    /* s=7  */ 256, 253, 251, 249, 247, 245, 241,
    /* s=8  */ 256, 253, 251, 249, 247, 245, 241, 239,
    /* s=9  */ 256, 253, 251, 249, 247, 245, 241, 239, 233,
    /* s=10 */ 256, 253, 251, 249, 247, 245, 241, 239, 233, 229,
    /* s=11 */ 256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227,
    /* s=12 */ 256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227, 223,
    /* s=13 */ 256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227, 223, 211,
    /* s=14 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211,
    /* s=15 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199,
    /* s=16 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197,
    /* s=17 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193,
    /* s=18 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191,
    /* s=19 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181,
    /* s=20 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
    /* s=21 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
               173,
    /* s=22 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
               173, 167,
    /* s=23 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
               173, 167, 163,
    /* s=24 */ 256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197, 193, 191, 181, 179,
               173, 167, 163, 157,
};

const uint16_t* oz2o_moduli(int s) {
    if (s < 2 || s > OZ2O_MAX_MODULI) return NULL;
    int off = 0;
    for (int q = 2; q < s; ++q) off += q;
    return kSets + off;
}

/* ------------------------------------------------------------------------ */
/* 256-bit helpers (little-endian u32 words)                                 */
/* ------------------------------------------------------------------------ */
static int big_mul_small(uint32_t* w, uint32_t f) {
    uint64_t c = 0;
    for (int i = 0; i < OZ2O_WORDS; ++i) {
        uint64_t t = (uint64_t)w[i] * f + c;
        w[i] = (uint32_t)t;
        c = t >> 32;
    }
    return c != 0;
}
static uint32_t big_divmod_small(uint32_t* w, uint32_t d) {
    uint64_t r = 0;
    for (int i = OZ2O_WORDS - 1; i >= 0; --i) {
        uint64_t cur = (r << 32) | w[i];
        w[i] = (uint32_t)(cur / d);
        r = cur % d;
    }
    return (uint32_t)r;
}
static int big_cmp(const uint32_t* a, const uint32_t* b) {
    for (int i = OZ2O_WORDS - 1; i >= 0; --i)
        if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
    return 0;
}
/* log2 estimate exactly as hpp:129-135 (top three words as a double) */
static double big_log2(const uint32_t* w) {
    int top = OZ2O_WORDS - 1;
    while (top > 0 && w[top] == 0) --top;
    double v = 0;
    int lo = top - 2 < 0 ? 0 : top - 2;
    for (int i = top; i >= lo; --i) v = v * 4294967296.0 + w[i];
    return log2(v) + 32.0 * (top - 2 > 0 ? top - 2 : 0);
}
/* k << sh compared with M; returns 1 if k*2^sh <= M (sh small enough not to exceed 320 bits) */
static int shifted_le(uint64_t k, int sh, const uint32_t* M) {
    uint32_t x[OZ2O_WORDS + 3];
    memset(x, 0, sizeof x);
    int wq = sh / 32, b = sh % 32;
    u128 v = (u128)k << b; /* at most 95 bits */
    for (int i = 0; i < 3; ++i) {
        if (wq + i < OZ2O_WORDS + 3) x[wq + i] = (uint32_t)(v >> (32 * i));
    }
    for (int i = OZ2O_WORDS + 2; i >= OZ2O_WORDS; --i)
        if (x[i]) return 0;
    return big_cmp(x, M) <= 0;
}
static int64_t mod_inverse(int64_t a, int64_t m) {
    int64_t t = 0, nt = 1, r = m, nr = ((a % m) + m) % m;
    while (nr) {
        int64_t q = r / nr, tmp;
        tmp = t - q * nt; t = nt; nt = tmp;
        tmp = r - q * nr; r = nr; nr = tmp;
    }
    if (r != 1) return -1;
    return (t % m + m) % m;
}

/* make_plan, hpp:159-200 */
int oz2o_make_plan(int s, int64_t k, oz2o_plan* p) {
    if (s < 2 || s > OZ2O_MAX_MODULI) return 1;
    if (k < 1) return 1;
    memset(p, 0, sizeof *p);
    p->s = s;
    const uint16_t* mods = oz2o_moduli(s);
    p->M[0] = 1;
    for (int t = 0; t < s; ++t) {
        if (mods[t] < 2 || mods[t] > 256) return 2;
        if (big_mul_small(p->M, mods[t])) return 2;
    }
    p->log2M = big_log2(p->M);
    int nl = (int)floor(p->log2M / 32.0) + 1;
    p->NL = nl < 2 ? 2 : nl;
    int tk = (int)floor((p->log2M - 2.0 - log2((double)k)) / 2.0);
    if (tk > 100) tk = 100;
    for (;; --tk) {
        if (tk < 1) return 1;
        if (shifted_le((uint64_t)k, 2 * tk + 2, p->M)) break;
    }
    p->tk = tk;
    for (int t = 0; t < s; ++t) {
        uint32_t m = mods[t];
        uint32_t Mt[OZ2O_WORDS];
        memcpy(Mt, p->M, sizeof Mt);
        if (big_divmod_small(Mt, m) != 0) return 2;
        uint32_t tmp[OZ2O_WORDS];
        memcpy(tmp, Mt, sizeof tmp);
        int64_t y = mod_inverse(big_divmod_small(tmp, m), m);
        if (y < 0) return 2;
        memcpy(p->W[t], Mt, sizeof Mt);
        big_mul_small(p->W[t], (uint32_t)y);
        p->m[t] = m;
        p->y[t] = (uint32_t)y;
    }
    return 0;
}

int oz2o_precision_bits(int64_t k, int s) {
    oz2o_plan p;
    return oz2o_make_plan(s, k, &p) == 0 ? p.tk : -1;
}

/* hpp:693-700: smallest s with tk >= bits, else 24 */
int oz2o_moduli_for_bits(int64_t k, int bits) {
    for (int s = 2; s <= OZ2O_MAX_MODULI; ++s) {
        int tk = oz2o_precision_bits(k, s);
        if (tk >= 0 && tk >= bits) return s;
    }
    return OZ2O_MAX_MODULI;
}

/* ------------------------------------------------------------------------ */
/* Scaling (hpp:309-377)                                                     */
/* ------------------------------------------------------------------------ */
static inline uint64_t absbits(double v) {
    uint64_t b;
    memcpy(&b, &v, 8);
    return b & 0x7FFFFFFFFFFFFFFFull;
}

int32_t oz2o_exponent_from_maxbits(uint64_t b, int tk) {
    if (b >= 0x7FF0000000000000ull) return INT_MIN;
    if (b == 0) return 0;
    double v;
    memcpy(&v, &b, 8);
    int e2;
    frexp(v, &e2);
    return tk - e2;
}

/* element (r, c) of op(X) for X column-major with leading dimension ldx */
static inline int64_t op_index(int trans, int64_t r, int64_t c, int64_t ldx) {
    return trans == 0 ? r + c * ldx : c + r * ldx;
}

void oz2o_exponents(int trans, int is_b, int64_t len, int64_t k, const double* X, int64_t ldx, int tk, int cplx,
                    int32_t* ex) {
    for (int64_t i = 0; i < len; ++i) {
        uint64_t mx = 0;
        for (int64_t p = 0; p < k; ++p) {
            /* row i of op(A): element (i,p); column i of op(B): element (p,i) */
            int64_t idx = is_b ? op_index(trans, p, i, ldx) : op_index(trans, i, p, ldx);
            if (cplx) {
                uint64_t a = absbits(X[2 * idx]), b = absbits(X[2 * idx + 1]);
                if (a > mx) mx = a;
                if (b > mx) mx = b;
            } else {
                uint64_t a = absbits(X[idx]);
                if (a > mx) mx = a;
            }
        }
        ex[i] = oz2o_exponent_from_maxbits(mx, tk);
    }
}

/* ------------------------------------------------------------------------ */
/* Exact split (hpp:251-303, 407-455)                                        */
/* ------------------------------------------------------------------------ */
void oz2o_scaled_int(double v, int e, int* neg, uint64_t* hi, uint64_t* lo) {
    uint64_t bits;
    memcpy(&bits, &v, 8);
    *neg = (int)(bits >> 63);
    int bexp = (int)((bits >> 52) & 0x7FF);
    uint64_t mant = bits & 0xFFFFFFFFFFFFFull;
    if (bexp) mant |= 1ull << 52;
    int sh = (bexp ? bexp : 1) - 1075 + e;
    u128 mag;
    if (mant == 0) {
        mag = 0;
    } else if (sh >= 0) {
        mag = sh < 75 ? ((u128)mant << sh) : 0; /* callers keep |v*2^e| < 2^101 */
    } else {
        int rs = -sh;
        if (rs >= 64) {
            mag = 0;
        } else {
            uint64_t q = mant >> rs;
            uint64_t rem = mant & ((1ull << rs) - 1);
            uint64_t half = 1ull << (rs - 1);
            mag = q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
        }
    }
    *hi = (uint64_t)(mag >> 64);
    *lo = (uint64_t)mag;
}

static inline int8_t sym_residue(int neg, uint64_t hi, uint64_t lo, uint32_t m) {
    u128 mag = ((u128)hi << 64) | lo;
    uint32_t r = (uint32_t)(mag % m);
    if (neg) r = (m - r) % m;
    uint32_t h = (m - 1) / 2;
    int v = (int)((r + h) % m) - (int)h;
    return (int8_t)(uint8_t)(v & 0xFF);
}

void oz2o_split_values(const double* v, const int32_t* e, int64_t n, int s, int8_t* planes) {
    const uint16_t* mods = oz2o_moduli(s);
    for (int64_t i = 0; i < n; ++i) {
        int neg = 0;
        uint64_t hi = 0, lo = 0;
        if (e[i] != INT_MIN) oz2o_scaled_int(v[i], e[i], &neg, &hi, &lo);
        for (int t = 0; t < s; ++t) planes[(int64_t)t * n + i] = sym_residue(neg, hi, lo, mods[t]);
    }
}

/* ------------------------------------------------------------------------ */
/* Exact conversion of a signed big integer to double: RNE to 53 bits, then  */
/* one ldexp (the reference's order, hpp:576-586).                           */
/* ------------------------------------------------------------------------ */
static double mag_to_double(const uint32_t* w, int nw, int neg, int shift) {
    int top = nw * 32 - 1;
    while (top >= 0 && !((w[top / 32] >> (top % 32)) & 1u)) --top;
    if (top < 0) return 0.0;
#define BIT(i) ((i) >= 0 ? ((w[(i) / 32] >> ((i) % 32)) & 1u) : 0u)
    uint64_t m53 = 0;
    for (int i = top; i > top - 53; --i) m53 = (m53 << 1) | BIT(i);
    int rb = top - 53;
    unsigned round = BIT(rb);
    int sticky = 0;
    for (int i = rb - 1; i >= 0; --i)
        if (BIT(i)) { sticky = 1; break; }
#undef BIT
    int e = top - 52;
    if (round && (sticky || (m53 & 1))) {
        ++m53;
        if (m53 == (1ull << 53)) { m53 >>= 1; ++e; }
    }
    /* m53 * 2^(e - shift); ldexp rounds once more only in the subnormal range */
    long long ex = (long long)e - shift;
    if (ex > 100000) ex = 100000;
    if (ex < -100000) ex = -100000;
    double r = ldexp((double)m53, (int)ex);
    return neg ? -r : r;
}

/* CRT of residues in [0, m_t): x = sum r_t W_t - q M with q = round(sum r_t y_t / m_t),
 * exact whenever |C'| <= M/4 (hpp:526-588); the result is RNE(x) * 2^-shift. */
double oz2o_crt_value(const oz2o_plan* p, const uint32_t* r, int shift) {
    /* 9 words so that sum r_t W_t (< 24 * 256 * M) never wraps */
    uint64_t acc[OZ2O_WORDS + 1];
    memset(acc, 0, sizeof acc);
    double f = 0.0;
    for (int t = 0; t < p->s; ++t) {
        f += (double)r[t] * (double)p->y[t] / (double)p->m[t];
        uint64_t c = 0;
        for (int i = 0; i < OZ2O_WORDS; ++i) {
            uint64_t v = acc[i] + (uint64_t)p->W[t][i] * r[t] + c;
            acc[i] = v & 0xFFFFFFFFull;
            c = v >> 32;
        }
        acc[OZ2O_WORDS] += c;
    }
    int64_t q = (int64_t)llround(f); /* f >= 0 */
    /* w = acc - q*M  (mod 2^288, two's complement) */
    uint32_t qm[OZ2O_WORDS + 1];
    uint64_t c = 0;
    for (int i = 0; i < OZ2O_WORDS; ++i) {
        uint64_t v = (uint64_t)p->M[i] * (uint64_t)q + c;
        qm[i] = (uint32_t)v;
        c = v >> 32;
    }
    qm[OZ2O_WORDS] = (uint32_t)c;
    int64_t br = 0;
    for (int i = 0; i < OZ2O_WORDS + 1; ++i) {
        int64_t v = (int64_t)acc[i] - (int64_t)qm[i] - br;
        br = v < 0;
        if (v < 0) v += 0x100000000ll;
        acc[i] = (uint64_t)v;
    }
    uint32_t w[OZ2O_WORDS + 1];
    for (int i = 0; i < OZ2O_WORDS + 1; ++i) w[i] = (uint32_t)acc[i];
    int neg = (w[OZ2O_WORDS] >> 31) & 1u;
    if (neg) {
        uint64_t cy = 1;
        for (int i = 0; i < OZ2O_WORDS + 1; ++i) {
            uint64_t v = (uint64_t)(uint32_t)~w[i] + cy;
            w[i] = (uint32_t)v;
            cy = v >> 32;
        }
    }
    return mag_to_double(w, OZ2O_WORDS + 1, neg, shift);
}

/* ------------------------------------------------------------------------ */
/* Residue-route DGEMM / ZGEMM                                               */
/* ------------------------------------------------------------------------ */
static int check_args(int ta, int tb, int64_t m, int64_t n, int64_t k, int64_t lda, int64_t ldb, int64_t ldc) {
    /* hpp:745-748 */
    if (m < 0 || n < 0 || k < 0) return 1;
    int64_t ra = ta == 0 ? m : k, rb = tb == 0 ? k : n;
    if (lda < (ra > 1 ? ra : 1) || ldb < (rb > 1 ? rb : 1) || ldc < (m > 1 ? m : 1)) return 1;
    return 0;
}

#define KCHUNK 131071

/* exact int8 products of plane rows: out[i + j*m] = sum over chunks of (chunk dot mod mt), in [0, mt) */
static void plane_gemm(const int8_t* Ap, const int8_t* Bp, int64_t m, int64_t n, int64_t k, uint32_t mt,
                       uint8_t* out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t j = 0; j < n; ++j) {
        const int8_t* b = Bp + j * k;
        for (int64_t i = 0; i < m; ++i) {
            const int8_t* a = Ap + i * k;
            int64_t res = 0;
            for (int64_t p0 = 0; p0 < k; p0 += KCHUNK) {
                int64_t p1 = p0 + KCHUNK < k ? p0 + KCHUNK : k;
                int32_t acc = 0;
                for (int64_t p = p0; p < p1; ++p) acc += (int32_t)a[p] * (int32_t)b[p];
                int64_t r = acc % (int64_t)mt;
                if (r < 0) r += mt;
                res += r;
            }
            out[i + j * m] = (uint8_t)(res % mt);
        }
    }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
}

/* split rows of op(X) (len rows, k columns), per-row exponents ex, into planes [t][row][k].
 * Real: symmetric residues of RNE(x*2^e) (hpp:407-455).  Complex (group g): Re, Im (negated
 * for conj) or Re+Im, each reduced to the symmetric range [-h, m-1-h] (hpp:457-520). */
static void split_op(int trans, int is_b, int64_t len, int64_t k, const double* X, int64_t ldx, const int32_t* ex,
                     int s, int cplx, int conj, int group, int8_t* planes) {
    const uint16_t* mods = oz2o_moduli(s);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < len; ++i) {
        for (int64_t p = 0; p < k; ++p) {
            int64_t idx = is_b ? op_index(trans, p, i, ldx) : op_index(trans, i, p, ldx);
            int n0 = 0, n1 = 0;
            uint64_t h0 = 0, l0 = 0, h1 = 0, l1 = 0;
            if (ex[i] != INT_MIN) {
                if (!cplx) {
                    oz2o_scaled_int(X[idx], ex[i], &n0, &h0, &l0);
                } else {
                    oz2o_scaled_int(X[2 * idx], ex[i], &n0, &h0, &l0);
                    oz2o_scaled_int(X[2 * idx + 1], ex[i], &n1, &h1, &l1);
                }
            }
            for (int t = 0; t < s; ++t) {
                const uint32_t m = mods[t];
                int8_t out;
                if (!cplx) {
                    out = sym_residue(n0, h0, l0, m);
                } else {
                    const int h = (int)(m - 1) / 2;
                    int s0 = sym_residue(n0, h0, l0, m), s1 = sym_residue(n1, h1, l1, m);
                    if (s0 < -h) s0 += (int)m; /* m = 256 stores +128 as -128 */
                    if (s1 < -h) s1 += (int)m;
                    if (conj) s1 = -s1;
                    int v = group == 0 ? s0 : group == 1 ? s1 : s0 + s1;
                    v = ((v % (int)m) + (int)m) % (int)m;
                    v = ((v + h) % (int)m) - h;
                    out = (int8_t)(uint8_t)(v & 0xFF);
                }
                planes[((int64_t)t * len + i) * k + p] = out;
            }
        }
    }
}

static void scale_c_real(int64_t m, int64_t n, double beta, double* C, int64_t ldc) {
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double* cp = C + i + j * ldc;
            *cp = beta == 0.0 ? 0.0 : beta * *cp;
        }
}

int oz2o_dgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s, int nthreads) {
    if (check_args(ta, tb, m, n, k, lda, ldb, ldc)) return 1;
    if (m == 0 || n == 0) return 0;
    if (k == 0 || alpha == 0.0) { scale_c_real(m, n, beta, C, ldc); return 0; }
    oz2o_plan p;
    int rc = oz2o_make_plan(s, k, &p);
    if (rc) return rc;
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1; /* C == T for real data (hpp:1001) */
    int32_t* D = (int32_t*)malloc(sizeof(int32_t) * m);
    int32_t* E = (int32_t*)malloc(sizeof(int32_t) * n);
    int8_t* Ap = (int8_t*)malloc((size_t)s * m * k);
    int8_t* Bp = (int8_t*)malloc((size_t)s * n * k);
    uint8_t* R = (uint8_t*)malloc((size_t)s * m * n);
    if (!D || !E || !Ap || !Bp || !R) { free(D); free(E); free(Ap); free(Bp); free(R); return 2; }
    oz2o_exponents(ta_, 0, m, k, A, lda, p.tk, 0, D);
    oz2o_exponents(tb_, 1, n, k, B, ldb, p.tk, 0, E);
    split_op(ta_, 0, m, k, A, lda, D, s, 0, 0, 0, Ap);
    split_op(tb_, 1, n, k, B, ldb, E, s, 0, 0, 0, Bp);
    for (int t = 0; t < s; ++t)
        plane_gemm(Ap + (size_t)t * m * k, Bp + (size_t)t * n * k, m, n, k, p.m[t], R + (size_t)t * m * n);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            uint32_t r[OZ2O_MAX_MODULI];
            for (int t = 0; t < s; ++t) r[t] = R[(size_t)t * m * n + i + j * m];
            double val;
            if (D[i] == INT_MIN || E[j] == INT_MIN) val = NAN;
            else val = oz2o_crt_value(&p, r, D[i] + E[j]);
            double* cp = C + i + j * ldc;
            /* epilogue (hpp:623) evaluated as fma(alpha, val, beta*c) */
            *cp = beta == 0.0 ? alpha * val : fma(alpha, val, beta * *cp);
        }
    }
    free(D); free(E); free(Ap); free(Bp); free(R);
    return 0;
}

static void scale_c_cplx(int64_t m, int64_t n, const double* beta, double* C, int64_t ldc) {
    double br = beta[0], bi = beta[1];
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double* cp = C + 2 * (i + j * ldc);
            if (br == 0.0 && bi == 0.0) { cp[0] = 0.0; cp[1] = 0.0; }
            else {
                double r = cp[0], im = cp[1];
                cp[0] = fma(br, r, -(bi * im));
                cp[1] = fma(br, im, bi * r);
            }
        }
}

int oz2o_zgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* alpha, const double* A, int64_t lda,
               const double* B, int64_t ldb, const double* beta, double* C, int64_t ldc, int s, int nthreads) {
    if (check_args(ta, tb, m, n, k, lda, ldb, ldc)) return 1;
    if (m == 0 || n == 0) return 0;
    if (k == 0 || (alpha[0] == 0.0 && alpha[1] == 0.0)) { scale_c_cplx(m, n, beta, C, ldc); return 0; }
    oz2o_plan p;
    int rc = oz2o_make_plan(s, 2 * k, &p); /* hpp:778-781 */
    if (rc) return rc;
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
    int32_t* D = (int32_t*)malloc(sizeof(int32_t) * m);
    int32_t* E = (int32_t*)malloc(sizeof(int32_t) * n);
    int8_t* Ap = (int8_t*)malloc((size_t)3 * s * m * k);
    int8_t* Bp = (int8_t*)malloc((size_t)3 * s * n * k);
    uint8_t* R = (uint8_t*)malloc((size_t)3 * s * m * n);
    if (!D || !E || !Ap || !Bp || !R) { free(D); free(E); free(Ap); free(Bp); free(R); return 2; }
    oz2o_exponents(ta_, 0, m, k, A, lda, p.tk, 1, D);
    oz2o_exponents(tb_, 1, n, k, B, ldb, p.tk, 1, E);
    for (int g = 0; g < 3; ++g) {
        split_op(ta_, 0, m, k, A, lda, D, s, 1, ta == 2, g, Ap + (size_t)g * s * m * k);
        split_op(tb_, 1, n, k, B, ldb, E, s, 1, tb == 2, g, Bp + (size_t)g * s * n * k);
    }
    for (int g = 0; g < 3; ++g)
        for (int t = 0; t < s; ++t)
            plane_gemm(Ap + ((size_t)g * s + t) * m * k, Bp + ((size_t)g * s + t) * n * k, m, n, k, p.m[t],
                       R + ((size_t)g * s + t) * m * n);
    const double ar = alpha[0], ai = alpha[1], br = beta[0], bi = beta[1];
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < m; ++i) {
            uint32_t rr[OZ2O_MAX_MODULI], ri[OZ2O_MAX_MODULI];
            for (int t = 0; t < s; ++t) {
                uint32_t mt = p.m[t];
                uint32_t p0 = R[((size_t)0 * s + t) * m * n + i + j * m];
                uint32_t p1 = R[((size_t)1 * s + t) * m * n + i + j * m];
                uint32_t p2 = R[((size_t)2 * s + t) * m * n + i + j * m];
                rr[t] = (p0 + 2 * mt - p1) % mt;          /* Re = P0 - P1      (hpp:648) */
                ri[t] = (p2 + 4 * mt - p0 - p1) % mt;     /* Im = P2 - P0 - P1 (hpp:649) */
            }
            double vr, vi;
            if (D[i] == INT_MIN || E[j] == INT_MIN) vr = vi = NAN;
            else {
                vr = oz2o_crt_value(&p, rr, D[i] + E[j]);
                vi = oz2o_crt_value(&p, ri, D[i] + E[j]);
            }
            double* cp = C + 2 * (i + j * ldc);
            double orr = fma(ar, vr, -(ai * vi)), oi = fma(ar, vi, ai * vr);
            if (br != 0.0 || bi != 0.0) {
                double cr = cp[0], ci = cp[1];
                orr = orr + fma(br, cr, -(bi * ci));
                oi = oi + fma(br, ci, bi * cr);
            }
            cp[0] = orr;
            cp[1] = oi;
        }
    }
    free(D); free(E); free(Ap); free(Bp); free(R);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Closed form: C_ij = RNE(sum_p x_ip y_pj) * 2^-(D_i+E_j) with exact         */
/* products (x = RNE(a*2^D), y = RNE(b*2^E)), summed in 256-bit integers.     */
/* Independent of moduli, residues and CRT.                                   */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t w[4]; } i256; /* two's complement */

static inline void i256_add_prod(i256* acc, int neg, u128 a, u128 b) {
    uint64_t a0 = (uint64_t)a, a1 = (uint64_t)(a >> 64), b0 = (uint64_t)b, b1 = (uint64_t)(b >> 64);
    /* unsigned 256-bit product (a, b < 2^128) */
    u128 p00 = (u128)a0 * b0, p01 = (u128)a0 * b1, p10 = (u128)a1 * b0, p11 = (u128)a1 * b1;
    uint64_t r[4];
    u128 t = (u128)(uint64_t)p00;
    r[0] = (uint64_t)t;
    t = (p00 >> 64) + (uint64_t)p01 + (uint64_t)p10;
    r[1] = (uint64_t)t;
    t = (t >> 64) + (p01 >> 64) + (p10 >> 64) + (uint64_t)p11;
    r[2] = (uint64_t)t;
    t = (t >> 64) + (p11 >> 64);
    r[3] = (uint64_t)t;
    if (neg) { /* r = -r */
        u128 c = 1;
        for (int i = 0; i < 4; ++i) {
            u128 v = (u128)(uint64_t)~r[i] + c;
            r[i] = (uint64_t)v;
            c = v >> 64;
        }
    }
    u128 c = 0;
    for (int i = 0; i < 4; ++i) {
        u128 v = (u128)acc->w[i] + r[i] + c;
        acc->w[i] = (uint64_t)v;
        c = v >> 64;
    }
}

static double i256_to_double(i256 x, int shift) {
    int neg = (int)(x.w[3] >> 63);
    if (neg) {
        u128 c = 1;
        for (int i = 0; i < 4; ++i) {
            u128 v = (u128)(uint64_t)~x.w[i] + c;
            x.w[i] = (uint64_t)v;
            c = v >> 64;
        }
    }
    uint32_t w[8];
    for (int i = 0; i < 4; ++i) { w[2 * i] = (uint32_t)x.w[i]; w[2 * i + 1] = (uint32_t)(x.w[i] >> 32); }
    return mag_to_double(w, 8, neg, shift);
}

typedef struct { int neg; u128 mag; } sint;

static sint scaled(double v, int e) {
    sint r;
    uint64_t hi, lo;
    oz2o_scaled_int(v, e, &r.neg, &hi, &lo);
    r.mag = ((u128)hi << 64) | lo;
    return r;
}

int oz2o_dgemm_closed(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int s, int nthreads) {
    if (check_args(ta, tb, m, n, k, lda, ldb, ldc)) return 1;
    if (m == 0 || n == 0) return 0;
    oz2o_plan p;
    int rc = oz2o_make_plan(s, k, &p);
    if (rc) return rc;
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
    int32_t* D = (int32_t*)malloc(sizeof(int32_t) * m);
    int32_t* E = (int32_t*)malloc(sizeof(int32_t) * n);
    sint* X = (sint*)malloc(sizeof(sint) * m * k);
    sint* Y = (sint*)malloc(sizeof(sint) * n * k);
    oz2o_exponents(ta_, 0, m, k, A, lda, p.tk, 0, D);
    oz2o_exponents(tb_, 1, n, k, B, ldb, p.tk, 0, E);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t q = 0; q < k; ++q) {
            sint z = {0, 0};
            if (D[i] != INT_MIN) z = scaled(A[op_index(ta_, i, q, lda)], D[i]);
            X[i * k + q] = z;
        }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t q = 0; q < k; ++q) {
            sint z = {0, 0};
            if (E[j] != INT_MIN) z = scaled(B[op_index(tb_, q, j, ldb)], E[j]);
            Y[j * k + q] = z;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            i256 acc = {{0, 0, 0, 0}};
            for (int64_t q = 0; q < k; ++q) {
                const sint x = X[i * k + q], y = Y[j * k + q];
                if (x.mag && y.mag) i256_add_prod(&acc, x.neg ^ y.neg, x.mag, y.mag);
            }
            double v = (D[i] == INT_MIN || E[j] == INT_MIN) ? NAN : i256_to_double(acc, D[i] + E[j]);
            C[i + j * ldc] = v;
        }
    free(D); free(E); free(X); free(Y);
    return 0;
}

int oz2o_zgemm_closed(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                      const double* B, int64_t ldb, double* C, int64_t ldc, int s, int nthreads) {
    if (check_args(ta, tb, m, n, k, lda, ldb, ldc)) return 1;
    if (m == 0 || n == 0) return 0;
    oz2o_plan p;
    int rc = oz2o_make_plan(s, 2 * k, &p);
    if (rc) return rc;
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
    int32_t* D = (int32_t*)malloc(sizeof(int32_t) * m);
    int32_t* E = (int32_t*)malloc(sizeof(int32_t) * n);
    sint* X = (sint*)malloc(sizeof(sint) * 2 * m * k);
    sint* Y = (sint*)malloc(sizeof(sint) * 2 * n * k);
    oz2o_exponents(ta_, 0, m, k, A, lda, p.tk, 1, D);
    oz2o_exponents(tb_, 1, n, k, B, ldb, p.tk, 1, E);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t q = 0; q < k; ++q) {
            sint zr = {0, 0}, zi = {0, 0};
            int64_t idx = op_index(ta_, i, q, lda);
            if (D[i] != INT_MIN) { zr = scaled(A[2 * idx], D[i]); zi = scaled(A[2 * idx + 1], D[i]); }
            if (ta == 2) zi.neg ^= 1;
            X[2 * (i * k + q)] = zr; X[2 * (i * k + q) + 1] = zi;
        }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t q = 0; q < k; ++q) {
            sint zr = {0, 0}, zi = {0, 0};
            int64_t idx = op_index(tb_, q, j, ldb);
            if (E[j] != INT_MIN) { zr = scaled(B[2 * idx], E[j]); zi = scaled(B[2 * idx + 1], E[j]); }
            if (tb == 2) zi.neg ^= 1;
            Y[2 * (j * k + q)] = zr; Y[2 * (j * k + q) + 1] = zi;
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            i256 re = {{0, 0, 0, 0}}, im = {{0, 0, 0, 0}};
            for (int64_t q = 0; q < k; ++q) {
                const sint xr = X[2 * (i * k + q)], xi = X[2 * (i * k + q) + 1];
                const sint yr = Y[2 * (j * k + q)], yi = Y[2 * (j * k + q) + 1];
                if (xr.mag && yr.mag) i256_add_prod(&re, xr.neg ^ yr.neg, xr.mag, yr.mag);
                if (xi.mag && yi.mag) i256_add_prod(&re, !(xi.neg ^ yi.neg), xi.mag, yi.mag);
                if (xr.mag && yi.mag) i256_add_prod(&im, xr.neg ^ yi.neg, xr.mag, yi.mag);
                if (xi.mag && yr.mag) i256_add_prod(&im, xi.neg ^ yr.neg, xi.mag, yr.mag);
            }
            double* cp = C + 2 * (i + j * ldc);
            if (D[i] == INT_MIN || E[j] == INT_MIN) { cp[0] = NAN; cp[1] = NAN; }
            else { cp[0] = i256_to_double(re, D[i] + E[j]); cp[1] = i256_to_double(im, D[i] + E[j]); }
        }
    free(D); free(E); free(X); free(Y);
    return 0;
}

static inline void dd_add_prod(double* sh, double* sl, double a, double b);

static int32_t exponent_of_row(int trans, int is_b, int64_t i, int64_t k, const double* X, int64_t ldx, int tk,
                               int cplx) {
    int32_t e;
    oz2o_exponents(trans, is_b, 1, k, is_b ? X + (trans == 0 ? i * ldx : i) * (cplx ? 2 : 1)
                                           : X + (trans == 0 ? i : i * ldx) * (cplx ? 2 : 1),
                   ldx, tk, cplx, &e);
    return e;
}

int oz2o_dgemm_entries(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                       int64_t ldb, int s, int64_t count, const int64_t* rows, const int64_t* cols, double* out) {
    oz2o_plan p;
    int rc = oz2o_make_plan(s, k, &p);
    if (rc) return rc;
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t e = 0; e < count; ++e) {
        const int64_t i = rows[e], j = cols[e];
        const int32_t Di = exponent_of_row(ta_, 0, i, k, A, lda, p.tk, 0);
        const int32_t Ej = exponent_of_row(tb_, 1, j, k, B, ldb, p.tk, 0);
        if (Di == INT_MIN || Ej == INT_MIN) { out[e] = NAN; continue; }
        i256 acc = {{0, 0, 0, 0}};
        for (int64_t q = 0; q < k; ++q) {
            const sint x = scaled(A[op_index(ta_, i, q, lda)], Di), y = scaled(B[op_index(tb_, q, j, ldb)], Ej);
            if (x.mag && y.mag) i256_add_prod(&acc, x.neg ^ y.neg, x.mag, y.mag);
        }
        out[e] = i256_to_double(acc, Di + Ej);
    }
    (void)m; (void)n;
    return 0;
}

int oz2o_zgemm_entries(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                       int64_t ldb, int s, int64_t count, const int64_t* rows, const int64_t* cols, double* out) {
    oz2o_plan p;
    int rc = oz2o_make_plan(s, 2 * k, &p);
    if (rc) return rc;
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t e = 0; e < count; ++e) {
        const int64_t i = rows[e], j = cols[e];
        const int32_t Di = exponent_of_row(ta_, 0, i, k, A, lda, p.tk, 1);
        const int32_t Ej = exponent_of_row(tb_, 1, j, k, B, ldb, p.tk, 1);
        if (Di == INT_MIN || Ej == INT_MIN) { out[2 * e] = NAN; out[2 * e + 1] = NAN; continue; }
        i256 re = {{0, 0, 0, 0}}, im = {{0, 0, 0, 0}};
        for (int64_t q = 0; q < k; ++q) {
            const int64_t ia = op_index(ta_, i, q, lda), ib = op_index(tb_, q, j, ldb);
            sint xr = scaled(A[2 * ia], Di), xi = scaled(A[2 * ia + 1], Di);
            sint yr = scaled(B[2 * ib], Ej), yi = scaled(B[2 * ib + 1], Ej);
            if (ta == 2) xi.neg ^= 1;
            if (tb == 2) yi.neg ^= 1;
            if (xr.mag && yr.mag) i256_add_prod(&re, xr.neg ^ yr.neg, xr.mag, yr.mag);
            if (xi.mag && yi.mag) i256_add_prod(&re, !(xi.neg ^ yi.neg), xi.mag, yi.mag);
            if (xr.mag && yi.mag) i256_add_prod(&im, xr.neg ^ yi.neg, xr.mag, yi.mag);
            if (xi.mag && yr.mag) i256_add_prod(&im, xi.neg ^ yr.neg, xi.mag, yr.mag);
        }
        out[2 * e] = i256_to_double(re, Di + Ej);
        out[2 * e + 1] = i256_to_double(im, Di + Ej);
    }
    (void)m; (void)n;
    return 0;
}

void oz2o_dd_entries(int ta, int tb, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                     int64_t count, const int64_t* rows, const int64_t* cols, double* hi, double* lo, int cplx) {
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t e = 0; e < count; ++e) {
        const int64_t i = rows[e], j = cols[e];
        if (!cplx) {
            double sh = 0, sl = 0;
            for (int64_t q = 0; q < k; ++q) dd_add_prod(&sh, &sl, A[op_index(ta_, i, q, lda)], B[op_index(tb_, q, j, ldb)]);
            hi[e] = sh;
            lo[e] = sl;
        } else {
            double h[2] = {0, 0}, l[2] = {0, 0};
            for (int64_t q = 0; q < k; ++q) {
                int64_t ia = op_index(ta_, i, q, lda), ib = op_index(tb_, q, j, ldb);
                double ar = A[2 * ia], ai = ta == 2 ? -A[2 * ia + 1] : A[2 * ia + 1];
                double br = B[2 * ib], bi = tb == 2 ? -B[2 * ib + 1] : B[2 * ib + 1];
                dd_add_prod(&h[0], &l[0], ar, br);
                dd_add_prod(&h[0], &l[0], -ai, bi);
                dd_add_prod(&h[1], &l[1], ar, bi);
                dd_add_prod(&h[1], &l[1], ai, br);
            }
            hi[2 * e] = h[0]; hi[2 * e + 1] = h[1];
            lo[2 * e] = l[0]; lo[2 * e + 1] = l[1];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Double-double references (test:56-84, 356-377)                            */
/* ------------------------------------------------------------------------ */
static inline void dd_add_prod(double* sh, double* sl, double a, double b) {
    const double ph = a * b;
    const double pl = fma(a, b, -ph);
    const double s = *sh + ph;
    const double bb = s - *sh;
    const double e = (*sh - (s - bb)) + (ph - bb);
    const double t = e + *sl + pl;
    *sh = s + t;
    *sl = t - (*sh - s);
}

void oz2o_dd_gemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                  int64_t ldb, double* hi, double* lo, int nthreads) {
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double sh = 0, sl = 0;
            for (int64_t p = 0; p < k; ++p)
                dd_add_prod(&sh, &sl, A[op_index(ta_, i, p, lda)], B[op_index(tb_, p, j, ldb)]);
            hi[i + j * m] = sh;
            lo[i + j * m] = sl;
        }
}

void oz2o_dd_zgemm(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                   int64_t ldb, double* hi, double* lo, int nthreads) {
    set_threads(nthreads);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double h[2] = {0, 0}, l[2] = {0, 0};
            for (int64_t p = 0; p < k; ++p) {
                int64_t ia = op_index(ta_, i, p, lda), ib = op_index(tb_, p, j, ldb);
                double ar = A[2 * ia], ai = ta == 2 ? -A[2 * ia + 1] : A[2 * ia + 1];
                double br = B[2 * ib], bi = tb == 2 ? -B[2 * ib + 1] : B[2 * ib + 1];
                dd_add_prod(&h[0], &l[0], ar, br);
                dd_add_prod(&h[0], &l[0], -ai, bi);
                dd_add_prod(&h[1], &l[1], ar, bi);
                dd_add_prod(&h[1], &l[1], ai, br);
            }
            hi[2 * (i + j * m)] = h[0]; hi[2 * (i + j * m) + 1] = h[1];
            lo[2 * (i + j * m)] = l[0]; lo[2 * (i + j * m) + 1] = l[1];
        }
}

/* ------------------------------------------------------------------------ */
/* Rigorous element-wise bound (test:128-178) and complex bound (384-433)     */
/* ------------------------------------------------------------------------ */
static int frexp_exp(double v) { int e; frexp(v, &e); return e; }

double oz2o_bound_ratio(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double alpha, double beta, const double* C0, const double* C,
                        int64_t ldc, const double* hi, const double* lo, int s) {
    int tk = oz2o_precision_bits(k, s);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
    int* D = (int*)malloc(sizeof(int) * m);
    int* E = (int*)malloc(sizeof(int) * n);
    double* ra = (double*)calloc(m, sizeof(double));
    double* cb = (double*)calloc(n, sizeof(double));
    for (int64_t i = 0; i < m; ++i) {
        double mx = 0;
        for (int64_t p = 0; p < k; ++p) {
            double v = fabs(A[op_index(ta_, i, p, lda)]);
            if (v > mx) mx = v;
            ra[i] += v;
        }
        D[i] = mx == 0 ? 0 : tk - frexp_exp(mx);
    }
    for (int64_t j = 0; j < n; ++j) {
        double mx = 0;
        for (int64_t p = 0; p < k; ++p) {
            double v = fabs(B[op_index(tb_, p, j, ldb)]);
            if (v > mx) mx = v;
            cb[j] += v;
        }
        E[j] = mx == 0 ? 0 : tk - frexp_exp(mx);
    }
    const double u = ldexp(1.0, -53);
    double worst = 0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            double got = C[i + j * ldc], c0 = C0[i + j * ldc];
            double rh = hi[i + j * m], rl = lo[i + j * m];
            double tb_err = 0.5 * ldexp(cb[j], -D[i]) + 0.5 * ldexp(ra[i], -E[j]) +
                            0.25 * (double)k * ldexp(1.0, -D[i] - E[j]);
            double err, bound;
            if (alpha == 1.0 && beta == 0.0) {
                err = fabs((got - rh) - rl);
                bound = tb_err * (1 + 2 * u) + u * fabs(got) + 1e-300;
            } else {
                double ref = alpha * (rh + rl) + beta * c0;
                err = fabs(got - ref);
                bound = fabs(alpha) * tb_err * (1 + 4 * u) + 4 * u * (fabs(alpha * rh) + fabs(beta * c0)) + 1e-300;
            }
            double r = err / bound;
            if (!(r <= worst)) worst = r; /* NaN propagates as failure */
        }
    free(D); free(E); free(ra); free(cb);
    return worst;
}

double oz2o_zbound_ratio(int ta, int tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                         const double* B, int64_t ldb, const double* alpha, const double* beta, const double* C0,
                         const double* C, int64_t ldc, const double* hi, const double* lo, int s) {
    int tk = oz2o_precision_bits(2 * k, s);
    int ta_ = ta == 0 ? 0 : 1, tb_ = tb == 0 ? 0 : 1;
    int* D = (int*)malloc(sizeof(int) * m);
    int* E = (int*)malloc(sizeof(int) * n);
    double* sa = (double*)calloc(m, sizeof(double));
    double* sb = (double*)calloc(n, sizeof(double));
    for (int64_t i = 0; i < m; ++i) {
        double mx = 0;
        for (int64_t p = 0; p < k; ++p) {
            int64_t idx = op_index(ta_, i, p, lda);
            double vr = fabs(A[2 * idx]), vi = fabs(A[2 * idx + 1]);
            if (vr > mx) mx = vr;
            if (vi > mx) mx = vi;
            sa[i] += vr + vi;
        }
        D[i] = mx == 0 ? 0 : tk - frexp_exp(mx);
    }
    for (int64_t j = 0; j < n; ++j) {
        double mx = 0;
        for (int64_t p = 0; p < k; ++p) {
            int64_t idx = op_index(tb_, p, j, ldb);
            double vr = fabs(B[2 * idx]), vi = fabs(B[2 * idx + 1]);
            if (vr > mx) mx = vr;
            if (vi > mx) mx = vi;
            sb[j] += vr + vi;
        }
        E[j] = mx == 0 ? 0 : tk - frexp_exp(mx);
    }
    const double u = ldexp(1.0, -53);
    const int plain = alpha[0] == 1.0 && alpha[1] == 0.0 && beta[0] == 0.0 && beta[1] == 0.0;
    const double aabs = hypot(alpha[0], alpha[1]);
    double worst = 0;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            const double* h = hi + 2 * (i + j * m);
            const double* l = lo + 2 * (i + j * m);
            const double tb_err = 0.5 * ldexp(sb[j], -D[i]) + 0.5 * ldexp(sa[i], -E[j]) +
                                  0.5 * (double)k * ldexp(1.0, -D[i] - E[j]);
            const double* g = C + 2 * (i + j * ldc);
            const double* c0 = C0 + 2 * (i + j * ldc);
            /* exact complex reference alpha*(hi+lo) + beta*c0 */
            double rr = h[0] + l[0], ri = h[1] + l[1];
            double exr = alpha[0] * rr - alpha[1] * ri + (beta[0] * c0[0] - beta[1] * c0[1]);
            double exi = alpha[0] * ri + alpha[1] * rr + (beta[0] * c0[1] + beta[1] * c0[0]);
            double babs = hypot(beta[0] * c0[0] - beta[1] * c0[1], beta[0] * c0[1] + beta[1] * c0[0]);
            for (int c = 0; c < 2; ++c) {
                double err, bound;
                if (plain) {
                    err = fabs((g[c] - h[c]) - l[c]);
                    bound = tb_err * (1 + 4 * u) + u * fabs(g[c]) + 1e-300;
                } else {
                    err = fabs(g[c] - (c ? exi : exr));
                    bound = aabs * tb_err * 2 * (1 + 8 * u) + 8 * u * (aabs * (fabs(h[0]) + fabs(h[1])) + babs) +
                            1e-300;
                }
                double r = err / bound;
                if (!(r <= worst)) worst = r;
            }
        }
    free(D); free(E); free(sa); free(sb);
    return worst;
}

/* ------------------------------------------------------------------------ */
/* Accuracy metrics (bench_accuracy.cpp:71-94)                               */
/* ------------------------------------------------------------------------ */
void oz2o_metrics(int64_t rows, int64_t cols, const double* C, int64_t ldc, const double* hi, const double* lo,
                  double* ulp_max, double* ulp_mean, double* cr_frac) {
    double mx = 0, sum = 0;
    int64_t cr = 0;
    for (int64_t j = 0; j < cols; ++j)
        for (int64_t i = 0; i < rows; ++i) {
            double c = C[i + j * ldc], h = hi[i + j * rows], l = lo[i + j * rows];
            double err = fabs((c - h) - l);
            double ulp = h == 0 ? 4.9406564584124654e-324 : nextafter(fabs(h), INFINITY) - fabs(h);
            double eu = err / ulp;
            if (eu > mx) mx = eu;
            sum += eu;
            if (c == h + l) ++cr;
        }
    int64_t cnt = rows * cols;
    *ulp_max = mx;
    *ulp_mean = cnt ? sum / cnt : 0;
    *cr_frac = cnt ? (double)cr / cnt : 1;
}
