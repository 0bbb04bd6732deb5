// scale_split.hip - scaling (row/column max -> power-of-two exponents) and the
// exact split of op(A) / op(B) into s int8 residue planes, for gfx950.
//
// Reference behaviour: src/ozaki2_gpu.hpp:309-377 (max |x| per row / column,
// exponent D = tk - frexp_exp(max), 0 for an all-zero row, kNonFinite for
// Inf/NaN) and hpp:251-303, 407-520 (x = RNE(v 2^e) taken exactly from the
// IEEE bits, symmetric residues per modulus; complex: Re, +-Im, Re+-Im).
//
// Both operands are written in one blocked plane layout (plane_offset: 1 KB
// blocks of 16 rows x 64 k), so the GEMM copies A and B tiles the same way.
// A thread's 16 (real) or 4 (complex) consecutive k stay inside one 16-byte
// chunk of a block, so each plane store is one aligned vector store.  Rows of op(X)
// that are strided in memory (op(A) = N, op(B) = T) are read column-by-column
// with lanes along rows, so no separate transpose pass is needed.
#include <algorithm>

#include "kernels.hpp"
#include "math.hpp"

namespace oz2 {
namespace k {
using namespace math;

namespace {

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
    const uint32_t lo = __shfl_xor(static_cast<uint32_t>(v), mask, 64);
    const uint32_t hi = __shfl_xor(static_cast<uint32_t>(v >> 32), mask, 64);
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ __forceinline__ uint64_t absbits(const double* p) {
    return __double_as_longlong(*p) & 0x7FFFFFFFFFFFFFFFll;
}

// One wave per row of op(X), row contiguous in memory: n doubles from base.
__global__ void __launch_bounds__(256) scale_kcontig_kernel(const double* __restrict__ X, int64_t ldx, int64_t len,
                                                            int64_t ndbl, int cplx, int tk, int32_t* __restrict__ ex) {
    const int lane = threadIdx.x & 63;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
    if (r >= len) return;
    const double* row = X + (cplx ? 2 : 1) * r * ldx;
    uint64_t mx = 0;
    if ((reinterpret_cast<uintptr_t>(row) & 15) == 0) {
        const int64_t n2 = ndbl >> 1;
        const double2* row2 = reinterpret_cast<const double2*>(row);
        for (int64_t i = lane; i < n2; i += 64) {
            const double2 v = row2[i];
            const uint64_t a = __double_as_longlong(v.x) & 0x7FFFFFFFFFFFFFFFll;
            const uint64_t b = __double_as_longlong(v.y) & 0x7FFFFFFFFFFFFFFFll;
            mx = max(mx, max(a, b));
        }
        if ((ndbl & 1) && lane == 0) mx = max(mx, absbits(row + ndbl - 1));
    } else {
        for (int64_t i = lane; i < ndbl; i += 64) mx = max(mx, absbits(row + i));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, shfl_xor_u64(mx, o));
    if (lane == 0) ex[r] = exponent_from_maxbits(mx, tk);
}

// Rows strided in memory: block = 64 rows x 4 k-lanes over a k range; partial max per (k block, row).
__global__ void __launch_bounds__(256) scale_kstrided_kernel(const double* __restrict__ X, int64_t ldx, int64_t len,
                                                             int64_t k, int64_t kper, int cplx,
                                                             uint64_t* __restrict__ partial) {
    __shared__ uint64_t red[4][64];
    const int rl = threadIdx.x & 63, kl = threadIdx.x >> 6;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * 64 + rl;
    const int64_t p0 = static_cast<int64_t>(blockIdx.y) * kper;
    const int64_t p1 = min(k, p0 + kper);
    uint64_t mx = 0;
    if (r < len) {
        if (cplx) {
            for (int64_t p = p0 + kl; p < p1; p += 4) {
                const double2 v = *reinterpret_cast<const double2*>(X + 2 * (r + p * ldx));
                mx = max(mx, max(static_cast<uint64_t>(__double_as_longlong(v.x) & 0x7FFFFFFFFFFFFFFFll),
                                 static_cast<uint64_t>(__double_as_longlong(v.y) & 0x7FFFFFFFFFFFFFFFll)));
            }
        } else {
            for (int64_t p = p0 + kl; p < p1; p += 4) mx = max(mx, absbits(X + r + p * ldx));
        }
    }
    red[kl][rl] = mx;
    __syncthreads();
    if (kl == 0 && r < len) {
        mx = max(max(red[0][rl], red[1][rl]), max(red[2][rl], red[3][rl]));
        partial[static_cast<int64_t>(blockIdx.y) * len + r] = mx;
    }
}

__global__ void scale_finalize_kernel(const uint64_t* __restrict__ partial, int64_t len, int nkb, int tk,
                                      int32_t* __restrict__ ex) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= len) return;
    uint64_t mx = 0;
    for (int b = 0; b < nkb; ++b) mx = max(mx, partial[static_cast<int64_t>(b) * len + r]);
    ex[r] = exponent_from_maxbits(mx, tk);
}

constexpr int64_t kScaleRowsPerBlock = 64;

int64_t scale_kblocks(int64_t len, int64_t k) {
    const int64_t rb = (len + kScaleRowsPerBlock - 1) / kScaleRowsPerBlock;
    int64_t kb = (4096 + rb - 1) / rb;
    kb = std::max<int64_t>(1, std::min<int64_t>(kb, (k + 255) / 256));
    return std::min<int64_t>(kb, 65535);
}

// ---------------------------------------------------------------------------
// split
// ---------------------------------------------------------------------------

// Real split: each thread owns 16 consecutive k of one row -> one 16-byte store per plane.
// KCONTIG: row r is contiguous (X[p + r*ldx]); block = 256 chunks along the row(s).
// !KCONTIG: rows strided (X[r + p*ldx]); block = 32 rows x 8 chunks (128 k).
template <bool WIDE, bool KCONTIG>
__global__ void __launch_bounds__(256) split_real_kernel(SplitArgs a, int lg_tpr) {
    const TabSet* __restrict__ tb = a.tab;
    int64_t r, ck;
    if (KCONTIG) {
        const int tpr = 1 << lg_tpr;  // threads per row in a block
        r = static_cast<int64_t>(blockIdx.x) * (256 >> lg_tpr) + (threadIdx.x >> lg_tpr);
        ck = static_cast<int64_t>(blockIdx.y) * tpr + (threadIdx.x & (tpr - 1));
    } else {
        r = static_cast<int64_t>(blockIdx.x) * 32 + (threadIdx.x & 31);
        ck = static_cast<int64_t>(blockIdx.y) * 8 + (threadIdx.x >> 5);
    }
    const int64_t nck = a.kp >> 4;
    if (r >= a.rows_pad || ck >= nck) return;
    const int64_t p0 = ck * 16;
    const bool rowok = r < a.rows;
    const int32_t e = rowok ? a.ex[r] : 0;
    const bool rowzero = !rowok || e == INT32_MIN;

    uint32_t limbs[16][5];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int64_t p = p0 + q;
        uint64_t bits = 0;
        const bool z = rowzero || p >= a.k;
        if (!z) {
            const double* src = KCONTIG ? a.X + r * a.ldx + p : a.X + r + p * a.ldx;
            bits = static_cast<uint64_t>(__double_as_longlong(*src));
        }
        to_limbs<WIDE>(bits, e, z, limbs[q]);
    }
    int8_t* out = a.P + plane_offset(r, p0, a.kp);
    for (int t = 0; t < a.s; ++t) {
        uint32_t w[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            uint32_t acc = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) acc |= sym_byte<WIDE>(limbs[4 * g + q], tb, t) << (8 * q);
            w[g] = acc;
        }
        *reinterpret_cast<uint4*>(out + t * a.plane) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// Complex split: three plane groups Re, Im (negated for conj), Re+Im, each
// reduced to the symmetric range [-h, m-1-h].  4 consecutive k per thread.
template <bool WIDE, bool KCONTIG, bool CONJ>
__global__ void __launch_bounds__(256) split_cplx_kernel(SplitArgs a, int lg_tpr) {
    const TabSet* __restrict__ tb = a.tab;
    int64_t r, ck;
    if (KCONTIG) {
        const int tpr = 1 << lg_tpr;
        r = static_cast<int64_t>(blockIdx.x) * (256 >> lg_tpr) + (threadIdx.x >> lg_tpr);
        ck = static_cast<int64_t>(blockIdx.y) * tpr + (threadIdx.x & (tpr - 1));
    } else {
        r = static_cast<int64_t>(blockIdx.x) * 32 + (threadIdx.x & 31);
        ck = static_cast<int64_t>(blockIdx.y) * 8 + (threadIdx.x >> 5);
    }
    const int64_t nck = a.kp >> 2;
    if (r >= a.rows_pad || ck >= nck) return;
    const int64_t p0 = ck * 4;
    const bool rowok = r < a.rows;
    const int32_t e = rowok ? a.ex[r] : 0;
    const bool rowzero = !rowok || e == INT32_MIN;

    uint32_t lr[4][5], li[4][5];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int64_t p = p0 + q;
        uint64_t br = 0, bi = 0;
        const bool z = rowzero || p >= a.k;
        if (!z) {
            const double* src = KCONTIG ? a.X + 2 * (r * a.ldx + p) : a.X + 2 * (r + p * a.ldx);
            br = static_cast<uint64_t>(__double_as_longlong(src[0]));
            bi = static_cast<uint64_t>(__double_as_longlong(src[1]));
        }
        to_limbs<WIDE>(br, e, z, lr[q]);
        to_limbs<WIDE>(bi, e, z, li[q]);
    }
    int8_t* out = a.P + plane_offset(r, p0, a.kp);
    for (int t = 0; t < a.s; ++t) {
        const int32_t m = static_cast<int32_t>(tb->m[t]);
        const int32_t h = static_cast<int32_t>(tb->h[t]);
        uint32_t w0 = 0, w1 = 0, w2 = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int32_t s0 = sym_value<WIDE>(lr[q], tb, t);
            int32_t s1 = sym_value<WIDE>(li[q], tb, t);
            if (CONJ) s1 = -s1;
            int32_t ss = s0 + s1;
            ss = ss > m - 1 - h ? ss - m : ss;
            ss = ss < -h ? ss + m : ss;
            w0 |= (static_cast<uint32_t>(s0) & 0xFFu) << (8 * q);
            w1 |= (static_cast<uint32_t>(s1) & 0xFFu) << (8 * q);
            w2 |= (static_cast<uint32_t>(ss) & 0xFFu) << (8 * q);
        }
        *reinterpret_cast<uint32_t*>(out + t * a.plane) = w0;
        *reinterpret_cast<uint32_t*>(out + a.group + t * a.plane) = w1;
        *reinterpret_cast<uint32_t*>(out + 2 * a.group + t * a.plane) = w2;
    }
}

template <typename K>
void launch_split(hipStream_t st, K kernel, const SplitArgs& a, int elems_per_thread) {
    const int64_t nck = a.kp / elems_per_thread;
    dim3 grid, block(256);
    int lg = 0;
    if (a.kcontig) {
        int tpr = 1;
        while (tpr < 256 && tpr < nck) {
            tpr <<= 1;
            ++lg;
        }
        grid = dim3(static_cast<unsigned>((a.rows_pad + (256 / tpr) - 1) / (256 / tpr)),
                    static_cast<unsigned>((nck + tpr - 1) / tpr));
    } else {
        grid = dim3(static_cast<unsigned>((a.rows_pad + 31) / 32), static_cast<unsigned>((nck + 7) / 8));
    }
    hipLaunchKernelGGL(kernel, grid, block, 0, st, a, lg);
}

}  // namespace

size_t scale_partial_elems(int64_t len, int64_t k, bool kcontig) {
    if (kcontig) return 0;
    return static_cast<size_t>(scale_kblocks(len, k)) * static_cast<size_t>(len);
}

void scale(hipStream_t st, const double* X, int64_t ldx, int64_t len, int64_t k, bool kcontig, bool cplx, int tk,
           uint64_t* partial, int32_t* ex) {
    if (len <= 0) return;
    if (kcontig) {
        const int64_t ndbl = cplx ? 2 * k : k;
        hipLaunchKernelGGL(scale_kcontig_kernel, dim3(static_cast<unsigned>((len + 3) / 4)), dim3(256), 0, st, X, ldx,
                           len, ndbl, cplx ? 1 : 0, tk, ex);
    } else {
        const int64_t nkb = scale_kblocks(len, k);
        const int64_t kper = (k + nkb - 1) / nkb;
        const unsigned gx = static_cast<unsigned>((len + kScaleRowsPerBlock - 1) / kScaleRowsPerBlock);
        hipLaunchKernelGGL(scale_kstrided_kernel, dim3(gx, static_cast<unsigned>(nkb)), dim3(256), 0, st, X, ldx, len,
                           k, kper, cplx ? 1 : 0, partial);
        hipLaunchKernelGGL(scale_finalize_kernel, dim3(static_cast<unsigned>((len + 255) / 256)), dim3(256), 0, st,
                           partial, len, static_cast<int>(nkb), tk, ex);
    }
}

void split(hipStream_t st, const SplitArgs& a) {
    if (!a.cplx) {
        if (a.wide) {
            if (a.kcontig) launch_split(st, split_real_kernel<true, true>, a, 16);
            else launch_split(st, split_real_kernel<true, false>, a, 16);
        } else {
            if (a.kcontig) launch_split(st, split_real_kernel<false, true>, a, 16);
            else launch_split(st, split_real_kernel<false, false>, a, 16);
        }
        return;
    }
#define OZ2_CPLX(W, KC)                                                                         \
    if (a.conj) launch_split(st, split_cplx_kernel<W, KC, true>, a, 4);                     \
    else launch_split(st, split_cplx_kernel<W, KC, false>, a, 4);
    if (a.wide) {
        if (a.kcontig) { OZ2_CPLX(true, true) } else { OZ2_CPLX(true, false) }
    } else {
        if (a.kcontig) { OZ2_CPLX(false, true) } else { OZ2_CPLX(false, false) }
    }
#undef OZ2_CPLX
}

}  // namespace k
}  // namespace oz2
