// ozaki1.hip - kernels of Ozaki Scheme I (slicing), the comparison baseline of the
// accuracy benchmark: the counterparts of the reference's launch_slices and
// launch_horner (benchmarks/ozaki1_gpu.hpp:11-85).  The digit products run on this
// library's INT8 GEMM (igemm_int32) and the exponents come from its scale kernels.
//
// Digit planes: every row of op(X) (rows of A, columns of B) keeps its L balanced
// base-128 digits side by side along k in one extended plane of L * kp columns, in
// the GEMM's blocked layout (plane_offset): digit i of A at k-slot i, digit i of B at
// k-slot L - 1 - i.  The anti-diagonal d = sum_{a <= d} A_a B_{d-a} is then one GEMM
// over k-slots 0 .. d of A and L - 1 - d .. L - 1 of B.
#include "kernels.hpp"
#include "math.hpp"

namespace oz2 {
namespace k {
using namespace math;

namespace {

// (row, 16-byte chunk) of a thread: k-contiguous rows take 4 threads per row (one 64-k
// block), strided rows 16 rows x 16 chunks with lanes along the rows (coalesced loads)
template <bool KC>
__device__ __forceinline__ void slice_thread(uint32_t bx, uint32_t by, int64_t& r, int64_t& ck) {
    if (KC) {
        r = static_cast<int64_t>(bx) * 64 + (threadIdx.x >> 2);
        ck = static_cast<int64_t>(by) * 4 + (threadIdx.x & 3);
    } else {
        r = static_cast<int64_t>(bx) * 16 + (threadIdx.x & 15);
        ck = static_cast<int64_t>(by) * 16 + (threadIdx.x >> 4);
    }
}

// One thread: 16 k of one row (chunk c of 64-k block blk, k in block_k order), x = RNE(v 2^e)
// exactly (|x| <= 2^(7L - 2) < 2^62), cut into L digits d = ((x + 64) & 127) - 64,
// x = (x - d) >> 7, least significant first; one 16-byte store per digit.  Rows past
// `rows`, k past `k` and non-finite rows give zero digits.
template <bool KC, int L>
__global__ void __launch_bounds__(256) oz1_slice_kernel(Oz1SliceArgs a) {
    int64_t r, ck;
    slice_thread<KC>(blockIdx.x, blockIdx.y, r, ck);
    if (r >= a.rows_pad || ck >= a.kp / 16) return;
    const int64_t blk = ck >> 2;
    const int c = static_cast<int>(ck & 3);
    const bool rowok = r < a.rows;
    const int32_t ex = rowok ? a.ex[r] : 0;
    const bool zero_row = !rowok || ex == INT32_MIN;
    const int e = zero_row ? 0 : ex;
    const int64_t rl = rowok ? r : 0;
    int64_t x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int64_t p = blk * 64 + block_k(c, q);
        const int64_t pc = p < a.k ? p : a.k - 1;  // in bounds; zeroed below
        const uint64_t bits =
            static_cast<uint64_t>(__double_as_longlong(KC ? a.X[rl * a.ldx + pc] : a.X[rl + pc * a.ldx]));
        const int64_t mag = static_cast<int64_t>(scaled_mag62(bits, e));
        x[q] = (zero_row || p >= a.k) ? 0 : ((bits >> 63) ? -mag : mag);
    }
    const int64_t kblocks = static_cast<int64_t>(L) * (a.kp >> 6);
    for (int i = L - 1; i >= 0; --i) {
        uint32_t w[4];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            uint32_t b[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                int64_t& v = x[4 * g + q];
                const int64_t d = ((v + 64) & 127) - 64;
                v = (v - d) >> 7;
                b[q] = static_cast<uint32_t>(d);
            }
            w[g] = pack_lo_bytes(b[0], b[1], b[2], b[3]);
        }
        const int64_t slot = a.rev ? L - 1 - i : i;
        int8_t* out = a.P + chunk_offset(r, slot * (a.kp >> 6) + blk, c, kblocks << 6);
        *reinterpret_cast<uint4*>(out) = make_uint4(w[0], w[1], w[2], w[3]);
    }
}

// One entry of C per thread: S = sum of the int32 chunk sums of anti-diagonal d (int64),
// the double-double Horner step T = 128 T + S, and on the last diagonal
// C = ldexp(hi + lo, shift - D_i - E_j), NaN for a non-finite row or column.
__global__ void __launch_bounds__(256) oz1_horner_kernel(Oz1HornerArgs a) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (t >= a.m * a.n) return;
    const int64_t i = t % a.m, j = t / a.m;
    int64_t S = 0;
    for (int c = 0; c < a.nchunks; ++c) S += a.Ct[c * a.chunk + j * a.ldt + i];
    const int64_t idx = i + j * a.m;
    double hi = a.first ? 0.0 : a.Th[idx] * 128.0, lo = a.first ? 0.0 : a.Tl[idx] * 128.0;
    const double sd = static_cast<double>(S);
    const double s = hi + sd, bb = s - hi;
    const double err = (hi - (s - bb)) + (sd - bb) + lo;
    hi = s + err;
    lo = err - (hi - s);
    if (a.last) {
        const int32_t De = a.D[i], Ee = a.E[j];
        double v = ldexp(hi + lo, a.shift - De - Ee);
        if (De == INT32_MIN || Ee == INT32_MIN) v = __builtin_nan("");
        a.C[i + j * a.ldc] = v;
    } else {
        a.Th[idx] = hi;
        a.Tl[idx] = lo;
    }
}

template <bool KC>
void launch_slices_kc(hipStream_t st, const Oz1SliceArgs& a, int L) {
    const int64_t nck = a.kp / 16;
    const dim3 grid = KC ? dim3(static_cast<uint32_t>((a.rows_pad + 63) / 64), static_cast<uint32_t>((nck + 3) / 4))
                         : dim3(static_cast<uint32_t>((a.rows_pad + 15) / 16), static_cast<uint32_t>((nck + 15) / 16));
    switch (L) {
#define OZ1_CASE(LL) \
    case LL: hipLaunchKernelGGL((oz1_slice_kernel<KC, LL>), grid, dim3(256), 0, st, a); break;
        OZ1_CASE(2) OZ1_CASE(3) OZ1_CASE(4) OZ1_CASE(5) OZ1_CASE(6) OZ1_CASE(7) OZ1_CASE(8) OZ1_CASE(9)
#undef OZ1_CASE
    }
}

}  // namespace

void oz1_slices(hipStream_t st, const Oz1SliceArgs& a, int L) {
    if (a.rows_pad <= 0 || a.kp <= 0) return;
    if (a.kcontig) launch_slices_kc<true>(st, a, L);
    else launch_slices_kc<false>(st, a, L);
}

void oz1_horner(hipStream_t st, const Oz1HornerArgs& a) {
    if (a.m <= 0 || a.n <= 0) return;
    hipLaunchKernelGGL(oz1_horner_kernel, dim3(static_cast<uint32_t>((a.m * a.n + 255) / 256)), dim3(256), 0, st, a);
}

}  // namespace k
}  // namespace oz2
