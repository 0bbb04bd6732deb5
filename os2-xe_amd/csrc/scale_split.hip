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
// A thread's 16 k lie in one 16-byte chunk of a block (block_k order), so each
// plane store is one aligned vector store.  Rows of op(X)
// that are strided in memory (op(A) = N, op(B) = T) are read column-by-column
// with lanes along rows, so no separate transpose pass is needed.
#include <algorithm>
#include <stdexcept>

#include "grouped.hpp"
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
struct ScaleKcItem {
    const double* X;
    int64_t ldx, len, ndbl;
    int cplx, tk;
    int32_t* ex;
    uint32_t gx;
};
__device__ __forceinline__ void scale_kcontig_body(const ScaleKcItem& it, uint32_t bx) {
    const int lane = threadIdx.x & 63;
    const int64_t r = static_cast<int64_t>(bx) * 4 + (threadIdx.x >> 6);
    if (r >= it.len) return;
    const int64_t ndbl = it.ndbl;
    const double* row = it.X + (it.cplx ? 2 : 1) * r * it.ldx;
    uint64_t mx = 0;
    if ((reinterpret_cast<uintptr_t>(row) & 15) == 0) {
        const int64_t n2 = ndbl >> 1;
        const double2* row2 = reinterpret_cast<const double2*>(row);
        constexpr int U = 8;  // loads in flight per step, all at i < n2
        int64_t i = lane;
        for (; i + 64 * (U - 1) < n2; i += 64 * U) {
            double2 v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = row2[i + 64 * u];
#pragma unroll
            for (int u = 0; u < U; ++u)
                mx = max(mx, max(static_cast<uint64_t>(__double_as_longlong(v[u].x) & 0x7FFFFFFFFFFFFFFFll),
                                 static_cast<uint64_t>(__double_as_longlong(v[u].y) & 0x7FFFFFFFFFFFFFFFll)));
        }
        for (; i < n2; i += 64) {
            const double2 v = row2[i];
            const uint64_t a = __double_as_longlong(v.x) & 0x7FFFFFFFFFFFFFFFll;
            const uint64_t b = __double_as_longlong(v.y) & 0x7FFFFFFFFFFFFFFFll;
            mx = max(mx, max(a, b));
        }
        if ((ndbl & 1) && lane == 0) mx = max(mx, absbits(row + ndbl - 1));
    } else {  // a row that is only 8-byte aligned (odd leading dimension): single doubles
        constexpr int U = 8;
        int64_t i = lane;
        for (; i + 64 * (U - 1) < ndbl; i += 64 * U) {
            uint64_t v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = absbits(row + i + 64 * u);
#pragma unroll
            for (int u = 0; u < U; ++u) mx = max(mx, v[u]);
        }
        for (; i < ndbl; i += 64) mx = max(mx, absbits(row + i));
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) mx = max(mx, shfl_xor_u64(mx, o));
    if (lane == 0) it.ex[r] = exponent_from_maxbits(mx, it.tk);
}

// Rows strided in memory: block = 64 rows x 4 k-lanes over a k range; partial max per (k block, row).
struct ScaleKsItem {
    const double* X;
    int64_t ldx, len, k, kper;
    int cplx;
    uint64_t* partial;
    uint32_t gx;
};
__device__ __forceinline__ void scale_kstrided_body(const ScaleKsItem& it, uint32_t bx, uint32_t by) {
    __shared__ uint64_t red[4][64];
    const int rl = threadIdx.x & 63, kl = threadIdx.x >> 6;
    const int64_t r = static_cast<int64_t>(bx) * 64 + rl;
    const int64_t p0 = static_cast<int64_t>(by) * it.kper;
    const int64_t p1 = min(it.k, p0 + it.kper);
    const double* X = it.X;
    const int64_t ldx = it.ldx;
    uint64_t mx = 0;
    if (r < it.len) {
        // U loads in flight per step (all of them at k < p1), then the remainder one by one:
        // a single load per step leaves each wave waiting out the full memory latency
        constexpr int U = 8;
        int64_t p = p0 + kl;
        if (it.cplx) {
            for (; p + 4 * (U - 1) < p1; p += 4 * U) {
                double2 v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = *reinterpret_cast<const double2*>(X + 2 * (r + (p + 4 * u) * ldx));
#pragma unroll
                for (int u = 0; u < U; ++u)
                    mx = max(mx, max(static_cast<uint64_t>(__double_as_longlong(v[u].x) & 0x7FFFFFFFFFFFFFFFll),
                                     static_cast<uint64_t>(__double_as_longlong(v[u].y) & 0x7FFFFFFFFFFFFFFFll)));
            }
            for (; p < p1; p += 4) {
                const double2 v = *reinterpret_cast<const double2*>(X + 2 * (r + p * ldx));
                mx = max(mx, max(static_cast<uint64_t>(__double_as_longlong(v.x) & 0x7FFFFFFFFFFFFFFFll),
                                 static_cast<uint64_t>(__double_as_longlong(v.y) & 0x7FFFFFFFFFFFFFFFll)));
            }
        } else {
            for (; p + 4 * (U - 1) < p1; p += 4 * U) {
                uint64_t v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) v[u] = absbits(X + r + (p + 4 * u) * ldx);
#pragma unroll
                for (int u = 0; u < U; ++u) mx = max(mx, v[u]);
            }
            for (; p < p1; p += 4) mx = max(mx, absbits(X + r + p * ldx));
        }
    }
    red[kl][rl] = mx;
    __syncthreads();
    if (kl == 0 && r < it.len) {
        mx = max(max(red[0][rl], red[1][rl]), max(red[2][rl], red[3][rl]));
        it.partial[static_cast<int64_t>(by) * it.len + r] = mx;
    }
}

struct ScaleFinItem {
    const uint64_t* partial;
    int64_t len;
    int nkb, tk;
    int32_t* ex;
    uint32_t gx;
};
__device__ __forceinline__ void scale_finalize_body(const ScaleFinItem& it, uint32_t bx) {
    const int64_t r = static_cast<int64_t>(bx) * 256 + threadIdx.x;
    if (r >= it.len) return;
    uint64_t mx = 0;
    int b = 0;
    for (; b + 3 < it.nkb; b += 4) {  // four loads in flight, all at b < nkb
        uint64_t v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = it.partial[static_cast<int64_t>(b + u) * it.len + r];
        mx = max(mx, max(max(v[0], v[1]), max(v[2], v[3])));
    }
    for (; b < it.nkb; ++b) mx = max(mx, it.partial[static_cast<int64_t>(b) * it.len + r]);
    it.ex[r] = exponent_from_maxbits(mx, it.tk);
}

__global__ void __launch_bounds__(256) scale_kcontig_kernel(ScaleKcItem it) { scale_kcontig_body(it, blockIdx.x); }
__global__ void __launch_bounds__(256) scale_kstrided_kernel(ScaleKsItem it) {
    scale_kstrided_body(it, blockIdx.x, blockIdx.y);
}
__global__ void __launch_bounds__(256) scale_finalize_kernel(ScaleFinItem it) { scale_finalize_body(it, blockIdx.x); }

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

// Bit pattern of op(X)[r][p] for a row that exists (r < rows) and p < k.
template <bool KCONTIG, bool CPLX>
__device__ __forceinline__ uint64_t load_bits(const double* __restrict__ X, int64_t ldx, int64_t r, int64_t p, int part) {
    const int64_t i = KCONTIG ? r * ldx + p : r + p * ldx;
    return static_cast<uint64_t>(__double_as_longlong(CPLX ? X[2 * i + part] : X[i]));
}

// Thread slot of a split: Q bytes (Q = 16: a whole 16-byte chunk, Q = 8: half of one) of
// one row of the planes.  ck counts the slots along the row.
template <int Q>
struct SplitSlot {
    static_assert(Q == 16 || Q == 8, "a slot is a chunk or half a chunk");
    int64_t blk;  // 64-k block
    int c;        // chunk in the block
    int q0;       // first byte of the chunk filled by this slot
    int64_t kb;   // byte q of the chunk holds k = kb + 8 (q >> 1) + (q & 1)
    __device__ __forceinline__ explicit SplitSlot(int64_t ck) {
        constexpr int per_blk = 64 / Q;
        blk = ck / per_blk;
        const int ib = static_cast<int>(ck % per_blk);
        c = ib / (16 / Q);
        q0 = (ib % (16 / Q)) * Q;
        kb = blk * 64 + 2 * c;
    }
    __device__ __forceinline__ int64_t k_of(int q) const { return kb + 8 * ((q0 + q) >> 1) + ((q0 + q) & 1); }
};

// (row, slot) of a thread.  KCONTIG: row r is contiguous (X[p + r*ldx]); a block
// covers 256 slots along the row(s).  !KCONTIG: rows strided (X[r + p*ldx]); a block
// is 16 rows x 16 slots, lanes along rows, so a wave (16 rows x 4 slots) fills whole
// 1 KB plane blocks and reads one 128-byte line per load.
template <bool KCONTIG>
__device__ __forceinline__ void split_thread(int lg_tpr, uint32_t bx, uint32_t by, int64_t& r, int64_t& ck) {
    if (KCONTIG) {
        const int tpr = 1 << lg_tpr;  // threads per row in a block
        r = static_cast<int64_t>(bx) * (256 >> lg_tpr) + (threadIdx.x >> lg_tpr);
        ck = static_cast<int64_t>(by) * tpr + (threadIdx.x & (tpr - 1));
    } else {
        r = static_cast<int64_t>(bx) * 16 + (threadIdx.x & 15);
        ck = static_cast<int64_t>(by) * 16 + (threadIdx.x >> 4);
    }
}

// Q bytes packed little-endian from 32-bit words: one Q-byte store
template <int Q>
__device__ __forceinline__ void store_bytes(int8_t* p, const uint32_t (&w)[4]) {
    if (Q == 16) *reinterpret_cast<uint4*>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    else *reinterpret_cast<uint2*>(p) = make_uint2(w[0], w[1]);
}

// Real split: Q bytes (Q of the k of one row, block_k order) per thread -> one
// Q-byte store per plane.  All loads are issued before any arithmetic and stay
// inside the matrix: padding rows read row 0 and k past the end reads k - 1, and
// those values are then zeroed.
template <int FORM, bool KCONTIG, int Q>
__device__ __forceinline__ void split_real_body(const SplitArgs& a, int lg_tpr, uint32_t bx, uint32_t by) {
    const auto tb = const_tab(a.tab);
    int64_t r, ck;
    split_thread<KCONTIG>(lg_tpr, bx, by, r, ck);
    if (r >= a.rows_pad || ck >= a.kp / Q) return;
    const SplitSlot<Q> sl(ck);
    const bool rowok = r < a.rows;
    const int32_t ex = rowok ? a.ex[r] : 0;
    const bool rowzero = !rowok || ex == INT32_MIN;
    const int e = rowzero ? 0 : ex;
    const int64_t rl = rowok ? r : 0;

    uint64_t bits[Q];
    const bool full = sl.blk * 64 + 64 <= a.k;
    const double* rowp = a.X + rl * a.ldx;
    if (KCONTIG && full && (reinterpret_cast<uintptr_t>(rowp + sl.kb) & 15) == 0) {
#pragma unroll
        for (int i = 0; i < Q / 2; ++i) {
            const uint4 v = *reinterpret_cast<const uint4*>(rowp + sl.k_of(2 * i));
            bits[2 * i] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
            bits[2 * i + 1] = static_cast<uint64_t>(v.z) | (static_cast<uint64_t>(v.w) << 32);
        }
    } else if (full) {
#pragma unroll
        for (int q = 0; q < Q; ++q) bits[q] = load_bits<KCONTIG, false>(a.X, a.ldx, rl, sl.k_of(q), 0);
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int64_t p = sl.k_of(q);
            bits[q] = load_bits<KCONTIG, false>(a.X, a.ldx, rl, p < a.k ? p : a.k - 1, 0);
            bits[q] = p < a.k ? bits[q] : 0;
        }
    }

    // the 8, 12 or 14 bytes of the offset value (limb_form of tk)
    constexpr Limbs F = static_cast<Limbs>(FORM);
    uint32_t limbs[Q][4];
#pragma unroll
    for (int q = 0; q < Q; ++q) to_limbs<F>(rowzero ? 0 : bits[q], e, false, limbs[q]);
    int8_t* out = a.P + chunk_offset(r, sl.blk, sl.c, a.kp) + sl.q0;
    uint32_t w[4] = {0, 0, 0, 0};
    // plane 0 has m = 256 in every moduli set: its residue byte is the low byte of u
#pragma unroll
    for (int g = 0; g < Q / 4; ++g)
        w[g] = pack_lo_bytes(limbs[4 * g][0], limbs[4 * g + 1][0], limbs[4 * g + 2][0], limbs[4 * g + 3][0]);
    store_bytes<Q>(out, w);
    for (int t = 1; t < a.s; ++t) {
#pragma unroll
        for (int g = 0; g < Q / 4; ++g)
            w[g] = pack_lo_bytes(sym_residue<F>(limbs[4 * g], tb, t), sym_residue<F>(limbs[4 * g + 1], tb, t),
                                 sym_residue<F>(limbs[4 * g + 2], tb, t), sym_residue<F>(limbs[4 * g + 3], tb, t));
        store_bytes<Q>(out + t * a.plane, w);
    }
}

// Complex split: three plane groups Re, Im (negated for conj), Re+Im, each
// reduced to the symmetric range [-h, m-1-h]; slots, k order and loads as in the
// real split, one Q-byte store per plane of each group.
template <int FORM, bool KCONTIG, bool CONJ, int Q>
__device__ __forceinline__ void split_cplx_body(const SplitArgs& a, int lg_tpr, uint32_t bx, uint32_t by) {
    const auto tb = const_tab(a.tab);
    int64_t r, ck;
    split_thread<KCONTIG>(lg_tpr, bx, by, r, ck);
    if (r >= a.rows_pad || ck >= a.kp / Q) return;
    const SplitSlot<Q> sl(ck);
    const bool rowok = r < a.rows;
    const int32_t ex = rowok ? a.ex[r] : 0;
    const bool rowzero = !rowok || ex == INT32_MIN;
    const int e = rowzero ? 0 : ex;
    const int64_t rl = rowok ? r : 0;

    uint64_t br[Q], bi[Q];
    const bool full = sl.blk * 64 + 64 <= a.k;
    const double* rowp = a.X + 2 * rl * a.ldx;
    if (KCONTIG && full && (reinterpret_cast<uintptr_t>(rowp + 2 * sl.kb) & 15) == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const uint4 v = *reinterpret_cast<const uint4*>(rowp + 2 * sl.k_of(q));
            br[q] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
            bi[q] = static_cast<uint64_t>(v.z) | (static_cast<uint64_t>(v.w) << 32);
        }
    } else {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int64_t p = sl.k_of(q);
            const int64_t pc = p < a.k ? p : a.k - 1;
            br[q] = load_bits<KCONTIG, true>(a.X, a.ldx, rl, pc, 0);
            bi[q] = load_bits<KCONTIG, true>(a.X, a.ldx, rl, pc, 1);
            br[q] = p < a.k ? br[q] : 0;
            bi[q] = p < a.k ? bi[q] : 0;
        }
    }
    constexpr Limbs F = static_cast<Limbs>(FORM);
    uint32_t lr[Q][4], li[Q][4];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        to_limbs<F>(rowzero ? 0 : br[q], e, false, lr[q]);
        to_limbs<F>(rowzero ? 0 : bi[q], e, false, li[q]);
    }
    int8_t* out = a.P + chunk_offset(r, sl.blk, sl.c, a.kp) + sl.q0;
    for (int t = 0; t < a.s; ++t) {
        const int32_t m = static_cast<int32_t>(tb->m[t]);
        const int32_t h = static_cast<int32_t>(tb->h[t]);
        uint32_t w0[4] = {0, 0, 0, 0}, w1[4] = {0, 0, 0, 0}, w2[4] = {0, 0, 0, 0};
#pragma unroll
        for (int g = 0; g < Q / 4; ++g) {
            uint32_t b0[4], b1[4], b2[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int32_t s0 = sym_value<F>(lr[4 * g + q], tb, t);
                int32_t s1 = sym_value<F>(li[4 * g + q], tb, t);
                if (CONJ) s1 = -s1;
                int32_t ss = s0 + s1;
                ss = ss > m - 1 - h ? ss - m : ss;
                ss = ss < -h ? ss + m : ss;
                b0[q] = static_cast<uint32_t>(s0);
                b1[q] = static_cast<uint32_t>(s1);
                b2[q] = static_cast<uint32_t>(ss);
            }
            w0[g] = pack_lo_bytes(b0[0], b0[1], b0[2], b0[3]);
            w1[g] = pack_lo_bytes(b1[0], b1[1], b1[2], b1[3]);
            w2[g] = pack_lo_bytes(b2[0], b2[1], b2[2], b2[3]);
        }
        store_bytes<Q>(out + t * a.plane, w0);
        store_bytes<Q>(out + a.group + t * a.plane, w1);
        store_bytes<Q>(out + 2 * a.group + t * a.plane, w2);
    }
}

// split variants: bits 0-1 limb form (0, 1, 2), bit 2 kcontig, 3 complex, 4 conj
struct SplitItem {
    SplitArgs a;
    int lg;
    uint32_t gx;
};
// bytes per thread slot: half chunks for wide complex values (their 2 x 4 octet words
// per element would not fit 16 elements in registers), whole chunks otherwise
constexpr int split_q(int form, bool cplx) { return form != 0 && cplx ? 8 : 16; }
template <int V>
__device__ __forceinline__ void split_body(const SplitItem& it, uint32_t bx, uint32_t by) {
    constexpr int FM = V & 3;
    constexpr bool KC = (V >> 2) & 1, CX = (V >> 3) & 1, CJ = (V >> 4) & 1;
    if constexpr (CX) split_cplx_body<FM, KC, CJ, split_q(FM, true)>(it.a, it.lg, bx, by);
    else split_real_body<FM, KC, split_q(FM, false)>(it.a, it.lg, bx, by);
}
template <int V>
__global__ void __launch_bounds__(256) split_kernel(SplitItem it) { split_body<V>(it, blockIdx.x, blockIdx.y); }

int split_variant(const SplitArgs& a) {
    return a.form | (a.kcontig ? 4 : 0) | (a.cplx ? 8 : 0) | (a.cplx && a.conj ? 16 : 0);
}

// grid of one split: KCONTIG: rows x slot groups (tpr threads per row); strided: 16 rows x 16 slots.
// KCONTIG takes at most one 64-k block per row (tpr <= 64 / Q), so a wave fills whole
// 1 KB plane blocks (16 rows x 64 k; 8 rows for 8-byte slots) and every 128-byte line of
// a plane is written by one wave.  With a whole row per workgroup (256 slots) the two
// rows sharing a line came from different workgroups, usually on different XCDs, and
// the L2s wrote half lines: 0.79 against 0.69 ms for both operands at 8192^3 s = 16.
SplitItem split_item(const SplitArgs& a, uint32_t& gy) {
    const int q = split_q(a.form, a.cplx);
    const int64_t nck = a.kp / q;
    SplitItem it{a, 0, 0};
    if (a.kcontig) {
        int tpr = 1;
        while (tpr < 64 / q && tpr < nck) {
            tpr <<= 1;
            ++it.lg;
        }
        it.gx = static_cast<uint32_t>((a.rows_pad + (256 / tpr) - 1) / (256 / tpr));
        gy = static_cast<uint32_t>((nck + tpr - 1) / tpr);
    } else {
        it.gx = static_cast<uint32_t>((a.rows_pad + 15) / 16);
        gy = static_cast<uint32_t>((nck + 15) / 16);
    }
    return it;
}

template <int V>
void launch_split_v(hipStream_t st, const SplitItem& it, uint32_t gy) {
    hipLaunchKernelGGL((split_kernel<V>), dim3(it.gx, gy), dim3(256), 0, st, it);
}

// ---- grouped forms: block b of the launch runs block (b - start[p]) of item p
enum GroupKind { kScaleKc = 0, kScaleKs = 1, kScaleFin = 2, kSplitBase = 16 };

template <typename Item, int K>
__global__ void __launch_bounds__(256) grouped_scale_kernel(const Item* __restrict__ items,
                                                            const uint32_t* __restrict__ start, uint32_t n) {
    const uint32_t p = owner_of(start, n, blockIdx.x);
    const Item& it = items[p];
    const uint32_t l = blockIdx.x - start[p];
    if constexpr (K == kScaleKc) scale_kcontig_body(it, l);
    else if constexpr (K == kScaleKs) scale_kstrided_body(it, l % it.gx, l / it.gx);
    else scale_finalize_body(it, l);
}

template <int V>
__global__ void __launch_bounds__(256) grouped_split_kernel(const SplitItem* __restrict__ items,
                                                            const uint32_t* __restrict__ start, uint32_t n) {
    const uint32_t p = owner_of(start, n, blockIdx.x);
    const SplitItem& it = items[p];
    const uint32_t l = blockIdx.x - start[p];
    split_body<V>(it, l % it.gx, l / it.gx);
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
        const ScaleKcItem it{X, ldx, len, cplx ? 2 * k : k, cplx ? 1 : 0, tk, ex, static_cast<uint32_t>((len + 3) / 4)};
        hipLaunchKernelGGL(scale_kcontig_kernel, dim3(it.gx), dim3(256), 0, st, it);
    } else {
        const int64_t nkb = scale_kblocks(len, k);
        const int64_t kper = (k + nkb - 1) / nkb;
        const ScaleKsItem it{X, ldx, len, k, kper, cplx ? 1 : 0, partial,
                             static_cast<uint32_t>((len + kScaleRowsPerBlock - 1) / kScaleRowsPerBlock)};
        hipLaunchKernelGGL(scale_kstrided_kernel, dim3(it.gx, static_cast<unsigned>(nkb)), dim3(256), 0, st, it);
        const ScaleFinItem fin{partial, len, static_cast<int>(nkb), tk, ex, static_cast<uint32_t>((len + 255) / 256)};
        hipLaunchKernelGGL(scale_finalize_kernel, dim3(fin.gx), dim3(256), 0, st, fin);
    }
}

void split(hipStream_t st, const SplitArgs& a) {
    uint32_t gy = 0;
    const SplitItem it = split_item(a, gy);
    switch (split_variant(a)) {
#define OZ2_V(V) case V: launch_split_v<V>(st, it, gy); break;
#define OZ2_F(F) OZ2_V(F) OZ2_V(F + 4) OZ2_V(F + 8) OZ2_V(F + 12) OZ2_V(F + 24) OZ2_V(F + 28)
        OZ2_F(0) OZ2_F(1) OZ2_F(2)
#undef OZ2_F
#undef OZ2_V
        default: throw std::logic_error("oz2: no split kernel for this variant");
    }
}

void prep_scale(std::vector<GroupedLaunch>& out, const std::vector<ScaleJob>& jobs) {
    GroupedLaunch kc, ks, fin;
    kc.kind = kScaleKc;
    ks.kind = kScaleKs;
    fin.kind = kScaleFin;
    for (const ScaleJob& j : jobs) {
        if (j.len <= 0) continue;
        if (j.kcontig) {
            const ScaleKcItem it{j.X, j.ldx, j.len, j.cplx ? 2 * j.k : j.k, j.cplx ? 1 : 0, j.tk, j.ex,
                                 static_cast<uint32_t>((j.len + 3) / 4)};
            kc.add(it, it.gx);
        } else {
            const int64_t nkb = scale_kblocks(j.len, j.k);
            const int64_t kper = (j.k + nkb - 1) / nkb;
            const ScaleKsItem it{j.X, j.ldx, j.len, j.k, kper, j.cplx ? 1 : 0, j.partial,
                                 static_cast<uint32_t>((j.len + kScaleRowsPerBlock - 1) / kScaleRowsPerBlock)};
            ks.add(it, it.gx * static_cast<uint32_t>(nkb));
            const ScaleFinItem f{j.partial, j.len, static_cast<int>(nkb), j.tk, j.ex,
                                 static_cast<uint32_t>((j.len + 255) / 256)};
            fin.add(f, f.gx);
        }
    }
    for (GroupedLaunch* g : {&kc, &ks, &fin})
        if (g->n) out.push_back(std::move(*g));
}

void prep_split(std::vector<GroupedLaunch>& out, const std::vector<SplitArgs>& jobs) {
    GroupedLaunch by[32];
    for (const SplitArgs& a : jobs) {
        if (a.rows_pad <= 0) continue;
        uint32_t gy = 0;
        const SplitItem it = split_item(a, gy);
        const int v = split_variant(a);
        by[v].kind = kSplitBase + v;
        by[v].add(it, it.gx * gy);
    }
    for (GroupedLaunch& g : by)
        if (g.n) out.push_back(std::move(g));
}

bool launch_scale_split_grouped(hipStream_t st, const GroupedLaunch& g) {
    if (g.n == 0) return true;
    const dim3 grid(g.blocks), block(256);
    switch (g.kind) {
        case kScaleKc:
            hipLaunchKernelGGL((grouped_scale_kernel<ScaleKcItem, kScaleKc>), grid, block, 0, st,
                               static_cast<const ScaleKcItem*>(g.d_items), g.d_start, g.n);
            return true;
        case kScaleKs:
            hipLaunchKernelGGL((grouped_scale_kernel<ScaleKsItem, kScaleKs>), grid, block, 0, st,
                               static_cast<const ScaleKsItem*>(g.d_items), g.d_start, g.n);
            return true;
        case kScaleFin:
            hipLaunchKernelGGL((grouped_scale_kernel<ScaleFinItem, kScaleFin>), grid, block, 0, st,
                               static_cast<const ScaleFinItem*>(g.d_items), g.d_start, g.n);
            return true;
#define OZ2_V(V)                                                                                  \
    case kSplitBase + V:                                                                          \
        hipLaunchKernelGGL((grouped_split_kernel<V>), grid, block, 0, st,                         \
                           static_cast<const SplitItem*>(g.d_items), g.d_start, g.n);             \
        return true;
        // the batch takes real products only, whose plans all have tk <= 91 (forms 0 and 1)
        OZ2_V(0) OZ2_V(1) OZ2_V(4) OZ2_V(5)
#undef OZ2_V
        default: return false;
    }
}

}  // namespace k
}  // namespace oz2
