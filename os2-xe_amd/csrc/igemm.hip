// igemm.hip - batched INT8 x INT8 -> INT32 GEMM on gfx950 MFMA with the
// residue epilogue of Ozaki Scheme II.
//
// Replaces the reference's oneMKL gemm_batch int8 -> int32 calls
// (src/ozaki2_gpu.hpp:1023-1042 DGEMM, 880-891 ZGEMM): plane t computes
// Ct = op(A)_t * op(B)_t exactly in int32 (k-chunks <= 130944 keep
// |sum| <= kc * 2^14 < 2^31).  Instead of storing int32, the epilogue reduces
// every entry mod m_t to a byte, the only information the CRT needs
// (hpp:590-601), cutting the CRT's reads from 4s to s bytes per entry.
//
// Kernel shape: macro tile BM x BN, 256 threads = 4 waves in 2 x 2, each wave
// (BM/2) x (BN/2) built from v_mfma_i32_32x32x32_i8 blocks.  k advances 128
// bytes per LDS stage, two stages double-buffered; 16-byte chunks of a row
// are XOR-swizzled so both the 16-byte stores and the MFMA operand reads are
// free of LDS bank conflicts.
//
// MFMA 32x32x32 i8 operand layout (checked on gfx950): lane l holds row
// (l mod 32) of A / column (l mod 32) of B with 16 k-bytes at k = 16*(l/32);
// accumulator v of lane l is C[8*(v/4) + 4*(l/32) + v%4][l mod 32].
#include "kernels.hpp"
#include "math.hpp"

namespace oz2 {
namespace k {
using namespace math;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

namespace {

constexpr int KT = 128;  // k bytes per LDS stage (8 chunks of 16 bytes)

__device__ __forceinline__ int swz(int row, int c) { return c ^ ((row ^ (row >> 3)) & 7); }

// blockIdx -> work item with the 8 XCDs of MI355X each given a contiguous
// range (workgroups are dispatched round-robin over XCDs, so consecutive ids
// land on different L2s; remapping keeps tiles that share A/B rows together).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    if ((n & 7) != 0) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

template <int BM, int BN>
__global__ void __launch_bounds__(256, 1) igemm_residue_kernel(GemmArgs p) {
    constexpr int WTM = BM / 2, WTN = BN / 2;
    constexpr int FM = WTM / 32, FN = WTN / 32;
    constexpr int LA = BM / 32, LB = BN / 32;
    constexpr int STAGE = (BM + BN) * 8;  // int4 per stage
    __shared__ v4i smem[2 * STAGE];

    const uint32_t nblk = gridDim.x;
    const uint32_t bid = xcd_remap(blockIdx.x, nblk);
    const uint32_t tiles = static_cast<uint32_t>(p.tiles_m) * static_cast<uint32_t>(p.tiles_n);
    const uint32_t plane = bid / tiles;
    uint32_t tt = bid - plane * tiles;
    // grouped order: columns of 8 tile-rows, walking tile-columns inside a group
    uint32_t tm, tn;
    {
        const uint32_t G = 8;
        const uint32_t per_group = G * p.tiles_n;
        const uint32_t g = tt / per_group;
        const uint32_t first = g * G;
        const uint32_t gs = min(G, static_cast<uint32_t>(p.tiles_m) - first);
        const uint32_t in = tt - g * per_group;
        tm = first + in % gs;
        tn = in / gs;
    }

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int lrow = tid >> 3, lchk = tid & 7;

    const int8_t* Ag = p.A + static_cast<int64_t>(plane) * p.plane_a + static_cast<int64_t>(tm) * BM * p.kp + p.k0 +
                       static_cast<int64_t>(lrow) * p.kp + lchk * 16;
    const int8_t* Bg = p.B + static_cast<int64_t>(plane) * p.plane_b + static_cast<int64_t>(tn) * BN * p.kp + p.k0 +
                       static_cast<int64_t>(lrow) * p.kp + lchk * 16;
    const int64_t rowstep = 32 * p.kp;

    v4i ra[LA], rb[LB];
    v16i acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;

#define OZ2_GLOAD(KT_)                                                                             \
    {                                                                                              \
        const int64_t ko = static_cast<int64_t>(KT_) * KT;                                         \
        _Pragma("unroll") for (int j = 0; j < LA; ++j) ra[j] =                                      \
            *reinterpret_cast<const v4i*>(Ag + j * rowstep + ko);       \
        _Pragma("unroll") for (int j = 0; j < LB; ++j) rb[j] =                                      \
            *reinterpret_cast<const v4i*>(Bg + j * rowstep + ko);       \
    }
#define OZ2_SSTORE(ST_)                                                                            \
    {                                                                                              \
        v4i* sa_ = smem + (ST_) * STAGE;                                                          \
        v4i* sb_ = sa_ + BM * 8;                                                                  \
        _Pragma("unroll") for (int j = 0; j < LA; ++j) {                                           \
            const int row = lrow + 32 * j;                                                         \
            sa_[row * 8 + swz(row, lchk)] = ra[j];                                                 \
        }                                                                                          \
        _Pragma("unroll") for (int j = 0; j < LB; ++j) {                                           \
            const int row = lrow + 32 * j;                                                         \
            sb_[row * 8 + swz(row, lchk)] = rb[j];                                                 \
        }                                                                                          \
    }

    const int nkt = static_cast<int>(p.klen / KT);
    OZ2_GLOAD(0)
    OZ2_SSTORE(0)
    __syncthreads();
    const int fr = lane & 31, fc = lane >> 5;
    for (int kt = 0; kt < nkt; ++kt) {
        const int st = kt & 1;
        if (kt + 1 < nkt) OZ2_GLOAD(kt + 1)
        const v4i* sa = smem + st * STAGE;
        const v4i* sb = sa + BM * 8;
#pragma unroll
        for (int ks = 0; ks < KT / 32; ++ks) {
            const int c = ks * 2 + fc;
            v4i fa[FM], fb[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                const int row = wm * WTM + i * 32 + fr;
                fa[i] = sa[row * 8 + swz(row, c)];
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                const int row = wn * WTN + j * 32 + fr;
                fb[j] = sb[row * 8 + swz(row, c)];
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[i], fb[j], acc[i][j], 0, 0, 0);
        }
        if (kt + 1 < nkt) OZ2_SSTORE(st ^ 1)
        __syncthreads();
    }
#undef OZ2_GLOAD
#undef OZ2_SSTORE

    // epilogue: int32 -> residue mod m_t, 4 rows packed per 32-bit store
    const int t = static_cast<int>(plane % static_cast<uint32_t>(p.s));
    const TabSet* __restrict__ tb = p.tab;
    uint8_t* Rb = p.R + static_cast<int64_t>(plane) * p.plane_r;
    const int64_t col0 = static_cast<int64_t>(tn) * BN + wn * WTN + fr;
    const int64_t row0 = static_cast<int64_t>(tm) * BM + wm * WTM + 4 * fc;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                uint32_t w = 0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    w |= acc_residue(static_cast<uint32_t>(acc[i][j][4 * q + b]), tb, t) << (8 * b);
                }
                *reinterpret_cast<uint32_t*>(Rb + (col0 + j * 32) * p.ldr + row0 + i * 32 + 8 * q) = w;
            }
}

}  // namespace

int tile_rows(Tile t) { return t == Tile::T256 ? 256 : t == Tile::T128 ? 128 : 64; }

void igemm_residue(hipStream_t st, Tile tile, const GemmArgs& a) {
    const uint32_t nblk = static_cast<uint32_t>(a.tiles_m) * a.tiles_n * a.nplanes;
    if (nblk == 0 || a.klen <= 0) return;
    switch (tile) {
        case Tile::T256: hipLaunchKernelGGL((igemm_residue_kernel<256, 256>), dim3(nblk), dim3(256), 0, st, a); break;
        case Tile::T128: hipLaunchKernelGGL((igemm_residue_kernel<128, 128>), dim3(nblk), dim3(256), 0, st, a); break;
        case Tile::T64: hipLaunchKernelGGL((igemm_residue_kernel<64, 64>), dim3(nblk), dim3(256), 0, st, a); break;
    }
}

}  // namespace k
}  // namespace oz2
