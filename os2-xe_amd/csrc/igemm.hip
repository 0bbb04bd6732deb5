// igemm.hip - batched INT8 x INT8 -> INT32 GEMM on gfx950 MFMA with the
// residue epilogue of Ozaki Scheme II.
//
// Replaces the reference's oneMKL gemm_batch int8 -> int32 calls
// (src/ozaki2_gpu.hpp:1023-1042 DGEMM, 880-891 ZGEMM): plane t computes
// Ct = op(A)_t * op(B)_t exactly in int32 (k-chunks <= 65408 keep
// |sum| <= kc * 2^14 < 2^30).  Instead of storing int32, the epilogue reduces
// every entry mod m_t to a byte, the only information the CRT needs
// (hpp:590-601), cutting the CRT's reads from 4s to s bytes per entry.
//
// Operand planes are stored blocked (plane_offset() in kernels.hpp): 1 KB per
// 16 rows x 64 k, the four 16-byte k-chunks of a row XOR-permuted by
// (row >> 2) & 3.  One stage of a macro tile is then a run of whole 1 KB
// blocks, each fetched by one global_load_lds_dwordx4 (lane l -> bytes
// 16l..16l+15) straight into LDS in the same order, and the MFMA operand reads
// (lane l: row l mod 32, chunk 2 ks + l / 32) are free of bank conflicts:
// ds_read_b128 serves 16 lanes per pass from the quads {0,3,5,6} and {1,2,4,7}
// of each half-wave over 64 banks, and rows 4q..4q+3 of such a quad set land
// on four distinct chunk positions.
//
// Kernel shape: macro tile BM x BN, 256 threads = 4 waves in 2 x 2, each wave
// (BM/2) x (BN/2) from v_mfma_i32_32x32x32_i8 blocks.  k advances 64 bytes per
// stage; NS stages are in flight (global -> LDS by DMA, no staging
// registers) and the operand fragments are double-buffered in registers, so
// each half-iteration issues, in a fixed order, the MFMAs of the current
// stage, the ds_reads of the next stage and the DMAs of stage + NS - 1; a
// wave's VMEM and LDS issue never sits in front of a run of its MFMAs.
// 256 x 256 tiles run one workgroup per CU with NS = 4 (128 KB of LDS);
// 256 x 128 tiles run two per CU with NS = 3 (72 KB each): with short k loops
// the residue epilogue of one workgroup then overlaps the MFMAs of the other.
//
// MFMA 32x32x32 i8 operand layout (checked on gfx950): lane l holds row
// (l mod 32) of A / column (l mod 32) of B with 16 k-bytes at k = 16*(l/32);
// accumulator v of lane l is C[8*(v/4) + 4*(l/32) + v%4][l mod 32].
#include <algorithm>

#include "grouped.hpp"
#include "kernels.hpp"
#include "math.hpp"

namespace oz2 {
namespace k {
using namespace math;

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef __attribute__((address_space(3))) void* lds_ptr;
typedef __attribute__((address_space(1))) void* glb_ptr;

namespace {

constexpr int KT = 64;   // k bytes per LDS stage = one plane block column

// s_waitcnt immediate waiting for vmcnt <= n only (gfx9 encoding: vmcnt in bits
// [3:0] and [15:14]; expcnt [6:4] and lgkmcnt [11:8] left at their maximum)
constexpr int vm_wait(int n) { return 0x0F70 | (n & 15) | ((n >> 4) << 14); }
constexpr int kLgkmWait0 = 0xC07F & ~0x0F00;  // lgkmcnt(0) only

// blockIdx -> work item with the 8 XCDs of MI355X each given a contiguous
// range (workgroups are dispatched round-robin over XCDs, so consecutive ids
// land on different L2s; remapping keeps tiles that share A/B rows together).
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t n) {
    if ((n & 7) != 0) return b;
    return (b & 7) * (n >> 3) + (b >> 3);
}

// The same per plane: work items stay plane-major and each plane's tiles are
// split over the XCDs in contiguous ranges, so all XCDs work on one plane at a
// time and share its A and B planes in the last-level cache (with whole ranges
// of planes per XCD, 8 plane pairs compete for it).
__device__ __forceinline__ uint32_t plane_xcd_remap(uint32_t b, uint32_t tiles, uint32_t n) {
    if ((tiles & 7) != 0) return xcd_remap(b, n);
    const uint32_t plane = b / tiles;
    const uint32_t r = b - plane * tiles;
    return plane * tiles + (r & 7) * (tiles >> 3) + (r >> 3);
}

// Residues of one wave tile (WTM x WTN) staged through LDS: 4 residues per
// 32-bit LDS store into a [col][row] byte tile whose 16-byte chunks are
// XOR-permuted by column, then 16-byte global stores that write whole runs of
// WTM bytes of each column of R.
template <int FM, int FN>
__device__ __forceinline__ void residue_epilogue(const GemmArgs& p, v16i (&acc)[FM][FN], uint32_t plane,
                                                 int64_t row0, int64_t col0, char* wlds, int lane) {
    constexpr int WTM = FM * 32, WTN = FN * 32, NCH = WTM / 16;
    const int t = static_cast<int>((static_cast<uint32_t>(p.plane0) + plane) % static_cast<uint32_t>(p.s));
    const auto tb = const_tab(p.tab);
    const uint32_t m = tb->m[t];
    const int fr = lane & 31, fc = lane >> 5;
    auto lds_at = [&](int c, int r) { return wlds + c * WTM + ((((r >> 4) ^ (c & (NCH - 1))) << 4) | (r & 15)); };
    if (m == 256) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint32_t w = pack_lo_bytes(static_cast<uint32_t>(acc[i][j][4 * q]), static_cast<uint32_t>(acc[i][j][4 * q + 1]),
                                                     static_cast<uint32_t>(acc[i][j][4 * q + 2]), static_cast<uint32_t>(acc[i][j][4 * q + 3]));
                    *reinterpret_cast<uint32_t*>(lds_at(j * 32 + fr, i * 32 + 8 * q + 4 * fc)) = w;
                }
    } else {
        const uint32_t mu = tb->mu[t], c16 = tb->c16[t], c30 = tb->c30[t];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    uint32_t r[4];
#pragma unroll
                    for (int b = 0; b < 4; ++b) r[b] = residue_m30(static_cast<uint32_t>(acc[i][j][4 * q + b]), m, mu, c16, c30);
                    *reinterpret_cast<uint32_t*>(lds_at(j * 32 + fr, i * 32 + 8 * q + 4 * fc)) = pack_lo_bytes(r[0], r[1], r[2], r[3]);
                }
            }
    }
    __syncthreads();
    uint8_t* Rb = p.R + static_cast<int64_t>(plane) * p.plane_r;
    constexpr int CPI = 64 / NCH;  // columns per pass
#pragma unroll
    for (int it = 0; it < WTN / CPI; ++it) {
        const int c = it * CPI + lane / NCH, ch = lane % NCH;
        const v4i v = *reinterpret_cast<const v4i*>(wlds + c * WTM + ((ch ^ (c & (NCH - 1))) << 4));
        *reinterpret_cast<v4i*>(Rb + (col0 + c) * p.ldr + row0 + ch * 16) = v;
    }
}

// tile (tm, tn) of plane `plane` inside a problem's grid: columns of 8 tile-rows,
// walking tile-columns inside a group (tiles that run together share A and B rows)
__device__ __forceinline__ void tile_coords(const GemmArgs& p, uint32_t bid, uint32_t& plane, uint32_t& tm,
                                            uint32_t& tn) {
    const uint32_t tiles = static_cast<uint32_t>(p.tiles_m) * static_cast<uint32_t>(p.tiles_n);
    plane = bid / tiles;
    const uint32_t tt = bid - plane * tiles;
    const uint32_t G = 8;
    const uint32_t per_group = G * p.tiles_n;
    const uint32_t g = tt / per_group;
    const uint32_t first = g * G;
    const uint32_t gs = min(G, static_cast<uint32_t>(p.tiles_m) - first);
    const uint32_t in = tt - g * per_group;
    tm = first + in % gs;
    tn = in / gs;
}

template <int BM, int BN, int NS>
struct TileShape {
    static constexpr int STAGE = (BM + BN) * KT;  // bytes
    static constexpr int SMEM = NS * STAGE;
};

// int32 sums of one wave tile stored as they are: accumulator v of lane l holds row
// 8 (v / 4) + 4 (l / 32) + v % 4 and column l % 32 of its 32 x 32 block, so each lane
// stores four 16-byte runs of 4 rows per block
template <int FM, int FN>
__device__ __forceinline__ void int32_epilogue(const GemmArgs& p, v16i (&acc)[FM][FN], uint32_t plane, int64_t row0,
                                               int64_t col0, int lane) {
    const int fr = lane & 31, fc = lane >> 5;
    int32_t* Rp = p.R32 + static_cast<int64_t>(plane) * p.plane_r;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const v4i v = {acc[i][j][4 * q], acc[i][j][4 * q + 1], acc[i][j][4 * q + 2], acc[i][j][4 * q + 3]};
                *reinterpret_cast<v4i*>(Rp + (col0 + j * 32 + fr) * p.ldr + row0 + i * 32 + 8 * q + 4 * fc) = v;
            }
}

// one BM x BN macro tile of one plane: k loop and residue (or int32) epilogue; smem holds NS stages
struct NoHook {
    __device__ void operator()() const {}
};

// after_loop runs once every DMA of the tile has landed (vm_wait(0)) and before the epilogue
template <int BM, int BN, int NS, bool I32 = false, typename Hook = NoHook>
__device__ __forceinline__ void gemm_tile(const GemmArgs& p, uint32_t plane, uint32_t tm, uint32_t tn, v4i* smem,
                                          const Hook& after_loop = Hook{}) {
    constexpr int WTM = BM / 2, WTN = BN / 2;
    constexpr int FM = WTM / 32, FN = WTN / 32;
    constexpr int LA = BM / 64, LB = BN / 64;  // 1 KB DMAs per wave per stage
    constexpr int STAGE = TileShape<BM, BN, NS>::STAGE;
    constexpr int NR = 2 * (FM + FN);           // fragment reads per stage
    constexpr int NMF = 2 * FM * FN;            // MFMAs per stage
    constexpr int ND = LA + LB;                 // DMAs per stage
    constexpr int VM_AHEAD = (NS - 3) * ND;    // DMAs still in flight when a stage must have landed
    static_assert(WTM * WTN * 4 <= NS * STAGE, "epilogue staging must fit in the stage buffers");

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int fr = lane & 31, fc = lane >> 5;
    const int nkt = static_cast<int>(p.klen / KT);  // even: klen is a multiple of 128
    const int64_t kblocks = p.kp / KT;              // plane blocks per 16 rows
    // this lane's 16 bytes of each of its DMA blocks at stage 0
    const int8_t* Ag = p.A + static_cast<int64_t>(plane) * p.plane_a +
                       ((static_cast<int64_t>(tm) * (BM / 16) + wave * LA) * kblocks + p.k0 / KT) * 1024 + lane * 16;
    const int8_t* Bg = p.B + static_cast<int64_t>(plane) * p.plane_b +
                       ((static_cast<int64_t>(tn) * (BN / 16) + wave * LB) * kblocks + (p.k0 + p.kb_extra) / KT) * 1024 +
                       lane * 16;
    const int64_t bstride = kblocks * 1024;  // next 16-row block
    char* sbase = reinterpret_cast<char*>(smem);
    // LDS byte offsets of this lane's fragments (row-major 64-byte rows; the
    // chunk permutation depends on row bits 2-3 only, so fragment i is +2 KB * i)
    int offa[2], offb[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
        const int ra = wm * WTM + fr, rb = wn * WTN + fr;
        offa[ks] = ra * 64 + (((2 * ks + fc) ^ ((ra >> 2) & 3)) << 4);
        offb[ks] = BM * KT + rb * 64 + (((2 * ks + fc) ^ ((rb >> 2) & 3)) << 4);
    }

    v16i acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[i][j][v] = 0;
    v4i fa0[2][FM], fb0[2][FN], fa1[2][FM], fb1[2][FN];

    // stage KT_ (clamped to the last one: past the end the DMA refills a
    // buffer that is never read again, keeping the vmcnt arithmetic uniform)
#define OZ2_DMA(KT_, D_)                                                                                         \
    {                                                                                                            \
        char* st_ = sbase + ((KT_) % NS) * STAGE;                                                               \
        const int64_t ko_ = static_cast<int64_t>(min((KT_), nkt - 1)) * 1024;                                    \
        if ((D_) < LA)                                                                                           \
            __builtin_amdgcn_global_load_lds((glb_ptr)(Ag + (D_) * bstride + ko_),                               \
                                             (lds_ptr)(st_ + (wave * LA + (D_)) * 1024), 16, 0, 0);            \
        else                                                                                                     \
            __builtin_amdgcn_global_load_lds((glb_ptr)(Bg + ((D_) - LA) * bstride + ko_),                        \
                                             (lds_ptr)(st_ + BM * KT + (wave * LB + (D_) - LA) * 1024), 16, 0, 0); \
    }
#define OZ2_READ(NA, NB, KT_, R_)                                                                                \
    {                                                                                                            \
        const char* s_ = sbase + ((KT_) % NS) * STAGE;                                                          \
        const int ks_ = (R_) / (FM + FN), w_ = (R_) % (FM + FN);                                                 \
        if (w_ < FM) NA[ks_][w_] = *reinterpret_cast<const v4i*>(s_ + offa[ks_] + w_ * 2048);                   \
        else NB[ks_][w_ - FM] = *reinterpret_cast<const v4i*>(s_ + offb[ks_] + (w_ - FM) * 2048);               \
    }
    // one stage: MFMAs on (FA, FB), reads of stage KTR into (NA, NB), DMAs of stage KTI,
    // spread evenly in a fixed order
#define OZ2_HALF(FA, FB, NA, NB, KTR, KTI)                                                                       \
    {                                                                                                            \
        _Pragma("unroll") for (int u = 0; u < NMF; ++u) {                                                        \
            _Pragma("unroll") for (int r = 0; r < NR; ++r) if (r * NMF / NR == u) OZ2_READ(NA, NB, KTR, r)       \
            __builtin_amdgcn_sched_barrier(0);                                                                   \
            {                                                                                                    \
                const int ks = u / (FM * FN), i = (u / FN) % FM, j = u % FN;                                     \
                acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(FA[ks][i], FB[ks][j], acc[i][j], 0, 0, 0);     \
            }                                                                                                    \
            __builtin_amdgcn_sched_barrier(0);                                                                   \
            _Pragma("unroll") for (int d = 0; d < ND; ++d) if (d * NMF / ND == u) OZ2_DMA(KTI, d)                \
            __builtin_amdgcn_sched_barrier(0);                                                                   \
        }                                                                                                        \
    }
    // top of a stage: its successor landed (own DMAs, then everyone's) and the
    // fragments read during the previous stage are in registers
#define OZ2_WAIT()                                                      \
    {                                                                   \
        __builtin_amdgcn_sched_barrier(0);                              \
        __builtin_amdgcn_s_waitcnt(vm_wait(VM_AHEAD));                  \
        __builtin_amdgcn_s_waitcnt(kLgkmWait0);                         \
        __builtin_amdgcn_s_barrier();                                   \
        __builtin_amdgcn_sched_barrier(0);                              \
    }

    // prologue: stages 0 .. NS-2 in flight, stage 0 landed and read
#pragma unroll
    for (int s0 = 0; s0 < NS - 1; ++s0)
#pragma unroll
        for (int d = 0; d < ND; ++d) OZ2_DMA(s0, d)
    __builtin_amdgcn_s_waitcnt(vm_wait((NS - 2) * ND));
    __builtin_amdgcn_s_barrier();
#pragma unroll
    for (int r = 0; r < NR; ++r) OZ2_READ(fa0, fb0, 0, r)

    for (int kt = 0; kt < nkt; kt += 2) {
        OZ2_WAIT()
        OZ2_HALF(fa0, fb0, fa1, fb1, kt + 1, kt + NS - 1)
        OZ2_WAIT()
        OZ2_HALF(fa1, fb1, fa0, fb0, kt + 2, kt + NS)  // the last pass reads a stage that is not used
    }
#undef OZ2_DMA
#undef OZ2_READ
#undef OZ2_HALF
#undef OZ2_WAIT
    __builtin_amdgcn_s_waitcnt(vm_wait(0));  // no DMA may still target LDS
    after_loop();
    if constexpr (I32) {
        int32_epilogue<FM, FN>(p, acc, plane, static_cast<int64_t>(tm) * BM + wm * WTM,
                               static_cast<int64_t>(tn) * BN + wn * WTN, lane);
        return;
    }
    __syncthreads();                          // every wave is done with the stage buffers
    residue_epilogue<FM, FN>(p, acc, plane, static_cast<int64_t>(tm) * BM + wm * WTM,
                             static_cast<int64_t>(tn) * BN + wn * WTN, sbase + wave * (WTM * WTN), lane);
}

template <int BM, int BN, int MINB, int NS>
__global__ void __launch_bounds__(256, MINB) igemm_residue_kernel(GemmArgs p) {
    __shared__ v4i smem[TileShape<BM, BN, NS>::SMEM / 16];
    uint32_t plane, tm, tn;
    const uint32_t tiles = static_cast<uint32_t>(p.tiles_m) * static_cast<uint32_t>(p.tiles_n);
    tile_coords(p, plane_xcd_remap(blockIdx.x, tiles, gridDim.x), plane, tm, tn);
    gemm_tile<BM, BN, NS>(p, plane, tm, tn, smem);
}

template <int BM, int BN, int MINB, int NS>
__global__ void __launch_bounds__(256, MINB) igemm_int32_kernel(GemmArgs p) {
    __shared__ v4i smem[TileShape<BM, BN, NS>::SMEM / 16];
    uint32_t plane, tm, tn;
    const uint32_t tiles = static_cast<uint32_t>(p.tiles_m) * static_cast<uint32_t>(p.tiles_n);
    tile_coords(p, plane_xcd_remap(blockIdx.x, tiles, gridDim.x), plane, tm, tn);
    gemm_tile<BM, BN, NS, true>(p, plane, tm, tn, smem);
}

// Grouped forms over many problems; owner[w] is the item (problem, k-chunk) of
// work index w and start[] its first index, items listed largest first.
// Persistent: a fixed grid takes work indices from counters until all are done.  The
// indices are split over the 8 XCDs in contiguous ranges, as in the one-tile-per-block
// form (workgroup b runs on XCD b % 8), so the tiles that share A and B bands meet in
// one L2; each range has a counter of its own (counter[0..7]), and a workgroup whose
// range is done takes work from the next ranges in turn.  Every counter only grows, so
// each workgroup leaves after at most 8 empty draws.  Every drawn index is examined at the
// top of the loop: one below its range's end is a tile, processed once (the counters hand
// out each index once), one past it moves the workgroup to the next range.
template <int BM, int BN, int MINB, int NS>
__global__ void __launch_bounds__(256, MINB) igemm_grouped_kernel(const GemmArgs* __restrict__ items,
                                                                  const uint32_t* __restrict__ start,
                                                                  const uint32_t* __restrict__ owner,
                                                                  uint32_t total, uint32_t* __restrict__ counter) {
    // the work index is broadcast through one more 16-byte slot of the tile buffer: with a
    // second __shared__ object the compiler cannot tell which LDS the tile's DMA writes and
    // waits for all of it (vmcnt(0)) before every fragment read of the k-loop
    __shared__ v4i smem[TileShape<BM, BN, NS>::SMEM / 16 + 1];
    uint32_t& next = *reinterpret_cast<uint32_t*>(smem + TileShape<BM, BN, NS>::SMEM / 16);
    const uint32_t x0 = blockIdx.x & 7;
    uint32_t tried = 0;  // ranges found empty, from x0 on
    auto range = [&](uint32_t x, uint32_t& lo, uint32_t& hi) {
        lo = static_cast<uint32_t>(static_cast<uint64_t>(total) * x / 8);
        hi = static_cast<uint32_t>(static_cast<uint64_t>(total) * (x + 1) / 8);
    };
    // the draw for a tile is issued one tile ahead, when the previous tile starts, and its
    // result is parked in LDS once that tile's k-loop is done: the draw's return is then waited
    // for before the previous tile's residue stores are issued, not after them (a wave's
    // vector memory counter covers stores as well as loads, so a draw issued after the stores
    // would wait for their write acknowledgements)
    uint32_t x = x0, lo, hi;
    range(x, lo, hi);
    if (threadIdx.x == 0) next = lo + atomicAdd(&counter[x], 1u);
    __syncthreads();
    uint32_t w = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();
    for (;;) {
        if (w >= hi) {  // this range is done: draw from the next one
            if (++tried == 8) break;
            x = (x0 + tried) & 7;
            range(x, lo, hi);
            if (threadIdx.x == 0) next = lo + atomicAdd(&counter[x], 1u);
            __syncthreads();
            w = __builtin_amdgcn_readfirstlane(next);
            __syncthreads();
            continue;
        }
        uint32_t ticket = 0;
        if (threadIdx.x == 0) ticket = atomicAdd(&counter[x], 1u);  // the tile after this one
        const uint32_t it = owner[w];
        const GemmArgs& p = items[it];
        uint32_t plane, tm, tn;
        tile_coords(p, w - start[it], plane, tm, tn);
        gemm_tile<BM, BN, NS>(p, plane, tm, tn, smem, [&] {
            if (threadIdx.x == 0) next = lo + ticket;
        });
        __syncthreads();  // the epilogue is done with LDS; next is visible
        w = __builtin_amdgcn_readfirstlane(next);
        __syncthreads();
    }
}

// One workgroup per tile, work indices spread over the XCDs in contiguous ranges.
template <int BM, int BN, int MINB, int NS>
__global__ void __launch_bounds__(256, MINB) igemm_grouped_blocks_kernel(const GemmArgs* __restrict__ items,
                                                                         const uint32_t* __restrict__ start,
                                                                         const uint32_t* __restrict__ owner) {
    __shared__ v4i smem[TileShape<BM, BN, NS>::SMEM / 16];
    const uint32_t w = xcd_remap(blockIdx.x, gridDim.x);
    const uint32_t it = owner[w];
    const GemmArgs& p = items[it];
    uint32_t plane, tm, tn;
    tile_coords(p, w - start[it], plane, tm, tn);
    gemm_tile<BM, BN, NS>(p, plane, tm, tn, smem);
}

__global__ void reset_counters(uint32_t* c, int n) {
    if (static_cast<int>(threadIdx.x) < n) c[threadIdx.x] = 0;
}

}  // namespace

void prep_gemm(std::vector<GroupedLaunch>& out, const std::vector<GemmJob>& jobs) {
    GroupedLaunch by[4];
    for (const GemmJob& j : jobs) {
        const uint32_t nblk = static_cast<uint32_t>(j.args.tiles_m) * j.args.tiles_n * j.args.nplanes;
        if (nblk == 0 || j.args.klen <= 0) continue;
        const int t = static_cast<int>(j.tile);
        by[t].kind = t;
        by[t].owner.insert(by[t].owner.end(), nblk, by[t].n);
        by[t].add(j.args, nblk);
    }
    for (GroupedLaunch& g : by)
        if (g.n) out.push_back(std::move(g));
}

// Persistent forms: 128 x 128 and 64 x 64 keep their workgroups per CU; 256 x 128 runs one
// workgroup per CU there (with two, its loop state on top of the 252 registers the tile
// needs spills 23 dwords; with one it has the accumulation registers to itself: 194 + 128,
// no spill).  The loop around a 256 x 256 tile spills with one (256 accumulators per
// lane), so those groups always take one workgroup per tile.
static bool persistent_form(Tile t, bool persistent) { return persistent && t != Tile::T256; }
static int persistent_blocks_per_cu(Tile t) { return t == Tile::T256x128 ? 1 : tile_blocks_per_cu(t); }

int gemm_grouped_launches(const GroupedLaunch& g, bool persistent) {
    if (g.n == 0) return 0;
    return persistent_form(static_cast<Tile>(g.kind), persistent) ? 2 : 1;  // counter reset + GEMM
}

void launch_gemm_grouped(hipStream_t st, const GroupedLaunch& g, uint32_t* counter, int cus, bool persistent) {
    if (g.n == 0) return;
    const Tile t = static_cast<Tile>(g.kind);
    const GemmArgs* items = static_cast<const GemmArgs*>(g.d_items);
    if (persistent_form(t, persistent)) {
        hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, st, counter, kGemmCounters);
        const uint32_t grid = std::min<uint32_t>(g.blocks, static_cast<uint32_t>(cus * persistent_blocks_per_cu(t)));
#define OZ2_LAUNCH(BM, BN, MINB, NS)                                                                             \
    hipLaunchKernelGGL((igemm_grouped_kernel<BM, BN, MINB, NS>), dim3(grid), dim3(256), 0, st, items, g.d_start, \
                       g.d_owner, g.blocks, counter)
        if (t == Tile::T128) OZ2_LAUNCH(128, 128, 2, 4);
        else if (t == Tile::T256x128) OZ2_LAUNCH(256, 128, 1, 4);
        else OZ2_LAUNCH(64, 64, 4, 4);
#undef OZ2_LAUNCH
        return;
    }
#define OZ2_LAUNCH(BM, BN, MINB, NS)                                                                             \
    hipLaunchKernelGGL((igemm_grouped_blocks_kernel<BM, BN, MINB, NS>), dim3(g.blocks), dim3(256), 0, st, items, \
                       g.d_start, g.d_owner)
    switch (t) {
        case Tile::T256: OZ2_LAUNCH(256, 256, 1, 4); break;
        case Tile::T128: OZ2_LAUNCH(128, 128, 2, 4); break;
        case Tile::T64: OZ2_LAUNCH(64, 64, 4, 4); break;
        case Tile::T256x128: OZ2_LAUNCH(256, 128, 2, 3); break;
    }
#undef OZ2_LAUNCH
}

void igemm_int32(hipStream_t st, Tile tile, const GemmArgs& a) {
    const uint32_t nblk = static_cast<uint32_t>(a.tiles_m) * a.tiles_n * a.nplanes;
    if (nblk == 0 || a.klen <= 0) return;
#define OZ2_LAUNCH(BM, BN, MINB, NS) \
    hipLaunchKernelGGL((igemm_int32_kernel<BM, BN, MINB, NS>), dim3(nblk), dim3(256), 0, st, a)
    switch (tile) {
        case Tile::T256: OZ2_LAUNCH(256, 256, 1, 4); break;
        case Tile::T128: OZ2_LAUNCH(128, 128, 2, 4); break;
        case Tile::T64: OZ2_LAUNCH(64, 64, 4, 4); break;
        case Tile::T256x128: OZ2_LAUNCH(256, 128, 2, 3); break;
    }
#undef OZ2_LAUNCH
}

void igemm_residue(hipStream_t st, Tile tile, const GemmArgs& a) {
    const uint32_t nblk = static_cast<uint32_t>(a.tiles_m) * a.tiles_n * a.nplanes;
    if (nblk == 0 || a.klen <= 0) return;
#define OZ2_LAUNCH(BM, BN, MINB, NS) \
    hipLaunchKernelGGL((igemm_residue_kernel<BM, BN, MINB, NS>), dim3(nblk), dim3(256), 0, st, a)
    switch (tile) {
        case Tile::T256: OZ2_LAUNCH(256, 256, 1, 4); break;
        case Tile::T128: OZ2_LAUNCH(128, 128, 2, 4); break;
        case Tile::T64: OZ2_LAUNCH(64, 64, 4, 4); break;
        case Tile::T256x128: OZ2_LAUNCH(256, 128, 2, 3); break;
    }
#undef OZ2_LAUNCH
}

}  // namespace k
}  // namespace oz2
