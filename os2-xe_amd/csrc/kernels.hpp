// kernels.hpp - host-side launch interface of the gfx950 kernels.
//
// Data layout in HBM (all owned by the engine's workspace):
//   planes   int8  [group][t] then blocked: 1 KB per 16 rows x 64 k of op(A) (or op(B)^T),
//                  blocks k-major within a 16-row band (plane_offset); rows padded with zeros to
//                  the GEMM tile, k to kp
//   residues uint8 [chunk][group][t][j][ldr]  GEMM results reduced mod m_t, i contiguous
//   D, E     int32 exponents per row of op(A) / column of op(B)
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include <vector>

#include "grouped.hpp"
#include "plan.hpp"
#include "planner.hpp"

namespace oz2 {
namespace k {

// Byte offset of element (row r, k p) inside one operand plane with kp columns:
// 16-row bands of kp/64 consecutive 1 KB blocks; inside a block, 64 bytes per row
// as four 16-byte chunks.  Byte b of chunk c holds k = block_k(c, b) of the block:
// chunk c takes the pair (2c, 2c+1) of each group of 8 consecutive k, so the four
// threads that fill the chunks of one row read 64 contiguous bytes of a
// k-contiguous row per load.  The GEMM's dot products do not depend on the k order
// inside a block, which both operands share.  The chunks of row r are permuted by
// XOR with (r >> 2) & 3: the GEMM's DMA copies blocks verbatim to LDS, where that
// permutation makes the MFMA operand reads free of bank conflicts.
__host__ __device__ constexpr int block_k(int c, int b) { return 8 * (b >> 1) + 2 * c + (b & 1); }
__host__ __device__ inline int64_t chunk_offset(int64_t r, int64_t blk, int c, int64_t kp) {
    return (((r >> 4) * (kp >> 6) + blk) << 10) + ((r & 15) << 6) + ((c ^ static_cast<int>((r >> 2) & 3)) << 4);
}
__host__ __device__ inline int64_t plane_offset(int64_t r, int64_t p, int64_t kp) {
    const int q = static_cast<int>(p & 63);
    return chunk_offset(r, p >> 6, (q >> 1) & 3, kp) + (((q >> 3) << 1) | (q & 1));
}

// ---- scaling -----------------------------------------------------------------
// Exponents of the `len` rows of op(X) over k columns.  kcontig: row r of op(X)
// is contiguous (X[p + r*ldx]); otherwise rows are strided (X[r + p*ldx]).
// Complex data (interleaved re, im) scales by the max over both parts.
// `partial` needs partial_elems(len, k) u64 entries when !kcontig.
size_t scale_partial_elems(int64_t len, int64_t k, bool kcontig);
void scale(hipStream_t st, const double* X, int64_t ldx, int64_t len, int64_t k, bool kcontig, bool cplx, int tk,
           uint64_t* partial, int32_t* ex);

// ---- split -------------------------------------------------------------------
struct SplitArgs {
    const double* X;      // op(X) rows start at row r0 (already offset)
    int64_t ldx;
    int64_t rows;         // valid rows of this block
    int64_t rows_pad;     // rows written (>= rows, multiple of 32)
    int64_t k;            // valid k
    int64_t kp;           // padded k (multiple of 128)
    const int32_t* ex;    // exponent per row (already offset)
    int8_t* P;            // planes
    int64_t plane;        // bytes per plane (rows_pad * kp)
    int64_t group;        // bytes per group of s planes (complex)
    int s;
    bool kcontig, cplx, conj;
    int form;             // limb_form(tk): octets of the offset value 8, 12 or 14
    const TabSet* tab;    // device tables of this s
};
void split(hipStream_t st, const SplitArgs& a);

// ---- INT8 GEMM with residue epilogue ------------------------------------------
struct GemmArgs {
    const int8_t* A;      // planes of op(A): nplanes x (mpad x kp, blocked by plane_offset)
    const int8_t* B;      // planes of op(B): nplanes x (npad x kp, blocked by plane_offset)
    uint8_t* R;           // residues [nplanes][npad_c][ldr]
    int64_t kp;           // row stride of planes (bytes)
    int64_t plane_a, plane_b, plane_r;
    int64_t ldr;          // column stride of residues (= mpad)
    int64_t k0, klen;     // k range (multiples of 128)
    int tiles_m, tiles_n;
    int nplanes;          // planes of this launch (s or 3s; 1 for per-plane launches)
    int plane0;           // global index of this launch's first plane: its modulus is (plane0 + plane) mod s
    int s;
    const TabSet* tab;    // device tables of this s
    int64_t kb_extra;     // op(B) reads k range k0 + kb_extra .. (0 for Scheme II)
    int32_t* R32;         // igemm_int32: sums [nplanes][npad_c][ldr] as int32 (R unused)
};
void igemm_residue(hipStream_t st, Tile tile, const GemmArgs& a);
// the same GEMM with the int32 sums stored as they are (|sums| < 2^31 is the caller's to keep):
// the digit products of Ozaki Scheme I
void igemm_int32(hipStream_t st, Tile tile, const GemmArgs& a);

// ---- CRT + alpha/beta epilogue -------------------------------------------------
struct CrtArgs {
    const uint8_t* R;     // residues of this tile
    int64_t plane_r;      // bytes per (plane) residue matrix
    int64_t chunk_r;      // bytes per k-chunk (nplanes * plane_r)
    int kch;
    int64_t ldr;
    int64_t mb, nb;       // tile extent
    const int32_t* D;     // per row (offset to tile)
    const int32_t* E;     // per column (offset to tile)
    double* C;            // offset to tile; complex: interleaved
    int64_t ldc;
    double ar, ai, br, bi;
    bool cplx;
    int s, L16;
    const TabSet* tab;    // device tables of this s
};
void crt(hipStream_t st, const CrtArgs& a);

// ---- Ozaki Scheme I (the accuracy benchmark's comparison baseline) --------------------
// L balanced base-128 digits of RNE(op(X)[r][p] 2^ex[r]) per element into one extended
// plane of rows_pad x (L kp) bytes (blocked as plane_offset with L kp columns): digit i
// (most significant first) at k-slot i, or L - 1 - i with rev
struct Oz1SliceArgs {
    const double* X;
    int64_t ldx, rows, rows_pad, k, kp;
    const int32_t* ex;
    int8_t* P;
    bool kcontig, rev;
};
void oz1_slices(hipStream_t st, const Oz1SliceArgs& a, int L);
// Horner step of anti-diagonal d over the int32 chunk sums Ct[c][j][ldt] (m x n entries)
struct Oz1HornerArgs {
    const int32_t* Ct;
    int64_t ldt, chunk;   // column stride and chunk stride of Ct (elements)
    int nchunks;
    double *Th, *Tl;      // running double-double T (m x n)
    bool first, last;
    int shift;            // 7 (L - 1)
    const int32_t *D, *E;
    double* C;
    int64_t m, n, ldc;
};
void oz1_horner(hipStream_t st, const Oz1HornerArgs& a);

// ---- C = beta * C for k == 0 or alpha == 0 --------------------------------------
void scale_c(hipStream_t st, double* C, int64_t ldc, int64_t m, int64_t n, double br, double bi, bool cplx);

// ---- grouped launches over many problems (batched GEMMs) -------------------------
// prep_* append launches (one per kernel variant the jobs need) whose blocks
// run exactly the single-problem kernels' blocks; launch order = vector order.
struct ScaleJob {
    const double* X;
    int64_t ldx, len, k;
    bool kcontig, cplx;
    int tk;
    uint64_t* partial;    // scale_partial_elems(len, k, kcontig) entries (strided rows only)
    int32_t* ex;
};
struct GemmJob {
    Tile tile;
    GemmArgs args;
};
void prep_scale(std::vector<GroupedLaunch>& out, const std::vector<ScaleJob>& jobs);
void prep_split(std::vector<GroupedLaunch>& out, const std::vector<SplitArgs>& jobs);
void prep_gemm(std::vector<GroupedLaunch>& out, const std::vector<GemmJob>& jobs);
void prep_crt(std::vector<GroupedLaunch>& out, const std::vector<CrtArgs>& jobs);
bool launch_scale_split_grouped(hipStream_t st, const GroupedLaunch& g);
// persistent (128 x 128 and 64 x 64 tiles at their blocks per CU, 256 x 128 at one per CU):
// min(blocks, cus x that) workgroups pull tiles through counter[0 .. kGemmCounters) (reset
// first; one per XCD range of the work); else, and always for 256 x 256 tiles, one workgroup
// per tile (counters unused)
inline constexpr int kGemmCounters = 8;
void launch_gemm_grouped(hipStream_t st, const GroupedLaunch& g, uint32_t* counter, int cus, bool persistent);
int gemm_grouped_launches(const GroupedLaunch& g, bool persistent);  // kernels launch_gemm_grouped enqueues
void launch_crt_grouped(hipStream_t st, const GroupedLaunch& g);

}  // namespace k
}  // namespace oz2
