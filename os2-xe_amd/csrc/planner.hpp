// planner.hpp - workspace layout and panel planner of one DGEMM/ZGEMM call.
//
// Host only (no HIP types), so the CPU test build links it as well.
//
// Counterpart of the reference's tiling planner (src/ozaki2_gpu.hpp:1072-1166):
// C is cut into pm x pn panels so that the planes and residues of one panel
// pair fit the workspace budget.  The choice here is this library's own: for
// every distinct row-panel count the largest column panel that fits is found by
// bisection on the exact workspace size, a few neighbouring column-panel counts
// are priced too, and the cheapest pair under a cost model of this library's
// kernels on MI355X wins (panel_cost).
#pragma once
#include <cstddef>
#include <cstdint>

namespace oz2 {

struct TileInfo {
    int64_t mb = 0, nb = 0;
    int64_t pm = 0, pn = 0;
    int64_t k_chunks = 0;
    size_t workspace_bytes = 0;
};

namespace k {
// Macro tiles of the INT8 GEMM (rows of op(A) x columns of op(B) per workgroup of
// 4 waves).  T256: one workgroup per CU, 4 LDS stages (long k).  T256x128: two
// workgroups per CU, 3 stages, so one workgroup's residue epilogue runs while the
// other's MFMAs do (short k).  T128, T64: small problems.
enum class Tile { T256 = 0, T128 = 1, T64 = 2, T256x128 = 3 };
constexpr int tile_m(Tile t) { return t == Tile::T256 || t == Tile::T256x128 ? 256 : t == Tile::T128 ? 128 : 64; }
constexpr int tile_n(Tile t) { return t == Tile::T256 ? 256 : t == Tile::T64 ? 64 : 128; }
constexpr int tile_code(Tile t) { return t == Tile::T256x128 ? 256128 : tile_m(t); }
// workgroups per CU (LDS: 128 / 64 / 32 / 72 KB per workgroup)
constexpr int tile_blocks_per_cu(Tile t) { return t == Tile::T256 ? 1 : t == Tile::T64 ? 4 : 2; }
inline bool tile_from_code(int code, Tile& t) {
    switch (code) {
        case 256: t = Tile::T256; return true;
        case 128: t = Tile::T128; return true;
        case 64: t = Tile::T64; return true;
        case 256128: t = Tile::T256x128; return true;
        default: return false;
    }
}
}  // namespace k

// What the planner needs to know of a call.
struct CallShape {
    int64_t m = 0, n = 0, k = 0;
    int s = 0;
    bool cplx = false;
    size_t partial_a = 0, partial_b = 0;  // bytes of the strided-row maxima of op(A) / op(B)
    int cus = 256;
    bool pad_planes = true;   // Tuning::pad_lda
    bool pow2 = true;         // Tuning::pow2_tiles
    bool per_plane = false;   // !Tuning::batched_gemm (one GEMM launch per plane)
    int forced_tile = -1;     // Tuning::gemm_tile
};
bool same_call(const CallShape& a, const CallShape& b);

// Workspace of one call with mb x nb panels: exponents, maxima, planes of one row
// panel of op(A) and one column panel of op(B), residues of one panel pair.
struct Layout {
    k::Tile tile = k::Tile::T256;
    int64_t TM = 256, TN = 256, kp = 0, klen = 0, mpad = 0, npad = 0, kch = 1;
    int g = 1, planes = 1;
    int64_t planeA = 0, planeB = 0, planeR = 0;
    size_t oD = 0, oE = 0, oPA = 0, oPB = 0, oA = 0, oB = 0, oR = 0, total = 0;
};

// GEMM macro tile for a panel of mb x nb with k loops of klen; forced_code from Tuning::gemm_tile
k::Tile pick_tile(int64_t mb, int64_t nb, int64_t klen, int planes, int forced_code);
Layout layout(const CallShape& c, int64_t mb, int64_t nb);
// modelled seconds of the whole call with mb x nb panels on MI355X
double panel_cost(const CallShape& c, int64_t mb, int64_t nb);
// cheapest panels whose workspace fits `limit`; std::runtime_error if none does
TileInfo plan_tiling(const CallShape& c, size_t limit);

}  // namespace oz2
