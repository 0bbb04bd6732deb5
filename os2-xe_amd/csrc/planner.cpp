// planner.cpp - workspace layout and panel planner (see planner.hpp).
#include "planner.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <string>

#include "plan.hpp"

namespace oz2 {

namespace {
inline int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
inline int64_t ceil_div(int64_t v, int64_t a) { return (v + a - 1) / a; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Plane strides that are multiples of 64 KiB put the same offset of every plane
// on the same HBM channels; the GEMM (one plane per XCD range) and the CRT (all
// planes of one entry) read those offsets at the same time.  Tuning::pad_lda
// staggers such strides by 4 KiB, the counterpart of the reference's ld pad
// for cache-set conflicts (ozaki2_gpu.hpp:669-672).
constexpr int64_t kPlanePadPeriod = 64 << 10;
constexpr int64_t kPlanePad = 4 << 10;
inline int64_t plane_stride(int64_t bytes, bool pad) {
    return pad && bytes % kPlanePadPeriod == 0 ? bytes + kPlanePad : bytes;
}

// Rates of this library's kernels on MI355X.  Only their ratios steer the choice.
// GEMM: a macro tile's k loop runs at `peak` (all CUs busy) and its prologue and
// residue epilogue cost as much as `k0` more k (fitted to 256 x 256 at 8192^3 2.5,
// 2048^3 2.1 and 4096 x 4096 x 256 0.77 Pop/s; 256 x 128 at 2.0 and 0.96).
struct TileRate {
    double peak;  // int8 op/s of the whole GPU
    double k0;    // k-equivalent fixed cost per tile
};
constexpr TileRate kTileRate[4] = {
    {2.60e15, 450.0},  // T256
    {2.07e15, 420.0},  // T128
    {1.00e15, 300.0},  // T64
    {2.12e15, 310.0},  // T256x128
};
// scale, split and CRT move their algorithmic bytes at 4.4-5.2 TB/s; one launch ~6 us
constexpr double kStreamRate = 4.7e12;
constexpr double kLaunchSec = 6e-6;

// distinct balanced panel extents of x: x itself, then round_up(ceil(x / p), 64)
// for p = 2, 3, ... down to 64 (each value once, largest first)
template <class F>
void for_each_panel(int64_t x, F&& f) {
    f(x);
    int64_t b = x;
    int64_t p = 2;
    while (b > 64) {
        const int64_t nb = round_up(ceil_div(x, p), 64);
        if (nb < b) {
            b = nb;
            f(b);
        }
        if (b <= 64) break;
        // the next panel count that gives a smaller extent: ceil(x / p') <= b - 64
        p = std::max(p + 1, ceil_div(x, b - 64));
    }
}

// GEMM seconds of one launch over r x c of the panel grid (tile t, k loops of klen)
double gemm_seconds(const CallShape& c, const Layout& L, int64_t rows, int64_t cols) {
    const TileRate& R = kTileRate[static_cast<int>(L.tile)];
    const double tiles = static_cast<double>(ceil_div(rows, L.TM)) * static_cast<double>(ceil_div(cols, L.TN)) *
                         L.planes;
    const double slots = static_cast<double>(c.cus) * k::tile_blocks_per_cu(L.tile);
    const double waves = std::ceil(tiles / slots);
    double t = 0;
    for (int64_t k0 = 0; k0 < L.kp; k0 += L.klen) {
        const double len = static_cast<double>(std::min(L.klen, L.kp - k0));
        // one wave of `slots` tiles takes 2 TM TN (len + k0) slots / peak
        t += waves * 2.0 * L.TM * L.TN * (len + R.k0) * slots / R.peak;
    }
    return t;
}
}  // namespace

bool same_call(const CallShape& a, const CallShape& b) {
    return a.m == b.m && a.n == b.n && a.k == b.k && a.s == b.s && a.cplx == b.cplx && a.partial_a == b.partial_a &&
           a.partial_b == b.partial_b && a.cus == b.cus && a.pad_planes == b.pad_planes && a.pow2 == b.pow2 &&
           a.per_plane == b.per_plane && a.forced_tile == b.forced_tile;
}

k::Tile pick_tile(int64_t mb, int64_t nb, int64_t klen, int planes, int forced_code) {
    k::Tile forced;
    if (k::tile_from_code(forced_code, forced)) return forced;
    // never pad a small tile up to a bigger macro tile
    const int64_t lim = std::min(mb, nb);
    if (lim <= 64) return k::Tile::T64;
    auto blocks = [&](int64_t tm, int64_t tn) { return ceil_div(mb, tm) * ceil_div(nb, tn) * planes; };
    // k loops of at most 1024: 256 x 128 tiles, two workgroups per CU (measured
    // ahead of both square tiles there, behind 256 x 256 for long k)
    if (klen <= 1024 && mb > 128 && nb > 64 && blocks(256, 128) >= 512) return k::Tile::T256x128;
    // largest square tile that still gives every CU two workgroups
    if (lim > 128 && blocks(256, 256) >= 512) return k::Tile::T256;
    if (blocks(128, 128) >= 512) return k::Tile::T128;
    return blocks(128, 128) >= 256 ? k::Tile::T128 : k::Tile::T64;
}

Layout layout(const CallShape& c, int64_t mb, int64_t nb) {
    Layout L;
    L.g = c.cplx ? 3 : 1;
    L.planes = L.g * c.s;
    L.kp = round_up(c.k, kKStep);
    L.klen = k_chunk_len(L.kp);
    L.kch = k_chunks(L.kp);
    L.tile = pick_tile(mb, nb, L.klen, L.planes, c.forced_tile);
    L.TM = k::tile_m(L.tile);
    L.TN = k::tile_n(L.tile);
    L.mpad = round_up(mb, L.TM);
    L.npad = round_up(nb, L.TN);
    L.planeA = plane_stride(L.mpad * L.kp, c.pad_planes);
    L.planeB = plane_stride(L.npad * L.kp, c.pad_planes);
    L.planeR = plane_stride(L.npad * L.mpad, c.pad_planes);
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    L.oD = carve(static_cast<size_t>(c.m) * 4);
    L.oE = carve(static_cast<size_t>(c.n) * 4);
    L.oPA = carve(c.partial_a);
    L.oPB = carve(c.partial_b);
    L.oA = carve(static_cast<size_t>(L.planes) * L.planeA);
    L.oB = carve(static_cast<size_t>(L.planes) * L.planeB);
    L.oR = carve(static_cast<size_t>(L.kch) * L.planes * L.planeR);
    L.total = off + 4096;
    return L;
}

double panel_cost(const CallShape& c, int64_t mb, int64_t nb) {
    const Layout L = layout(c, mb, nb);
    const int64_t pm = ceil_div(c.m, mb), pn = ceil_div(c.n, nb);
    const int64_t mlast = c.m - (pm - 1) * mb, nlast = c.n - (pn - 1) * nb;
    // GEMM launches: (pm - 1) x (pn - 1) full panel pairs, the ragged last row and column
    double gemm = 0;
    gemm += static_cast<double>((pm - 1) * (pn - 1)) * gemm_seconds(c, L, mb, nb);
    gemm += static_cast<double>(pn - 1) * gemm_seconds(c, L, mlast, nb);
    gemm += static_cast<double>(pm - 1) * gemm_seconds(c, L, mb, nlast);
    gemm += gemm_seconds(c, L, mlast, nlast);
    // bandwidth-bound bytes: each split reads its operand and writes the padded planes;
    // op(A) is split once when it is one row panel, else once per panel pair; the CRT
    // reads every residue byte and reads and writes C
    const double elem = c.cplx ? 16.0 : 8.0;
    const double mpad_all = static_cast<double>((pm - 1) * round_up(mb, L.TM) + round_up(mlast, L.TM));
    const double npad_all = static_cast<double>((pn - 1) * round_up(nb, L.TN) + round_up(nlast, L.TN));
    const double kd = static_cast<double>(c.k), kpd = static_cast<double>(L.kp);
    const double split_b = static_cast<double>(c.n) * kd * elem + L.planes * npad_all * kpd;
    const double split_a = static_cast<double>(pm > 1 ? pn : 1) * (static_cast<double>(c.m) * kd * elem + L.planes * mpad_all * kpd);
    const double crt = static_cast<double>(c.m) * static_cast<double>(c.n) * (L.planes * L.kch + 2.0 * elem);
    const double launches = static_cast<double>(pn + (pm > 1 ? pm * pn : 1) + pm * pn) +
                            static_cast<double>(pm * pn * L.kch) * (c.per_plane ? L.planes : 1);
    return gemm + (split_a + split_b + crt) / kStreamRate + launches * kLaunchSec;
}

TileInfo plan_tiling(const CallShape& c, size_t limit) {
    auto fits = [&](int64_t mb, int64_t nb) { return layout(c, mb, nb).total <= limit; };
    double best = std::numeric_limits<double>::infinity();
    TileInfo bt;
    auto consider = [&](int64_t mb, int64_t nb) {
        if (!fits(mb, nb)) return;
        const double t = panel_cost(c, mb, nb);
        // strictly cheaper only: among equal costs the first (fewest panels) stays
        if (t < best * (1.0 - 1e-12)) {
            best = t;
            bt = TileInfo{mb, nb, ceil_div(c.m, mb), ceil_div(c.n, nb), layout(c, mb, nb).kch,
                          layout(c, mb, nb).total};
        }
    };
    auto try_rows = [&](int64_t mb) {
        // largest column panel that fits: all of n, else bisection over multiples of 64
        int64_t nbmax = 0;
        if (fits(mb, c.n)) {
            nbmax = c.n;
        } else {
            int64_t lo = 0, hi = ceil_div(c.n, 64);  // fits(64 lo) (lo = 0: nothing yet), !fits(64 hi)
            while (hi - lo > 1) {
                const int64_t mid = (lo + hi) / 2;
                if (fits(mb, 64 * mid)) lo = mid;
                else hi = mid;
            }
            nbmax = 64 * lo;
        }
        if (nbmax == 0) return;
        // the fewest column panels that fit, and two more (less padding or better waves)
        const int64_t pn0 = ceil_div(c.n, nbmax);
        for (int64_t pn = pn0; pn <= pn0 + 2; ++pn) {
            const int64_t nb = pn == 1 ? c.n : round_up(ceil_div(c.n, pn), 64);
            if (nb > nbmax && nb != c.n) continue;
            if (ceil_div(c.n, nb) != pn) continue;  // this count is not reachable with balanced panels
            consider(mb, nb);
        }
        if (c.pow2)
            for (int64_t nb = 1024; nb < std::min(c.n, nbmax + 1); nb *= 2) consider(mb, nb);
    };
    for_each_panel(c.m, try_rows);
    if (c.pow2)
        for (int64_t mb = 1024; mb < c.m; mb *= 2) try_rows(mb);
    if (!std::isfinite(best))
        throw std::runtime_error("oz2: not enough device memory for workspace (limit " + std::to_string(limit >> 20) +
                                 " MiB)");
    return bt;
}

}  // namespace oz2
