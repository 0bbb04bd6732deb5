// batch.cpp - grouped launch sequence for many DGEMMs (see batch.hpp).
#include "batch.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

namespace oz2 {

namespace {
inline int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("oz2: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

BatchPlan::BatchPlan(Engine& eng, std::vector<GemmDesc> descs, const std::vector<int>& s,
                     const std::vector<int>& tiles, bool use_graph, BatchOptions opt)
    : eng_(eng), d_(std::move(descs)), opt_(opt) {
    const size_t np = d_.size();
    if (s.size() != np) throw std::invalid_argument("oz2: batch needs one number of moduli per problem");
    if (!tiles.empty() && tiles.size() != np) throw std::invalid_argument("oz2: batch needs one tile per problem");
    k::Tile unused;
    for (int t : tiles)
        if (t > 0 && !k::tile_from_code(t, unused))
            throw std::invalid_argument("oz2: batch tile must be 64, 128, 256 or 256128 (0: automatic)");
    s_ = s;
    tile_codes_.assign(np, 0);
    // validation (hpp:745-748, per problem)
    for (size_t i = 0; i < np; ++i) {
        const GemmDesc& g = d_[i];
        const std::string at = " (batch entry " + std::to_string(i) + ")";
        if (g.m < 0 || g.n < 0 || g.k < 0) throw std::invalid_argument("oz2: negative dimension" + at);
        if (g.lda < std::max<int64_t>(1, g.ta == Op::N ? g.m : g.k) ||
            g.ldb < std::max<int64_t>(1, g.tb == Op::N ? g.k : g.n) || g.ldc < std::max<int64_t>(1, g.m))
            throw std::invalid_argument("oz2: bad leading dimension" + at);
        if (s_[i] < kMinModuli || s_[i] > kMaxModuli)
            throw std::invalid_argument("oz2: num_moduli must be in [2, 24]" + at);
    }

    // per-problem layout inside one workspace: D, E, partial maxima, planes, residues
    struct P {
        Plan plan;
        k::Tile tile;
        int64_t kp, kch, mpad, npad, planeA, planeB, planeR;
        size_t oD, oE, oPA, oPB, oA, oB, oR;
    };
    std::vector<P> pp(np);
    std::vector<size_t> work;  // indices of problems that need the full path
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off = align_up(off + bytes, 256);
        return o;
    };
    for (size_t i = 0; i < np; ++i) {
        const GemmDesc& g = d_[i];
        if (g.m == 0 || g.n == 0) continue;
        if (g.k == 0 || g.alpha == 0.0) {
            special_.push_back(i);
            continue;
        }
        P& q = pp[i];
        q.plan = make_plan(s_[i], g.k);  // throws invalid_argument when s is too small for k (hpp:186)
        const int planes = s_[i];
        q.kp = round_up(g.k, kKStep);
        q.kch = k_chunks(q.kp);
        if (tiles.empty() || tiles[i] <= 0) q.tile = eng_.pick_tile(g.m, g.n, k_chunk_len(q.kp), planes);
        else k::tile_from_code(tiles[i], q.tile);
        tile_codes_[i] = k::tile_code(q.tile);
        q.mpad = round_up(g.m, k::tile_m(q.tile));
        q.npad = round_up(g.n, k::tile_n(q.tile));
        q.planeA = q.mpad * q.kp;
        q.planeB = q.npad * q.kp;
        q.planeR = q.npad * q.mpad;
        q.oD = carve(static_cast<size_t>(g.m) * 4);
        q.oE = carve(static_cast<size_t>(g.n) * 4);
        q.oPA = carve(k::scale_partial_elems(g.m, g.k, g.ta != Op::N) * 8);
        q.oPB = carve(k::scale_partial_elems(g.n, g.k, g.tb == Op::N) * 8);
        q.oA = carve(static_cast<size_t>(planes) * q.planeA);
        q.oB = carve(static_cast<size_t>(planes) * q.planeB);
        q.oR = carve(static_cast<size_t>(q.kch) * planes * q.planeR);
        work.push_back(i);
    }
    const size_t data_bytes = off;

    // grouped launches; GEMM work listed largest first so the persistent grid's tail is short
    std::vector<k::ScaleJob> sj;
    std::vector<k::SplitArgs> pj;
    std::vector<k::GemmJob> gj;
    std::vector<k::CrtArgs> cj;
    std::vector<size_t> by_cost = work;
    if (opt_.largest_first)
        std::stable_sort(by_cost.begin(), by_cost.end(), [&](size_t a, size_t b) {
            auto cost = [&](size_t i) { return static_cast<double>(pp[i].mpad) * pp[i].npad * pp[i].kp * s_[i]; };
            return cost(a) > cost(b);
        });
    const TabSet* tabs = eng_.device_tables();
    // the jobs hold device pointers into the workspace, whose size depends on the
    // job tables: build once against a null base to size them, then for real
    auto build = [&](std::byte* base) {
        sj.clear();
        pj.clear();
        gj.clear();
        cj.clear();
        for (size_t i : work) {
            const GemmDesc& g = d_[i];
            const P& q = pp[i];
            const int tk = q.plan.tk;
            int32_t* D = reinterpret_cast<int32_t*>(base + q.oD);
            int32_t* E = reinterpret_cast<int32_t*>(base + q.oE);
            sj.push_back({g.A, g.lda, g.m, g.k, g.ta != Op::N, false, tk, reinterpret_cast<uint64_t*>(base + q.oPA), D});
            sj.push_back({g.B, g.ldb, g.n, g.k, g.tb == Op::N, false, tk, reinterpret_cast<uint64_t*>(base + q.oPB), E});
            for (int which = 0; which < 2; ++which) {
                k::SplitArgs a{};
                a.kcontig = which == 0 ? g.ta != Op::N : g.tb == Op::N;
                a.X = which == 0 ? g.A : g.B;
                a.ldx = which == 0 ? g.lda : g.ldb;
                a.rows = which == 0 ? g.m : g.n;
                a.rows_pad = which == 0 ? q.mpad : q.npad;
                a.k = g.k;
                a.kp = q.kp;
                a.ex = which == 0 ? D : E;
                a.P = reinterpret_cast<int8_t*>(base + (which == 0 ? q.oA : q.oB));
                a.plane = which == 0 ? q.planeA : q.planeB;
                a.group = static_cast<int64_t>(s_[i]) * a.plane;
                a.s = s_[i];
                a.cplx = false;
                a.conj = false;
                a.form = limb_form(tk);
                a.tab = tabs + s_[i];
                pj.push_back(a);
            }
            k::CrtArgs c{};
            c.R = reinterpret_cast<const uint8_t*>(base + q.oR);
            c.plane_r = q.planeR;
            c.chunk_r = static_cast<int64_t>(s_[i]) * q.planeR;
            c.kch = static_cast<int>(q.kch);
            c.ldr = q.mpad;
            c.mb = g.m;
            c.nb = g.n;
            c.D = D;
            c.E = E;
            c.C = g.C;
            c.ldc = g.ldc;
            c.ar = g.alpha;
            c.ai = 0.0;
            c.br = g.beta;
            c.bi = 0.0;
            c.cplx = false;
            c.s = s_[i];
            c.L16 = 2 * q.plan.NL;
            c.tab = tabs + s_[i];
            cj.push_back(c);
        }
        for (size_t i : by_cost) {
            const P& q = pp[i];
            for (int64_t c = 0; c < q.kch; ++c) {
                k::GemmArgs a{};
                a.A = reinterpret_cast<const int8_t*>(base + q.oA);
                a.B = reinterpret_cast<const int8_t*>(base + q.oB);
                a.kp = q.kp;
                a.plane_a = q.planeA;
                a.plane_b = q.planeB;
                a.plane_r = q.planeR;
                a.ldr = q.mpad;
                a.k0 = c * k_chunk_len(q.kp);
                a.klen = std::min(k_chunk_len(q.kp), q.kp - a.k0);
                a.R = reinterpret_cast<uint8_t*>(base + q.oR) + c * s_[i] * q.planeR;
                a.tiles_m = static_cast<int>(q.mpad / k::tile_m(q.tile));
                a.tiles_n = static_cast<int>(q.npad / k::tile_n(q.tile));
                a.nplanes = s_[i];
                a.s = s_[i];
                a.tab = tabs + s_[i];
                gj.push_back({q.tile, a});
            }
        }
        scale_.clear();
        split_.clear();
        gemm_.clear();
        crt_.clear();
        k::prep_scale(scale_, sj);
        k::prep_split(split_, pj);
        k::prep_gemm(gemm_, gj);
        k::prep_crt(crt_, cj);
    };
    build(nullptr);
    size_t tables = 0;
    for (auto* v : {&scale_, &split_, &gemm_, &crt_})
        for (const k::GroupedLaunch& g : *v)
            tables += align_up(g.items.size(), 256) + align_up(g.start.size() * 4, 256) +
                      align_up(g.owner.size() * 4, 256);
    const size_t ncounters = gemm_.size() * k::kGemmCounters;
    ws_size_ = data_bytes + tables + align_up(std::max<size_t>(1, ncounters) * 4, 256) + 4096;
    if (ws_size_ > eng_.budget())
        throw std::runtime_error("oz2: not enough device memory for the batch workspace (" +
                                 std::to_string(ws_size_ >> 20) + " MiB)");
    void* p = nullptr;
    if (hipMalloc(&p, ws_size_) != hipSuccess || !p) {
        (void)hipGetLastError();
        throw std::runtime_error("oz2: batch workspace allocation failed");
    }
    ws_ = static_cast<std::byte*>(p);
    build(ws_);
    // upload the tables (once; the captured graph only reads them)
    size_t toff = data_bytes;
    std::vector<std::byte> host(tables);
    for (auto* v : {&scale_, &split_, &gemm_, &crt_})
        for (k::GroupedLaunch& g : *v) {
            std::memcpy(host.data() + (toff - data_bytes), g.items.data(), g.items.size());
            g.d_items = ws_ + toff;
            toff += align_up(g.items.size(), 256);
            std::memcpy(host.data() + (toff - data_bytes), g.start.data(), g.start.size() * 4);
            g.d_start = reinterpret_cast<const uint32_t*>(ws_ + toff);
            toff += align_up(g.start.size() * 4, 256);
            if (!g.owner.empty()) {
                std::memcpy(host.data() + (toff - data_bytes), g.owner.data(), g.owner.size() * 4);
                g.d_owner = reinterpret_cast<const uint32_t*>(ws_ + toff);
                toff += align_up(g.owner.size() * 4, 256);
            }
        }
    counters_ = reinterpret_cast<uint32_t*>(ws_ + toff);
    if (tables) hip_check(hipMemcpy(ws_ + data_bytes, host.data(), tables, hipMemcpyHostToDevice), "batch table upload");

    if (use_graph) {
        hipStream_t cs = nullptr;
        hip_check(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "stream create");
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            enqueue(cs);
            e = hipStreamEndCapture(cs, &graph_);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
        (void)hipStreamDestroy(cs);
        if (e != hipSuccess) {
            if (graph_) (void)hipGraphDestroy(graph_);
            graph_ = nullptr;
            exec_ = nullptr;
            (void)hipFree(ws_);
            ws_ = nullptr;
            hip_check(e, "batch graph capture");
        }
    }
}

BatchPlan::~BatchPlan() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (ws_) {
        (void)hipDeviceSynchronize();  // the plan may have been run on any stream
        (void)hipFree(ws_);
    }
}

int BatchPlan::launches() const {
    int n = static_cast<int>(special_.size());
    for (auto* v : {&scale_, &split_, &crt_}) n += static_cast<int>(v->size());
    for (const k::GroupedLaunch& g : gemm_) n += k::gemm_grouped_launches(g, opt_.persistent);
    return n;
}

void BatchPlan::enqueue(hipStream_t st) {
    for (size_t i : special_) {
        const GemmDesc& g = d_[i];
        k::scale_c(st, g.C, g.ldc, g.m, g.n, g.beta, 0.0, false);
    }
    for (const k::GroupedLaunch& g : scale_) k::launch_scale_split_grouped(st, g);
    for (const k::GroupedLaunch& g : split_) k::launch_scale_split_grouped(st, g);
    for (size_t i = 0; i < gemm_.size(); ++i)
        k::launch_gemm_grouped(st, gemm_[i], counters_ + k::kGemmCounters * i, eng_.compute_units(),
                               opt_.persistent);
    for (const k::GroupedLaunch& g : crt_) k::launch_crt_grouped(st, g);
}

void BatchPlan::run(hipStream_t st) {
    if (exec_) hip_check(hipGraphLaunch(exec_, st), "batch graph launch");
    else enqueue(st);
    hip_check(hipGetLastError(), "batch launch");
}

void BatchPlan::run_gemm_only(hipStream_t st) {
    for (size_t i = 0; i < gemm_.size(); ++i)
        k::launch_gemm_grouped(st, gemm_[i], counters_ + k::kGemmCounters * i, eng_.compute_units(),
                               opt_.persistent);
    hip_check(hipGetLastError(), "batch gemm launch");
}

namespace {
// INT8 GEMM rate (TOPS) of a macro tile at k-loop length kp, measured on MI355X with s = 16
// random planes of 4096 x 4096, k = 256 .. 8192: piecewise linear in log2(kp)
double tile_rate(k::Tile t, int64_t kp) {
    static const double lk[6] = {8, 9, 10, 11, 12, 13};  // log2 k
    static const double rate[4][6] = {
        {780, 1198, 1456, 1751, 2121, 2396},  // 256 x 256
        {688, 1202, 1646, 1866, 1976, 1933},  // 128 x 128
        {558, 817, 1082, 1269, 1351, 1179},   // 64 x 64
        {955, 1376, 1770, 2005, 2051, 2089},  // 256 x 128
    };
    const double* r = rate[static_cast<int>(t)];
    const double x = std::log2(static_cast<double>(std::max<int64_t>(kp, 128)));
    if (x <= lk[0]) return r[0] * std::pow(2.0, x - lk[0]);  // shorter k: epilogue-bound, rate ~ k
    if (x >= lk[5]) return r[5];
    int i = 0;
    while (x > lk[i + 1]) ++i;
    return r[i] + (r[i + 1] - r[i]) * (x - lk[i]);
}
}  // namespace

TuneResult autotune_batch(Engine& eng, const std::vector<GemmDesc>& d, const std::vector<int>& s, int reps) {
    TuneResult best;
    best.ms = 1e300;
    hipStream_t st = eng.stream();
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event create");
    hip_check(hipEventCreate(&e1), "event create");
    auto timed = [&](const std::vector<int>& tiles, BatchOptions opt) {
        BatchPlan plan(eng, d, s, tiles, false, opt);
        plan.run(st);  // warm (and fills the planes once)
        float bestms = 1e30f;
        for (int r = 0; r < std::max(1, reps); ++r) {
            hip_check(hipEventRecord(e0, st), "event record");
            plan.run(st);
            hip_check(hipEventRecord(e1, st), "event record");
            hip_check(hipEventSynchronize(e1), "event synchronize");
            float ms = 0;
            hip_check(hipEventElapsedTime(&ms, e0, e1), "event time");
            bestms = std::min(bestms, ms);
        }
        return static_cast<double>(bestms);
    };
    // square tile by min(m, n); problems with k loops up to kr (and room for the rows)
    // take 256 x 128 tiles instead of the square ones
    auto policy = [&](int64_t a, int64_t b, int64_t kr) {
        std::vector<int> t(d.size());
        for (size_t i = 0; i < d.size(); ++i) {
            const int64_t x = std::min(d[i].m, d[i].n);
            t[i] = x >= a ? 256 : x >= b ? 128 : 64;
            if (t[i] != 64 && round_up(d[i].k, kKStep) <= kr && d[i].m > 128 && d[i].n > 64) t[i] = 256128;
        }
        return t;
    };
    const int64_t inf = int64_t{1} << 40;
    struct Cand {
        std::string name;
        std::vector<int> tiles;
    };
    std::vector<Cand> cands{{"heuristic", {}}};
    // per problem, the tile with the least padded work / measured rate at its k: in a grouped
    // launch the union of the problems fills the GPU, so each problem only needs its fastest tile
    {
        std::vector<int> t(d.size());
        for (size_t i = 0; i < d.size(); ++i) {
            const int64_t kp = k_chunk_len(round_up(std::max<int64_t>(d[i].k, 1), kKStep));
            double best_t = 0;
            for (k::Tile c : {k::Tile::T256, k::Tile::T256x128, k::Tile::T128, k::Tile::T64}) {
                const double work = static_cast<double>(round_up(d[i].m, k::tile_m(c))) *
                                    static_cast<double>(round_up(d[i].n, k::tile_n(c))) * static_cast<double>(kp);
                const double est = work / tile_rate(c, kp);
                if (best_t == 0 || est < best_t) {
                    best_t = est;
                    t[i] = k::tile_code(c);
                }
            }
        }
        cands.push_back({"per-problem tile by padded work / rate", t});
    }
    for (int64_t kr : {int64_t{0}, int64_t{1024}})
        for (auto [a, b] : {std::pair<int64_t, int64_t>{192, 64}, {256, 96}, {384, 128}, {512, 160}, {768, 256},
                            {inf, 128}, {inf, 64}})
            cands.push_back({"tile256>=" + (a >= inf ? std::string("inf") : std::to_string(a)) + ",tile128>=" +
                                 std::to_string(b) + (kr ? ",tile256x128 for k<=" + std::to_string(kr) : ""),
                             policy(a, b, kr)});
    try {
        for (const Cand& c : cands)
            for (bool pers : {false, true}) {
                BatchOptions o;
                o.persistent = pers;
                const double ms = timed(c.tiles, o);
                best.trials.push_back({c.name + (pers ? ",persistent" : ",one-tile-per-block"), ms});
                if (ms < best.ms) {
                    best.ms = ms;
                    best.opt = o;
                    best.tiles = c.tiles;
                }
            }
        BatchOptions o = best.opt;
        o.largest_first = false;
        const double ms = timed(best.tiles, o);
        best.trials.push_back({"winner in problem order", ms});
        if (ms < best.ms) {
            best.ms = ms;
            best.opt = o;
        }
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (best.tiles.empty()) {  // the heuristic won: report its choices
        BatchPlan plan(eng, d, s, {}, false, best.opt);
        best.tiles = plan.tiles();
    }
    return best;
}

}  // namespace oz2
