// engine.cpp - host orchestrator (see engine.hpp).
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>

namespace oz2 {

namespace {
inline int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("oz2: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

Engine::Engine(hipStream_t stream) : stream_(stream) {
    std::vector<TabSet> host(kMaxModuli + 1);
    make_tabsets(host.data());
    void* p = nullptr;
    hip_check(hipMalloc(&p, host.size() * sizeof(TabSet)), "table allocation");
    dtabs_ = static_cast<TabSet*>(p);
    const hipError_t e = hipMemcpy(dtabs_, host.data(), host.size() * sizeof(TabSet), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(dtabs_);
        hip_check(e, "table upload");
    }
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        dev_ = dev;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0) cus_ = cus;
    }
    if (hipEventCreateWithFlags(&switch_ev_, hipEventDisableTiming) != hipSuccess) {
        (void)hipFree(dtabs_);
        throw std::runtime_error("oz2: event create failed");
    }
}

Engine::~Engine() {
    (void)hipStreamSynchronize(stream_);
    release_workspace();
    for (auto e : events_) (void)hipEventDestroy(e);
    if (switch_ev_) (void)hipEventDestroy(switch_ev_);
    if (dtabs_) (void)hipFree(dtabs_);
}

void Engine::set_stream(hipStream_t s) {
    if (s == stream_) return;
    // the workspace (and the selector's buffers) may still be in use by work on the old
    // stream: the new stream waits for all of it before anything else runs there
    hip_check(hipEventRecord(switch_ev_, stream_), "event record");
    hip_check(hipStreamWaitEvent(s, switch_ev_, 0), "stream wait");
    stream_ = s;
}

void Engine::release_workspace() {
    if (!ws_) return;
    (void)hipStreamSynchronize(stream_);
    (void)hipFree(ws_);
    ws_ = nullptr;
    ws_size_ = 0;
}

void Engine::reserve_workspace(size_t bytes) { ensure_workspace(bytes); }

void Engine::ensure_workspace(size_t bytes) {
    if (ws_size_ >= bytes) return;
    // the old workspace may still be read by queued work (hpp:1159-1166)
    hip_check(hipStreamSynchronize(stream_), "stream synchronize");
    if (ws_) (void)hipFree(ws_);
    ws_ = nullptr;
    ws_size_ = 0;
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess || !p) {
        (void)hipGetLastError();
        throw std::runtime_error("oz2: workspace allocation failed");
    }
    ws_ = static_cast<std::byte*>(p);
    ws_size_ = bytes;
}

const Plan& Engine::plan_for(int s, int64_t k) {
    if (!have_plan_ || plan_.s != s || plan_.k != k) {
        plan_ = make_plan(s, k);
        have_plan_ = true;
    }
    return plan_;
}

void Engine::record(Stage st) {
    if (!profiling_) return;
    const size_t idx = stages_.size();
    if (idx == events_.size()) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "event create");
        events_.push_back(e);
    }
    hip_check(hipEventRecord(events_[idx], stream_), "event record");
    stages_.emplace_back(st, idx);
}

StageTimes Engine::stage_times() {
    StageTimes t;
    if (!profiling_ || stages_.size() < 2) return t;
    (void)hipEventSynchronize(events_[stages_.back().second]);
    for (size_t i = 1; i < stages_.size(); ++i) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, events_[stages_[i - 1].second], events_[stages_[i].second]) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0;
        }
        switch (stages_[i].first) {
            case Stage::Begin: break;
            case Stage::Scale: t.scale_ms += ms; break;
            case Stage::Split: t.split_ms += ms; break;
            case Stage::Gemm: t.gemm_ms += ms; break;
            case Stage::Crt: t.crt_ms += ms; break;
        }
    }
    return t;
}

void Engine::profile_begin() {
    stages_.clear();
    calls_ = 0;
    accumulate_ = true;
}

StageTimes Engine::profile_end(int64_t* calls) {
    accumulate_ = false;
    if (calls) *calls = calls_;
    return stage_times();
}

size_t Engine::budget() const {
    if (ws_limit_) return ws_limit_;
    size_t free_b = 0, total = 0;
    if (hipMemGetInfo(&free_b, &total) != hipSuccess) {
        (void)hipGetLastError();
        return size_t{1} << 34;
    }
    free_b += ws_size_;
    const size_t reserve = (size_t{512} << 20) + total / 32;
    size_t lim = free_b > reserve ? free_b - reserve : 0;
    if (resident_hint_) lim = std::min(lim, total > reserve + resident_hint_ ? total - reserve - resident_hint_ : 0);
    return lim;
}

k::Tile Engine::pick_tile(int64_t mb, int64_t nb, int64_t klen, int planes) const {
    return oz2::pick_tile(mb, nb, klen, planes, tune_.gemm_tile);
}

CallShape Engine::call_shape(Op ta, Op tb, int64_t m, int64_t n, int64_t k, int s, bool cplx) const {
    CallShape c;
    c.m = m;
    c.n = n;
    c.k = k;
    c.s = s;
    c.cplx = cplx;
    c.partial_a = k::scale_partial_elems(m, k, ta != Op::N) * 8;
    c.partial_b = k::scale_partial_elems(n, k, tb == Op::N) * 8;
    c.cus = cus_;
    c.pad_planes = tune_.pad_lda;
    c.pow2 = tune_.pow2_tiles;
    c.per_plane = !tune_.batched_gemm;
    c.forced_tile = tune_.gemm_tile;
    return c;
}

TileInfo Engine::plan_for_call(const CallShape& c, size_t limit) {
    if (!have_memo_ || memo_limit_ != limit || !same_call(memo_shape_, c)) {
        memo_ti_ = plan_tiling(c, limit);
        memo_shape_ = c;
        memo_limit_ = limit;
        have_memo_ = true;
    }
    return memo_ti_;
}

size_t Engine::workspace_for(Op ta, Op tb, int64_t m, int64_t n, int64_t k, int s, bool cplx) {
    // a call with an empty C, k = 0 or alpha = 0 only scales C and needs no workspace
    if (m == 0 || n == 0 || k == 0) return 0;
    return plan_for_call(call_shape(ta, tb, m, n, k, s, cplx), budget()).workspace_bytes;
}

void Engine::gemm(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
                  const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s) {
    run(ta, tb, m, n, k, alpha, 0.0, A, lda, B, ldb, beta, 0.0, C, ldc, s, false, nullptr, 0);
}

void Engine::zgemm(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double ar, double ai, const double* A, int64_t lda,
                   const double* B, int64_t ldb, double br, double bi, double* C, int64_t ldc, int s) {
    run(ta, tb, m, n, k, ar, ai, A, lda, B, ldb, br, bi, C, ldc, s, true, nullptr, 0);
}

void Engine::gemm_fixed(std::byte* ws, size_t ws_bytes, Op ta, Op tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                        int64_t ldc, int s) {
    if (!ws) throw std::runtime_error("oz2: no fixed workspace on this device");
    run(ta, tb, m, n, k, alpha, 0.0, A, lda, B, ldb, beta, 0.0, C, ldc, s, false, ws, ws_bytes);
}

void Engine::run(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double ar, double ai, const double* A, int64_t lda,
                 const double* B, int64_t ldb, double br, double bi, double* C, int64_t ldc, int s, bool cplx,
                 std::byte* fixed_ws, size_t fixed_bytes) {
    // a call into a fixed workspace creates nothing: no stage events either
    const bool timed = profiling_ && !fixed_ws;
    if (!accumulate_) stages_.clear();
    ++calls_;
    // validation (hpp:745-748)
    if (m < 0 || n < 0 || k < 0) throw std::invalid_argument("oz2: negative dimension");
    if (lda < std::max<int64_t>(1, ta == Op::N ? m : k) || ldb < std::max<int64_t>(1, tb == Op::N ? k : n) ||
        ldc < std::max<int64_t>(1, m))
        throw std::invalid_argument("oz2: bad leading dimension");
    if (m == 0 || n == 0) return;
    if (k == 0 || (ar == 0.0 && ai == 0.0)) {  // C = beta * C (hpp:750-751, 1056-1070)
        k::scale_c(stream_, C, ldc, m, n, br, bi, cplx);
        hip_check(hipGetLastError(), "launch");
        return;
    }
    const Plan& p = plan_for(s, cplx ? 2 * k : k);  // complex plan at 2k (hpp:778-781)
    const int tk = p.tk;
    const bool wide = tk > kNarrowScaleBits;
    const TabSet* tab = dtabs_ + s;

    const CallShape cs = call_shape(ta, tb, m, n, k, s, cplx);
    if (fixed_ws) {
        // the smallest panels the planner can take: 64 x 64 (or all of a shorter m, n)
        const size_t need = layout(cs, std::min<int64_t>(m, 64), std::min<int64_t>(n, 64)).total;
        if (need > fixed_bytes)
            throw std::invalid_argument("oz2: product too large for the fixed workspace: its smallest panels need " +
                                        std::to_string((need + (size_t{1} << 20) - 1) >> 20) + " MiB, it holds " +
                                        std::to_string(fixed_bytes >> 20) + " MiB");
    }
    const TileInfo ti = plan_for_call(cs, fixed_ws ? fixed_bytes : budget());
    const int64_t mb = ti.mb, nb = ti.nb;
    const Layout L = layout(cs, mb, nb);
    const int64_t kp = L.kp, kch = L.kch, klen = L.klen;
    if (!fixed_ws) ensure_workspace(L.total);
    tiling_ = ti;
    tile_ = L.tile;
    std::byte* const ws = fixed_ws ? fixed_ws : ws_;

    int32_t* D = reinterpret_cast<int32_t*>(ws + L.oD);
    int32_t* E = reinterpret_cast<int32_t*>(ws + L.oE);
    uint64_t* PA = reinterpret_cast<uint64_t*>(ws + L.oPA);
    uint64_t* PB = reinterpret_cast<uint64_t*>(ws + L.oPB);
    int8_t* Ap = reinterpret_cast<int8_t*>(ws + L.oA);
    int8_t* Bp = reinterpret_cast<int8_t*>(ws + L.oB);
    uint8_t* R = reinterpret_cast<uint8_t*>(ws + L.oR);
    const int64_t cw = cplx ? 2 : 1;  // doubles per element

    if (timed) record(Stage::Begin);
    // scale: exponents of the rows of op(A) and the columns of op(B) (hpp:974-989)
    k::scale(stream_, A, lda, m, k, ta != Op::N, cplx, tk, PA, D);
    k::scale(stream_, B, ldb, n, k, tb == Op::N, cplx, tk, PB, E);
    if (timed) record(Stage::Scale);

    for (int64_t j0 = 0; j0 < n; j0 += nb) {
        const int64_t nbj = std::min(nb, n - j0);
        k::SplitArgs sb{};
        sb.kcontig = tb == Op::N;
        sb.X = sb.kcontig ? B + cw * j0 * ldb : B + cw * j0;
        sb.ldx = ldb;
        sb.rows = nbj;
        sb.rows_pad = round_up(nbj, L.TN);
        sb.k = k;
        sb.kp = kp;
        sb.ex = E + j0;
        sb.P = Bp;
        sb.plane = L.planeB;
        sb.group = static_cast<int64_t>(s) * L.planeB;
        sb.s = s;
        sb.cplx = cplx;
        sb.conj = cplx && tb == Op::C;
        sb.wide = wide;
        sb.tab = tab;
        k::split(stream_, sb);
        if (timed) record(Stage::Split);
        for (int64_t i0 = 0; i0 < m; i0 += mb) {
            const int64_t mbi = std::min(mb, m - i0);
            if (tiling_.pm > 1 || j0 == 0) {
                k::SplitArgs sa{};
                sa.kcontig = ta != Op::N;
                sa.X = sa.kcontig ? A + cw * i0 * lda : A + cw * i0;
                sa.ldx = lda;
                sa.rows = mbi;
                sa.rows_pad = round_up(mbi, L.TM);
                sa.k = k;
                sa.kp = kp;
                sa.ex = D + i0;
                sa.P = Ap;
                sa.plane = L.planeA;
                sa.group = static_cast<int64_t>(s) * L.planeA;
                sa.s = s;
                sa.cplx = cplx;
                sa.conj = cplx && ta == Op::C;
                sa.wide = wide;
                sa.tab = tab;
                k::split(stream_, sa);
                if (timed) record(Stage::Split);
            }
            k::GemmArgs ga{};
            ga.A = Ap;
            ga.B = Bp;
            ga.kp = kp;
            ga.plane_a = L.planeA;
            ga.plane_b = L.planeB;
            ga.plane_r = L.planeR;
            ga.ldr = L.mpad;
            ga.tiles_m = static_cast<int>((mbi + L.TM - 1) / L.TM);
            ga.tiles_n = static_cast<int>((nbj + L.TN - 1) / L.TN);
            ga.nplanes = L.planes;
            ga.s = s;
            ga.tab = tab;
            for (int64_t c = 0; c < kch; ++c) {
                ga.k0 = c * klen;
                ga.klen = std::min(klen, kp - ga.k0);
                ga.R = R + c * L.planes * L.planeR;
                if (tune_.batched_gemm) {
                    k::igemm_residue(stream_, tile_, ga);
                    continue;
                }
                // one launch per plane (the reference's per-plane GEMM loop)
                k::GemmArgs g1 = ga;
                g1.nplanes = 1;
                for (int t = 0; t < L.planes; ++t) {
                    g1.A = ga.A + t * L.planeA;
                    g1.B = ga.B + t * L.planeB;
                    g1.R = ga.R + t * L.planeR;
                    g1.plane0 = t;
                    k::igemm_residue(stream_, tile_, g1);
                }
            }
            if (timed) record(Stage::Gemm);
            k::CrtArgs ca{};
            ca.R = R;
            ca.plane_r = L.planeR;
            ca.chunk_r = static_cast<int64_t>(L.planes) * L.planeR;
            ca.kch = static_cast<int>(kch);
            ca.ldr = L.mpad;
            ca.mb = mbi;
            ca.nb = nbj;
            ca.D = D + i0;
            ca.E = E + j0;
            ca.C = C + cw * (i0 + j0 * ldc);
            ca.ldc = ldc;
            ca.ar = ar;
            ca.ai = ai;
            ca.br = br;
            ca.bi = bi;
            ca.cplx = cplx;
            ca.s = s;
            ca.L16 = 2 * p.NL;
            ca.tab = tab;
            k::crt(stream_, ca);
            if (timed) record(Stage::Crt);
        }
    }
    hip_check(hipGetLastError(), "kernel launch");
}

}  // namespace oz2
