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
    // the second stream of the column-block overlap (default priority)
    if (hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking) != hipSuccess) {
        (void)hipEventDestroy(switch_ev_);
        (void)hipFree(dtabs_);
        throw std::runtime_error("oz2: stream create failed");
    }
}

Engine::~Engine() {
    (void)hipStreamSynchronize(stream_);
    if (aux_) (void)hipStreamSynchronize(aux_);
    release_workspace();
    if (aux_) (void)hipStreamDestroy(aux_);
    for (auto e : events_) (void)hipEventDestroy(e);
    for (auto e : sync_events_) (void)hipEventDestroy(e);
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

hipEvent_t Engine::timing_event(size_t idx) {
    while (events_.size() <= idx) {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "event create");
        events_.push_back(e);
    }
    return events_[idx];
}

size_t Engine::stage_begin(Stage st, hipStream_t on) {
    const size_t idx = 2 * marks_.size();
    hip_check(hipEventRecord(timing_event(idx), on), "event record");
    marks_.push_back(Mark{st, idx});
    return marks_.size() - 1;
}

void Engine::stage_end(size_t mark, hipStream_t on) {
    hip_check(hipEventRecord(timing_event(marks_[mark].ev + 1), on), "event record");
}

hipEvent_t Engine::sync_event() {
    if (sync_used_ == sync_events_.size()) {
        hipEvent_t e;
        hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "event create");
        sync_events_.push_back(e);
    }
    return sync_events_[sync_used_++];
}

// `to` waits for everything enqueued on `from` so far
void Engine::order_after(hipStream_t to, hipStream_t from) {
    hipEvent_t e = sync_event();
    hip_check(hipEventRecord(e, from), "event record");
    hip_check(hipStreamWaitEvent(to, e, 0), "stream wait");
}

StageTimes Engine::stage_times() {
    StageTimes t;
    if (!profiling_ || marks_.empty()) return t;
    for (const Mark& mk : marks_) {
        (void)hipEventSynchronize(events_[mk.ev + 1]);
        float ms = 0;
        if (hipEventElapsedTime(&ms, events_[mk.ev], events_[mk.ev + 1]) != hipSuccess) {
            (void)hipGetLastError();
            ms = 0;
        }
        switch (mk.st) {
            case Stage::Scale: t.scale_ms += ms; break;
            case Stage::Split: t.split_ms += ms; break;
            case Stage::Gemm: t.gemm_ms += ms; break;
            case Stage::Crt: t.crt_ms += ms; break;
        }
    }
    return t;
}

void Engine::profile_begin() {
    marks_.clear();
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

int Engine::overlap_blocks() const {
    // automatic: none.  Measured on MI355X at DGEMM 8192^3 s = 16 (8.56 ms per call in one
    // stream): 2 blocks 8.62 ms, 4 blocks 8.75, 8 blocks 8.95 - the GEMM, power-limited at
    // about 1.7 GHz, slows by more than the split and CRT time it hides.
    if (tune_.overlap_blocks >= 0) return std::min(std::max(tune_.overlap_blocks, 1), kMaxOverlapBlocks);
    return 1;
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
    // under stream capture: no timing events, and the panels are planned for the workspace as it
    // is (it cannot grow there: that needs a synchronization and an allocation)
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(stream_, &cap) != hipSuccess) {
        (void)hipGetLastError();
        cap = hipStreamCaptureStatusNone;
    }
    const bool capturing = cap == hipStreamCaptureStatusActive;
    const bool timed = profiling_ && !fixed_ws && !capturing;
    if (!accumulate_) marks_.clear();
    sync_used_ = 0;
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
    const int form = limb_form(tk);
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
    const size_t limit = fixed_ws ? fixed_bytes : capturing && ws_limit_ == 0 ? ws_size_ : budget();
    const TileInfo ti = plan_for_call(cs, limit);
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

    // one stage's launches on stream `on`, bracketed by timing events when timed
    auto timed_on = [&](Stage st, hipStream_t on, auto&& launch) {
        if (!timed) {
            launch();
            return;
        }
        const size_t mk = stage_begin(st, on);
        launch();
        stage_end(mk, on);
    };
    // split of `rows` rows of op(X) starting at row r0 into the planes from plane row prow
    auto split_rows = [&](hipStream_t on, bool is_b, int64_t r0, int64_t rows, int64_t prow) {
        const Op op = is_b ? tb : ta;
        const double* X = is_b ? B : A;
        const int64_t ldx = is_b ? ldb : lda;
        k::SplitArgs sp{};
        sp.kcontig = is_b ? op == Op::N : op != Op::N;
        sp.X = sp.kcontig ? X + cw * r0 * ldx : X + cw * r0;
        sp.ldx = ldx;
        sp.rows = rows;
        sp.rows_pad = round_up(rows, is_b ? L.TN : L.TM);
        sp.k = k;
        sp.kp = kp;
        sp.ex = (is_b ? E : D) + r0;
        sp.P = (is_b ? Bp : Ap) + prow * kp;  // plane rows come in 16-row bands of kp bytes each
        sp.plane = is_b ? L.planeB : L.planeA;
        sp.group = static_cast<int64_t>(s) * sp.plane;
        sp.s = s;
        sp.cplx = cplx;
        sp.conj = cplx && op == Op::C;
        sp.form = form;
        sp.tab = tab;
        k::split(on, sp);
    };
    // INT8 GEMMs of `rows` rows of the A panel by columns c0 .. c0 + cols of the B panel
    auto gemm_block = [&](hipStream_t on, int64_t rows, int64_t c0, int64_t cols) {
        k::GemmArgs ga{};
        ga.A = Ap;
        ga.B = Bp + c0 * kp;
        ga.kp = kp;
        ga.plane_a = L.planeA;
        ga.plane_b = L.planeB;
        ga.plane_r = L.planeR;
        ga.ldr = L.mpad;
        ga.tiles_m = static_cast<int>((rows + L.TM - 1) / L.TM);
        ga.tiles_n = static_cast<int>((cols + L.TN - 1) / L.TN);
        ga.nplanes = L.planes;
        ga.s = s;
        ga.tab = tab;
        for (int64_t c = 0; c < kch; ++c) {
            ga.k0 = c * klen;
            ga.klen = std::min(klen, kp - ga.k0);
            ga.R = R + c * L.planes * L.planeR + c0 * L.mpad;
            if (tune_.batched_gemm) {
                k::igemm_residue(on, tile_, ga);
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
                k::igemm_residue(on, tile_, g1);
            }
        }
    };
    // CRT + alpha/beta epilogue of rows i0 .. i0 + rows by panel columns c0 .. c0 + cols of panel j0
    auto crt_block = [&](hipStream_t on, int64_t i0, int64_t rows, int64_t j0, int64_t c0, int64_t cols) {
        k::CrtArgs ca{};
        ca.R = R + c0 * L.mpad;
        ca.plane_r = L.planeR;
        ca.chunk_r = static_cast<int64_t>(L.planes) * L.planeR;
        ca.kch = static_cast<int>(kch);
        ca.ldr = L.mpad;
        ca.mb = rows;
        ca.nb = cols;
        ca.D = D + i0;
        ca.E = E + j0 + c0;
        ca.C = C + cw * (i0 + (j0 + c0) * ldc);
        ca.ldc = ldc;
        ca.ar = ar;
        ca.ai = ai;
        ca.br = br;
        ca.bi = bi;
        ca.cplx = cplx;
        ca.s = s;
        ca.L16 = 2 * p.NL;
        ca.tab = tab;
        k::crt(on, ca);
    };

    // scale: exponents of the rows of op(A) and the columns of op(B) (hpp:974-989)
    timed_on(Stage::Scale, stream_, [&] {
        k::scale(stream_, A, lda, m, k, ta != Op::N, cplx, tk, PA, D);
        k::scale(stream_, B, ldb, n, k, tb == Op::N, cplx, tk, PB, E);
    });
    // Column blocks of a panel pair: with more than one, the B blocks are split and the CRTs run on
    // the second stream while the GEMM of the next block runs on the caller's (a fixed workspace
    // keeps the one-stream order: it records no events)
    const int nblocks = fixed_ws ? 1 : overlap_blocks();
    bool forked = false;
    for (int64_t j0 = 0; j0 < n; j0 += nb) {
        const int64_t nbj = std::min(nb, n - j0);
        const int64_t wb = nblocks > 1 ? round_up((nbj + nblocks - 1) / nblocks, L.TN) : nbj;
        const int64_t nblk = (nbj + wb - 1) / wb;
        if (nblk == 1) timed_on(Stage::Split, stream_, [&] { split_rows(stream_, true, j0, nbj, 0); });
        for (int64_t i0 = 0; i0 < m; i0 += mb) {
            const int64_t mbi = std::min(mb, m - i0);
            if (ti.pm > 1 || j0 == 0) timed_on(Stage::Split, stream_, [&] { split_rows(stream_, false, i0, mbi, 0); });
            if (nblk == 1) {
                timed_on(Stage::Gemm, stream_, [&] { gemm_block(stream_, mbi, 0, nbj); });
                timed_on(Stage::Crt, stream_, [&] { crt_block(stream_, i0, mbi, j0, 0, nbj); });
                continue;
            }
            if (!forked) {  // the second stream starts after everything before it on the caller's
                order_after(aux_, stream_);
                forked = true;
            }
            hipEvent_t split_done[kMaxOverlapBlocks], gemm_done[kMaxOverlapBlocks];
            if (i0 == 0)  // the B panel's blocks, once per panel column, ahead of the CRTs
                for (int64_t c = 0; c < nblk; ++c) {
                    const int64_t c0 = c * wb, cols = std::min(wb, nbj - c0);
                    timed_on(Stage::Split, aux_, [&] { split_rows(aux_, true, j0 + c0, cols, c0); });
                    split_done[c] = sync_event();
                    hip_check(hipEventRecord(split_done[c], aux_), "event record");
                }
            for (int64_t c = 0; c < nblk; ++c) {
                const int64_t c0 = c * wb, cols = std::min(wb, nbj - c0);
                if (i0 == 0) hip_check(hipStreamWaitEvent(stream_, split_done[c], 0), "stream wait");
                timed_on(Stage::Gemm, stream_, [&] { gemm_block(stream_, mbi, c0, cols); });
                gemm_done[c] = sync_event();
                hip_check(hipEventRecord(gemm_done[c], stream_), "event record");
            }
            for (int64_t c = 0; c < nblk; ++c) {
                const int64_t c0 = c * wb, cols = std::min(wb, nbj - c0);
                hip_check(hipStreamWaitEvent(aux_, gemm_done[c], 0), "stream wait");
                timed_on(Stage::Crt, aux_, [&] { crt_block(aux_, i0, mbi, j0, c0, cols); });
            }
            // the next panel pair rewrites the residues the CRTs read, and its B split the planes
            // the GEMMs read (the GEMMs are done once the CRTs are: each waited for its GEMM)
            if (i0 + mb < m || j0 + nb < n) order_after(stream_, aux_);
        }
    }
    if (forked) order_after(stream_, aux_);  // the call ends on the caller's stream
    hip_check(hipGetLastError(), "kernel launch");
}

// Ozaki Scheme I (ozaki1_gpu.hpp:100-151): exponents for tk = 7L - 2 from the scale kernels,
// L balanced base-128 digits per element (oz1_slices), one int32-output INT8 GEMM per
// anti-diagonal over the digit k-slots (chunks keep |sums| < 2^31), and the double-double
// Horner over the anti-diagonals (oz1_horner).
void Engine::ozaki1(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb,
                    double* C, int64_t ldc, int L) {
    if (L < 2 || L > 9) throw std::invalid_argument("oz1: L must be in [2, 9]");
    if (k * 4096 >= (int64_t{1} << 31)) throw std::invalid_argument("oz1: k too large (int32 GEMM)");
    if (m < 0 || n < 0 || k < 0) throw std::invalid_argument("oz2: negative dimension");
    if (lda < std::max<int64_t>(1, m) || ldb < std::max<int64_t>(1, k) || ldc < std::max<int64_t>(1, m))
        throw std::invalid_argument("oz2: bad leading dimension");
    if (m == 0 || n == 0) return;
    if (k == 0) {  // every digit product is empty: C = 0
        k::scale_c(stream_, C, ldc, m, n, 0.0, 0.0, false);
        hip_check(hipGetLastError(), "launch");
        return;
    }
    const int tk = 7 * L - 2;
    const int64_t kp = round_up(k, kKStep), kpx = L * kp;
    const k::Tile tile = pick_tile(m, n, kp, L);
    const int64_t TM = k::tile_m(tile), TN = k::tile_n(tile);
    const int64_t mpad = round_up(m, TM), npad = round_up(n, TN);
    // chunks of at most kChunk k: |sum| <= kChunk 64^2 < 2^31
    constexpr int64_t kChunk = 523264;
    const int64_t maxch = (L * kp + kChunk - 1) / kChunk;
    size_t off = 0;
    auto carve = [&](size_t bytes) {
        const size_t o = off;
        off = (off + bytes + 255) / 256 * 256;
        return o;
    };
    const size_t oD = carve(static_cast<size_t>(m) * 4), oE = carve(static_cast<size_t>(n) * 4);
    const size_t oPA = carve(k::scale_partial_elems(m, k, false) * 8), oPB = carve(k::scale_partial_elems(n, k, true) * 8);
    const size_t oA = carve(static_cast<size_t>(mpad) * kpx), oB = carve(static_cast<size_t>(npad) * kpx);
    const size_t oC = carve(static_cast<size_t>(maxch) * mpad * npad * 4);
    const size_t oH = carve(static_cast<size_t>(m) * n * 8), oL = carve(static_cast<size_t>(m) * n * 8);
    ensure_workspace(off);
    int32_t* D = reinterpret_cast<int32_t*>(ws_ + oD);
    int32_t* E = reinterpret_cast<int32_t*>(ws_ + oE);
    int8_t* Ap = reinterpret_cast<int8_t*>(ws_ + oA);
    int8_t* Bp = reinterpret_cast<int8_t*>(ws_ + oB);
    int32_t* Ct = reinterpret_cast<int32_t*>(ws_ + oC);
    double* Th = reinterpret_cast<double*>(ws_ + oH);
    double* Tl = reinterpret_cast<double*>(ws_ + oL);

    k::scale(stream_, A, lda, m, k, false, false, tk, reinterpret_cast<uint64_t*>(ws_ + oPA), D);
    k::scale(stream_, B, ldb, n, k, true, false, tk, reinterpret_cast<uint64_t*>(ws_ + oPB), E);
    k::oz1_slices(stream_, k::Oz1SliceArgs{A, lda, m, mpad, k, kp, D, Ap, false, false}, L);
    k::oz1_slices(stream_, k::Oz1SliceArgs{B, ldb, n, npad, k, kp, E, Bp, true, true}, L);
    for (int d = 0; d < L; ++d) {
        const int64_t len = (d + 1) * kp;
        const int nch = static_cast<int>((len + kChunk - 1) / kChunk);
        k::GemmArgs ga{};
        ga.A = Ap;
        ga.B = Bp;
        ga.kp = kpx;
        ga.ldr = mpad;
        ga.tiles_m = static_cast<int>(mpad / TM);
        ga.tiles_n = static_cast<int>(npad / TN);
        ga.nplanes = 1;
        ga.s = kMinModuli;
        ga.tab = dtabs_ + kMinModuli;
        ga.kb_extra = (L - 1 - d) * kp;  // B's digits d .. 0 sit at k-slots L-1-d .. L-1
        for (int c = 0; c < nch; ++c) {
            ga.k0 = c * kChunk;
            ga.klen = std::min(kChunk, len - ga.k0);
            ga.R32 = Ct + static_cast<int64_t>(c) * mpad * npad;
            k::igemm_int32(stream_, tile, ga);
        }
        k::Oz1HornerArgs h{Ct, mpad, mpad * npad, nch, Th, Tl, d == 0, d == L - 1, 7 * (L - 1), D, E, C, m, n, ldc};
        k::oz1_horner(stream_, h);
    }
    hip_check(hipGetLastError(), "kernel launch");
}

}  // namespace oz2
