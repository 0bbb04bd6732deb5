// c_api.cpp - the extern "C" entry points of liboz2.so (declared in include/oz2/oz2_c.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>
#include <stdexcept>
#include <string>

#include "../../include/oz2/oz2_c.h"
#include "batch.hpp"
#include "selector.hpp"
#include "engine.hpp"
#include "capi.hpp"

struct oz2_batch_s {
    oz2_engine* h = nullptr;
    std::unique_ptr<oz2::BatchPlan> plan;
    int count = 0;
    // what the plan was made from, for oz2_batch_reselect
    std::vector<oz2_gemm_desc_t> descs;
    std::vector<int> tiles;
    int bits = 0, flags = 0;
};

namespace {
using oz2::capi::guard;
using oz2::capi::need;
using oz2::capi::signal;
using oz2::capi::to_op;
using oz2::capi::wait_deps;

oz2::SelProblem sel_problem(const oz2::GemmDesc& g, int bits) {
    return oz2::SelProblem{g.ta == oz2::Op::N ? 0 : 1, g.tb == oz2::Op::N ? 0 : 1, g.m, g.n, g.k, g.A, g.lda, g.B,
                           g.ldb, bits};
}

std::vector<oz2::GemmDesc> descs_of(int count, const oz2_gemm_desc_t* descs) {
    if (count < 0 || (count > 0 && !descs)) throw std::invalid_argument("oz2: bad problem list");
    std::vector<oz2::GemmDesc> d(count);
    for (int i = 0; i < count; ++i) {
        const oz2_gemm_desc_t& c = descs[i];
        if (c.m < 0 || c.n < 0 || c.k < 0) throw std::invalid_argument("oz2: negative dimension");
        d[i] = oz2::GemmDesc{to_op(c.ta), to_op(c.tb), c.m, c.n, c.k, c.alpha, c.A, c.lda,
                             c.B, c.ldb, c.beta, c.C, c.ldc, c.num_moduli};
    }
    return d;
}

// learned choice of s for the problems with num_moduli <= 0 (moduli_for_bits where the model abstains)
void select_moduli(oz2_handle h, const std::vector<oz2::GemmDesc>& d, int bits, std::vector<int>& s) {
    std::vector<oz2::SelProblem> probs;
    std::vector<size_t> idx;
    for (size_t i = 0; i < d.size(); ++i)
        if (d[i].s <= 0 && d[i].m > 0 && d[i].n > 0 && d[i].k > 0) {
            probs.push_back(sel_problem(d[i], bits));
            idx.push_back(i);
        }
    if (probs.empty()) return;
    const std::vector<int> c = h->sel.choose(h->eng.stream(), probs);
    for (size_t j = 0; j < idx.size(); ++j) s[idx[j]] = c[j];
}

// device tables for the kernel-level entry points: one copy per device, kept until exit
const oz2::TabSet* unit_tables() {
    static std::mutex mu;
    static std::map<int, oz2::TabSet*> per_dev;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("oz2: hipGetDevice failed");
    std::lock_guard<std::mutex> lock(mu);
    auto it = per_dev.find(dev);
    if (it != per_dev.end()) return it->second;
    std::vector<oz2::TabSet> host(oz2::kMaxModuli + 1);
    oz2::make_tabsets(host.data());
    void* p = nullptr;
    if (hipMalloc(&p, host.size() * sizeof(oz2::TabSet)) != hipSuccess)
        throw std::runtime_error("oz2: table allocation failed");
    if (hipMemcpy(p, host.data(), host.size() * sizeof(oz2::TabSet), hipMemcpyHostToDevice) != hipSuccess)
        throw std::runtime_error("oz2: table upload failed");
    per_dev[dev] = static_cast<oz2::TabSet*>(p);
    return per_dev[dev];
}

// descriptors and moduli of a batch as oz2_batch_create chooses them
void batch_inputs(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, int flags,
                  std::vector<oz2::GemmDesc>& d, std::vector<int>& s) {
    if (count < 0 || (count > 0 && !descs)) throw std::invalid_argument("oz2: bad batch");
    d.assign(count, oz2::GemmDesc{});
    s.assign(count, 0);
    for (int i = 0; i < count; ++i) {
        const oz2_gemm_desc_t& c = descs[i];
        d[i] = oz2::GemmDesc{to_op(c.ta), to_op(c.tb), c.m, c.n, c.k, c.alpha, c.A, c.lda,
                             c.B, c.ldb, c.beta, c.C, c.ldc, c.num_moduli};
        s[i] = c.num_moduli > 0 ? c.num_moduli : oz2::moduli_for_bits(std::max<int64_t>(c.k, 1), bits);
    }
    // shapes are validated by the plan; the selector only reads problems that would pass it
    for (const auto& g : d)
        if (g.m < 0 || g.n < 0 || g.k < 0) throw std::invalid_argument("oz2: negative dimension");
    if ((flags & OZ2_BATCH_SELECT) != 0) select_moduli(h, d, bits, s);
}

oz2::BatchOptions batch_options(int flags) {
    oz2::BatchOptions o;
    o.persistent = (flags & OZ2_BATCH_PERSISTENT) != 0;
    o.largest_first = (flags & OZ2_BATCH_PROBLEM_ORDER) == 0;
    return o;
}
}  // namespace

namespace oz2 {
namespace capi {
std::string& last_error() {
    thread_local std::string err;
    return err;
}
}  // namespace capi
}  // namespace oz2

extern "C" {

int oz2_create(void* stream, oz2_handle* out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        auto* h = new oz2_engine(static_cast<hipStream_t>(stream));
        // the device's scratch of oz2_dgemm_noalloc is looked up here, at set-up
        oz2::capi::noalloc_attach(h->eng.device());
        *out = h;
    });
}

int oz2_destroy(oz2_handle h) {
    return guard([&] { delete h; });
}

int oz2_set_stream(oz2_handle h, void* stream) {
    return guard([&] {
        need(h);
        h->eng.set_stream(static_cast<hipStream_t>(stream));
    });
}

int oz2_dgemm(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int num_moduli,
              void* const* deps, int ndeps, void* done) {
    return guard([&] {
        need(h);
        const oz2::Op a = to_op(ta), b = to_op(tb);
        wait_deps(h->eng.stream(), deps, ndeps);
        // Trans::C equals T for real data (hpp:1001)
        h->eng.gemm(a == oz2::Op::C ? oz2::Op::T : a, b == oz2::Op::C ? oz2::Op::T : b, m, n, k, alpha, A, lda, B,
                    ldb, beta, C, ldc, num_moduli);
        signal(h->eng.stream(), done);
    });
}

int oz2_zgemm(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, const double* alpha, const double* A,
              int64_t lda, const double* B, int64_t ldb, const double* beta, double* C, int64_t ldc, int num_moduli,
              void* const* deps, int ndeps, void* done) {
    return guard([&] {
        need(h);
        if (!alpha || !beta) throw std::invalid_argument("oz2: null alpha/beta");
        const oz2::Op a = to_op(ta), b = to_op(tb);
        wait_deps(h->eng.stream(), deps, ndeps);
        h->eng.zgemm(a, b, m, n, k, alpha[0], alpha[1], A, lda, B, ldb, beta[0], beta[1], C, ldc, num_moduli);
        signal(h->eng.stream(), done);
    });
}

int oz2_precision_bits(int64_t k, int num_moduli, int* err) {
    int tk = -1;
    const int rc = guard([&] { tk = oz2::make_plan(num_moduli, k).tk; });
    if (err) *err = rc;
    return rc == OZ2_OK ? tk : -1;
}

int oz2_precision_bits_complex(int64_t k, int num_moduli, int* err) {
    return oz2_precision_bits(2 * k, num_moduli, err);
}

int oz2_moduli_for_bits(int64_t k, int bits) { return oz2::moduli_for_bits(k, bits); }

int oz2_set_workspace_limit(oz2_handle h, size_t bytes) {
    return guard([&] { need(h); h->eng.set_workspace_limit(bytes); });
}
int oz2_set_resident_hint(oz2_handle h, size_t bytes) {
    return guard([&] { need(h); h->eng.set_resident_hint(bytes); });
}
int oz2_release_workspace(oz2_handle h) {
    return guard([&] { need(h); h->eng.release_workspace(); });
}
int oz2_reserve_workspace(oz2_handle h, size_t bytes) {
    return guard([&] { need(h); h->eng.reserve_workspace(bytes); });
}
int oz2_workspace_query(oz2_handle h, int ta, int tb, int64_t m, int64_t n, int64_t k, int num_moduli, int cplx,
                        size_t* bytes) {
    return guard([&] {
        need(h);
        if (!bytes) throw std::invalid_argument("oz2: null output pointer");
        oz2::Op a = to_op(ta), b = to_op(tb);
        if (!cplx) {
            a = a == oz2::Op::C ? oz2::Op::T : a;
            b = b == oz2::Op::C ? oz2::Op::T : b;
        }
        if (m < 0 || n < 0 || k < 0) throw std::invalid_argument("oz2: negative dimension");
        // m, n or k = 0 only scales C (as oz2_dgemm does): no plan and no workspace
        if (m == 0 || n == 0 || k == 0) {
            *bytes = 0;
            return;
        }
        oz2::make_plan(num_moduli, cplx ? 2 * k : k);
        *bytes = h->eng.workspace_for(a, b, m, n, k, num_moduli, cplx != 0);
    });
}

int oz2_stage_times(oz2_handle h, oz2_stage_times_t* out) {
    return guard([&] {
        need(h);
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        const oz2::StageTimes t = h->eng.stage_times();
        *out = oz2_stage_times_t{t.scale_ms, t.split_ms, t.gemm_ms, t.crt_ms};
    });
}

int oz2_last_tiling(oz2_handle h, oz2_tile_info_t* out) {
    return guard([&] {
        need(h);
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        const oz2::TileInfo& t = h->eng.last_tiling();
        *out = oz2_tile_info_t{t.mb, t.nb, t.pm, t.pn, t.k_chunks, t.workspace_bytes};
    });
}

int oz2_set_tuning(oz2_handle h, const oz2_tuning_t* t) {
    return guard([&] {
        need(h);
        if (!t) throw std::invalid_argument("oz2: null tuning");
        oz2::k::Tile tile;
        if (t->gemm_tile != -1 && !oz2::k::tile_from_code(t->gemm_tile, tile))
            throw std::invalid_argument("oz2: gemm_tile must be -1, 64, 128, 256 or 256128");
        oz2::Tuning u;
        u.batched_gemm = t->batched_gemm != 0;
        u.pad_lda = t->pad_lda != 0;
        u.pow2_tiles = t->pow2_tiles != 0;
        u.gemm_tile = t->gemm_tile;
        if (t->overlap_blocks < -1 || t->overlap_blocks > oz2::kMaxOverlapBlocks)
            throw std::invalid_argument("oz2: overlap_blocks must be -1 (automatic) or 0 .. 16");
        u.overlap_blocks = t->overlap_blocks;
        h->eng.set_tuning(u);
    });
}

int oz2_get_tuning(oz2_handle h, oz2_tuning_t* out) {
    return guard([&] {
        need(h);
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        const oz2::Tuning& u = h->eng.tuning();
        *out = oz2_tuning_t{u.batched_gemm, u.pad_lda, u.pow2_tiles, u.gemm_tile, u.overlap_blocks};
    });
}

int oz2_set_profiling(oz2_handle h, int on) {
    return guard([&] { need(h); h->eng.set_profiling(on != 0); });
}

int oz2_profile_begin(oz2_handle h) {
    return guard([&] { need(h); h->eng.profile_begin(); });
}

int oz2_profile_end(oz2_handle h, oz2_stage_times_t* sum, int64_t* calls) {
    return guard([&] {
        need(h);
        if (!sum) throw std::invalid_argument("oz2: null output pointer");
        const oz2::StageTimes t = h->eng.profile_end(calls);
        *sum = oz2_stage_times_t{t.scale_ms, t.split_ms, t.gemm_ms, t.crt_ms};
    });
}

int oz2_moduli(int num_moduli, uint16_t* out) {
    return guard([&] {
        const uint16_t* ms = oz2::moduli(num_moduli);
        if (!ms) throw std::invalid_argument("oz2: num_moduli must be in [2, 24]");
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        for (int t = 0; t < num_moduli; ++t) out[t] = ms[t];
    });
}


int oz2_batch_create(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, const int* tiles, int flags,
                     oz2_batch* out) {
    return guard([&] {
        need(h);
        if (!out) throw std::invalid_argument("oz2: null output");
        std::vector<oz2::GemmDesc> d;
        std::vector<int> s, t;
        batch_inputs(h, count, descs, bits, flags, d, s);
        if (tiles) t.assign(tiles, tiles + count);
        auto b = std::make_unique<oz2_batch_s>();
        b->h = h;
        b->count = count;
        b->descs.assign(descs, descs + count);
        b->tiles = t;
        b->bits = bits;
        b->flags = flags;
        b->plan = std::make_unique<oz2::BatchPlan>(h->eng, std::move(d), s, t, (flags & OZ2_BATCH_GRAPH) != 0,
                                                   batch_options(flags));
        *out = b.release();
    });
}

int oz2_batch_reselect(oz2_batch b, int* changed) {
    return guard([&] {
        if (!b) throw std::invalid_argument("oz2: null batch");
        int nchanged = 0;
        if ((b->flags & OZ2_BATCH_SELECT) != 0) {
            std::vector<oz2::GemmDesc> d;
            std::vector<int> s;
            // the selector reads the buffers' current contents (it synchronizes the stream first)
            batch_inputs(b->h, b->count, b->descs.data(), b->bits, b->flags, d, s);
            const std::vector<int>& old = b->plan->moduli();
            for (int i = 0; i < b->count; ++i) nchanged += s[i] != old[i];
            if (nchanged) {
                b->plan.reset();  // waits for the old plan's work and frees its workspace and graph
                b->plan = std::make_unique<oz2::BatchPlan>(b->h->eng, std::move(d), s, b->tiles,
                                                           (b->flags & OZ2_BATCH_GRAPH) != 0, batch_options(b->flags));
            }
        }
        if (changed) *changed = nchanged;
    });
}

int oz2_batch_autotune(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, int flags, int* tiles_out,
                       int* flags_out, double* best_ms) {
    return guard([&] {
        need(h);
        std::vector<oz2::GemmDesc> d;
        std::vector<int> s;
        batch_inputs(h, count, descs, bits, flags, d, s);
        const oz2::TuneResult r = oz2::autotune_batch(h->eng, d, s);
        if (tiles_out) std::copy(r.tiles.begin(), r.tiles.end(), tiles_out);
        if (flags_out)
            *flags_out = (r.opt.persistent ? OZ2_BATCH_PERSISTENT : 0) | (r.opt.largest_first ? 0 : OZ2_BATCH_PROBLEM_ORDER);
        if (best_ms) *best_ms = r.ms;
    });
}

int oz2_batch_run(oz2_batch b) {
    return guard([&] {
        if (!b) throw std::invalid_argument("oz2: null batch");
        b->plan->run(b->h->eng.stream());
    });
}

int oz2_batch_moduli(oz2_batch b, int* out) {
    return guard([&] {
        if (!b || !out) throw std::invalid_argument("oz2: null argument");
        std::copy(b->plan->moduli().begin(), b->plan->moduli().end(), out);
    });
}

int oz2_batch_tiles(oz2_batch b, int* out) {
    return guard([&] {
        if (!b || !out) throw std::invalid_argument("oz2: null argument");
        std::copy(b->plan->tiles().begin(), b->plan->tiles().end(), out);
    });
}

int oz2_batch_info(oz2_batch b, size_t* workspace_bytes, int* launches, int* graph) {
    return guard([&] {
        if (!b) throw std::invalid_argument("oz2: null batch");
        if (workspace_bytes) *workspace_bytes = b->plan->workspace_bytes();
        if (launches) *launches = b->plan->launches();
        if (graph) *graph = b->plan->uses_graph() ? 1 : 0;
    });
}

int oz2_batch_destroy(oz2_batch b) {
    return guard([&] { delete b; });
}

int oz2_selector_features(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, float* features) {
    return guard([&] {
        need(h);
        if (!features && count > 0) throw std::invalid_argument("oz2: null output");
        const std::vector<oz2::GemmDesc> d = descs_of(count, descs);
        std::vector<oz2::SelProblem> probs;
        for (const auto& g : d) probs.push_back(sel_problem(g, bits));
        std::vector<float> f;
        h->sel.evaluate(h->eng.stream(), probs, &f, nullptr);
        std::copy(f.begin(), f.end(), features);
    });
}

int oz2_selector_choose(oz2_handle h, int count, const oz2_gemm_desc_t* descs, int bits, float* probs, int* s_out,
                        int* learned) {
    return guard([&] {
        need(h);
        const std::vector<oz2::GemmDesc> d = descs_of(count, descs);
        std::vector<oz2::SelProblem> pr;
        for (const auto& g : d) pr.push_back(sel_problem(g, bits));
        if (probs) {
            std::vector<float> p;
            h->sel.evaluate(h->eng.stream(), pr, nullptr, &p);
            std::copy(p.begin(), p.end(), probs);
        }
        std::vector<bool> l;
        const std::vector<int> s = h->sel.choose(h->eng.stream(), pr, &l);
        if (s_out) std::copy(s.begin(), s.end(), s_out);
        if (learned)
            for (size_t i = 0; i < l.size(); ++i) learned[i] = l[i] ? 1 : 0;
    });
}

int oz2_selector_set_weights(oz2_handle h, const float* weights, size_t n, float threshold) {
    return guard([&] {
        need(h);
        if (n && !weights) throw std::invalid_argument("oz2: null weights");
        h->sel.set(weights, n, threshold);
    });
}

int oz2_selector_get_weights(oz2_handle h, float* weights, size_t n, float* threshold, size_t* count) {
    return guard([&] {
        need(h);
        const auto& w = h->sel.weights();
        if (weights && n) std::copy(w.begin(), w.begin() + std::min(n, w.size()), weights);
        if (threshold) *threshold = h->sel.threshold();
        if (count) *count = w.size();
    });
}

int oz2_mlp_forward(oz2_handle h, const float* X, int nb, const float* W, float* H1, float* H2, float* Z, float* P) {
    return guard([&] {
        need(h);
        if (nb < 0 || (nb > 0 && (!X || !W))) throw std::invalid_argument("oz2: bad mlp arguments");
        oz2::k::mlp_forward(h->eng.stream(), X, nb, W, H1, H2, Z, P);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: mlp launch failed");
    });
}

int oz2_mlp_backward(oz2_handle h, const float* X, int nb, const float* W, const float* H1, const float* H2,
                     const float* Z, const float* Y, const float* M, float inv_norm, float* D1, float* D2, float* D3,
                     float* G, float* loss) {
    return guard([&] {
        need(h);
        if (nb < 0 || (nb > 0 && (!X || !W || !H1 || !H2 || !Z || !Y || !D1 || !D2 || !D3 || !G)))
            throw std::invalid_argument("oz2: bad mlp arguments");
        oz2::k::mlp_backward(h->eng.stream(), X, nb, W, H1, H2, Z, Y, M, inv_norm, D1, D2, D3, G, loss);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: mlp launch failed");
    });
}

int oz2_mlp_adam(oz2_handle h, float* W, const float* G, float* M, float* V, int step, float lr, float beta1,
                 float beta2, float eps, float wd) {
    return guard([&] {
        need(h);
        if (!W || !G || !M || !V || step < 1) throw std::invalid_argument("oz2: bad adam arguments");
        oz2::k::mlp_adam(h->eng.stream(), W, G, M, V, step, lr, beta1, beta2, eps, wd);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: mlp launch failed");
    });
}

int oz2_unit_split(void* stream, const double* X, int64_t ldx, int64_t rows, int64_t k, int kcontig, int cplx, int conj,
                   const int32_t* ex, int num_moduli, int tk, int8_t* planes, int64_t rows_pad, int64_t kp) {
    return guard([&] {
        if (num_moduli < oz2::kMinModuli || num_moduli > oz2::kMaxModuli)
            throw std::invalid_argument("oz2: num_moduli must be in [2, 24]");
        if (rows < 0 || k < 0 || rows_pad < rows || rows_pad % 32 || kp < k || kp % 128 || tk < 1 ||
            tk > oz2::kMaxScaleBits)
            throw std::invalid_argument("oz2: bad unit_split shape");
        oz2::k::SplitArgs a{};
        a.X = X;
        a.ldx = ldx;
        a.rows = rows;
        a.rows_pad = rows_pad;
        a.k = k;
        a.kp = kp;
        a.ex = ex;
        a.P = planes;
        a.plane = rows_pad * kp;
        a.group = static_cast<int64_t>(num_moduli) * a.plane;
        a.s = num_moduli;
        a.kcontig = kcontig != 0;
        a.cplx = cplx != 0;
        a.conj = conj != 0;
        a.form = oz2::limb_form(tk);
        a.tab = unit_tables() + num_moduli;
        oz2::k::split(static_cast<hipStream_t>(stream), a);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: split launch failed");
    });
}

int oz2_unit_crt(void* stream, const uint8_t* R, int64_t ldr, int64_t mb, int64_t nb, const int32_t* D,
                 const int32_t* E, int num_moduli, double alpha, double beta, double* C, int64_t ldc) {
    return guard([&] {
        if (num_moduli < oz2::kMinModuli || num_moduli > oz2::kMaxModuli)
            throw std::invalid_argument("oz2: num_moduli must be in [2, 24]");
        if (mb < 0 || nb < 0 || ldr % 4 || ldr < ((mb + 3) / 4) * 4 || ldc < std::max<int64_t>(1, mb))
            throw std::invalid_argument("oz2: bad unit_crt shape");
        oz2::k::CrtArgs a{};
        a.R = R;
        a.plane_r = ldr * nb;
        a.chunk_r = a.plane_r * num_moduli;
        a.kch = 1;
        a.ldr = ldr;
        a.mb = mb;
        a.nb = nb;
        a.D = D;
        a.E = E;
        a.C = C;
        a.ldc = ldc;
        a.ar = alpha;
        a.br = beta;
        a.s = num_moduli;
        a.L16 = 2 * oz2::limbs_for(num_moduli);
        a.tab = unit_tables() + num_moduli;
        oz2::k::crt(static_cast<hipStream_t>(stream), a);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: crt launch failed");
    });
}

int oz2_unit_gemm(void* stream, const int8_t* A, const int8_t* B, int64_t mpad, int64_t npad, int64_t kp,
                  int num_moduli, int tile, uint8_t* R, int64_t ldr) {
    return guard([&] {
        if (num_moduli < oz2::kMinModuli || num_moduli > oz2::kMaxModuli)
            throw std::invalid_argument("oz2: num_moduli must be in [2, 24]");
        oz2::k::Tile t;
        if (!oz2::k::tile_from_code(tile, t)) throw std::invalid_argument("oz2: tile must be 64, 128, 256 or 256128");
        const int64_t TM = oz2::k::tile_m(t), TN = oz2::k::tile_n(t);
        if (mpad <= 0 || npad <= 0 || mpad % TM || npad % TN || kp <= 0 || kp % oz2::kKStep || kp > oz2::kMaxKChunk ||
            ldr < mpad || ldr % 16 || (mpad / TM) * (npad / TN) * num_moduli > (int64_t{1} << 31))
            throw std::invalid_argument("oz2: bad unit_gemm shape");
        oz2::k::GemmArgs a{};
        a.A = A;
        a.B = B;
        a.R = R;
        a.kp = kp;
        a.plane_a = mpad * kp;
        a.plane_b = npad * kp;
        a.plane_r = ldr * npad;
        a.ldr = ldr;
        a.k0 = 0;
        a.klen = kp;
        a.tiles_m = static_cast<int>(mpad / TM);
        a.tiles_n = static_cast<int>(npad / TN);
        a.nplanes = num_moduli;
        a.plane0 = 0;
        a.s = num_moduli;
        a.tab = unit_tables() + num_moduli;
        oz2::k::igemm_residue(static_cast<hipStream_t>(stream), t, a);
        if (hipGetLastError() != hipSuccess) throw std::runtime_error("oz2: gemm launch failed");
    });
}

int oz2_ozaki1_dgemm(oz2_handle h, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B,
                     int64_t ldb, double* C, int64_t ldc, int L) {
    return guard([&] {
        need(h);
        h->eng.ozaki1(m, n, k, A, lda, B, ldb, C, ldc, L);
    });
}

const char* oz2_last_error(void) { return oz2::capi::last_error().c_str(); }
const char* oz2_version(void) { return "oz2-mi355x 0.1"; }

}  // extern "C"
