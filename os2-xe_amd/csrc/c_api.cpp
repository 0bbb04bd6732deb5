// c_api.cpp - the extern "C" entry points of liboz2.so (declared in include/oz2/oz2_c.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include <map>
#include <mutex>
#include <new>
#include <vector>
#include <stdexcept>
#include <string>

#include "../../include/oz2/oz2_c.h"
#include "engine.hpp"

struct oz2_engine {
    oz2::Engine eng;
    explicit oz2_engine(hipStream_t s) : eng(s) {}
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return OZ2_OK;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return OZ2_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        g_err = e.what();
        return OZ2_OVERFLOW_ERROR;
    } catch (const std::logic_error& e) {
        g_err = e.what();
        return OZ2_LOGIC_ERROR;
    } catch (const std::bad_alloc&) {
        g_err = "oz2: host allocation failed";
        return OZ2_RUNTIME_ERROR;
    } catch (const std::exception& e) {
        g_err = e.what();
        return OZ2_RUNTIME_ERROR;
    } catch (...) {
        g_err = "oz2: unknown error";
        return OZ2_RUNTIME_ERROR;
    }
}

oz2::Op to_op(int t) {
    if (t == OZ2_TRANS_N) return oz2::Op::N;
    if (t == OZ2_TRANS_T) return oz2::Op::T;
    if (t == OZ2_TRANS_C) return oz2::Op::C;
    throw std::invalid_argument("oz2: bad transpose code");
}

void need(oz2_handle h) {
    if (!h) throw std::invalid_argument("oz2: null handle");
}

void wait_deps(hipStream_t st, void* const* deps, int ndeps) {
    for (int i = 0; i < ndeps; ++i)
        if (deps && deps[i])
            if (hipStreamWaitEvent(st, static_cast<hipEvent_t>(deps[i]), 0) != hipSuccess)
                throw std::runtime_error("oz2: hipStreamWaitEvent failed");
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

void signal(hipStream_t st, void* done) {
    if (done && hipEventRecord(static_cast<hipEvent_t>(done), st) != hipSuccess)
        throw std::runtime_error("oz2: hipEventRecord failed");
}
}  // namespace

extern "C" {

int oz2_create(void* stream, oz2_handle* out) {
    return guard([&] {
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        *out = new oz2_engine(static_cast<hipStream_t>(stream));
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

int oz2_moduli_for_bits(int64_t k, int bits) {
    // smallest s whose plan keeps >= bits; 24 if none does (hpp:693-700)
    for (int s = oz2::kMinModuli; s <= oz2::kMaxModuli; ++s) {
        try {
            if (oz2::make_plan(s, k).tk >= bits) return s;
        } catch (const std::invalid_argument&) {
        }
    }
    return oz2::kMaxModuli;
}

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
        if (t->gemm_tile != -1 && t->gemm_tile != 64 && t->gemm_tile != 128 && t->gemm_tile != 256)
            throw std::invalid_argument("oz2: gemm_tile must be -1, 64, 128 or 256");
        oz2::Tuning u;
        u.batched_gemm = t->batched_gemm != 0;
        u.pad_lda = t->pad_lda != 0;
        u.pow2_tiles = t->pow2_tiles != 0;
        u.gemm_tile = t->gemm_tile;
        h->eng.set_tuning(u);
    });
}

int oz2_get_tuning(oz2_handle h, oz2_tuning_t* out) {
    return guard([&] {
        need(h);
        if (!out) throw std::invalid_argument("oz2: null output pointer");
        const oz2::Tuning& u = h->eng.tuning();
        *out = oz2_tuning_t{u.batched_gemm, u.pad_lda, u.pow2_tiles, u.gemm_tile};
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
        a.wide = tk > oz2::kNarrowScaleBits;
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

const char* oz2_last_error(void) { return g_err.c_str(); }
const char* oz2_version(void) { return "oz2-mi355x 0.1"; }

}  // extern "C"
