// ozaki2.hpp - drop-in C++ interface of the MI355X Ozaki Scheme II DGEMM/ZGEMM
// library, mirroring the reference's class oz2::Ozaki2
// (src/ozaki2_gpu.hpp:26-52 types, 678-786 members).  Header-only over the C
// ABI in oz2_c.h (link with liboz2.so).  A sycl::queue becomes a hipStream_t;
// the returned sycl::event becomes oz2::Event (a hipEvent_t with wait()).
#pragma once
#include <hip/hip_runtime_api.h>

#include <complex>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "oz2_c.h"

namespace oz2 {

inline constexpr int kMaxModuli = 24;
inline constexpr int kMaxLimbs = 6;
inline constexpr int kMaxScaleBits = 100;
inline constexpr int kNarrowScaleBits = 62;
inline constexpr int kNonFinite = std::numeric_limits<int>::min();
// Longest k-chunk of one INT8 GEMM pass in this library (the reference's kMaxKChunk is 131071,
// hpp:32): chunks of at most 65408 keep every int32 sum below 2^30, and the chunks of one product
// have equal length, so last_tiling().k_chunks = ceil(round_up(k, 128) / 65408).
inline constexpr int64_t kMaxKChunk = 65408;

enum class Trans { N, T, C };

struct StageTimes {
    double scale_ms = 0, split_ms = 0, gemm_ms = 0, crt_ms = 0;
    double total_ms() const { return scale_ms + split_ms + gemm_ms + crt_ms; }
};

struct Tuning {
    bool batched_gemm = true;
    bool pad_lda = true;
    bool pow2_tiles = true;
    int gemm_tile = -1;  // MI355X extension: -1 auto, 256 / 128 / 64 / 256128 (256 x 128) forces the INT8 GEMM macro tile
    int overlap_blocks = -1;  // MI355X extension: column blocks overlapped over two streams, -1 auto, 0 or 1 none
};

struct TileInfo {
    int64_t mb = 0, nb = 0;
    int64_t pm = 0, pn = 0;
    int64_t k_chunks = 0;
    size_t workspace_bytes = 0;
};

namespace detail {
[[noreturn]] inline void raise(int rc) {
    const std::string msg = oz2_last_error();
    switch (rc) {
        case OZ2_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case OZ2_LOGIC_ERROR: throw std::logic_error(msg);
        case OZ2_OVERFLOW_ERROR: throw std::overflow_error(msg);
        default: throw std::runtime_error(msg);
    }
}
inline void check(int rc) {
    if (rc != OZ2_OK) raise(rc);
}
inline int code(Trans t) { return t == Trans::N ? OZ2_TRANS_N : t == Trans::T ? OZ2_TRANS_T : OZ2_TRANS_C; }
}  // namespace detail

// Completion of one gemm/zgemm call (the reference returns a sycl::event).
class Event {
public:
    Event() = default;
    void wait() const {
        if (ev_ && hipEventSynchronize(*ev_) != hipSuccess) throw std::runtime_error("oz2: hipEventSynchronize failed");
    }
    hipEvent_t native() const { return ev_ ? *ev_ : nullptr; }
    static Event make() {
        Event e;
        hipEvent_t h;
        if (hipEventCreateWithFlags(&h, hipEventDisableTiming) != hipSuccess)
            throw std::runtime_error("oz2: hipEventCreate failed");
        e.ev_ = std::shared_ptr<hipEvent_t>(new hipEvent_t(h), [](hipEvent_t* p) {
            (void)hipEventDestroy(*p);
            delete p;
        });
        return e;
    }

private:
    std::shared_ptr<hipEvent_t> ev_;
};

class Ozaki2 {
public:
    explicit Ozaki2(hipStream_t stream = nullptr) { detail::check(oz2_create(stream, &h_)); }
    ~Ozaki2() {
        if (h_) oz2_destroy(h_);
    }
    Ozaki2(const Ozaki2&) = delete;
    Ozaki2& operator=(const Ozaki2&) = delete;

    static int precision_bits(int64_t k, int num_moduli) {
        int err = 0;
        const int tk = oz2_precision_bits(k, num_moduli, &err);
        detail::check(err);
        return tk;
    }
    static int precision_bits_complex(int64_t k, int num_moduli) {
        int err = 0;
        const int tk = oz2_precision_bits_complex(k, num_moduli, &err);
        detail::check(err);
        return tk;
    }
    static int moduli_for_bits(int64_t k, int bits) { return oz2_moduli_for_bits(k, bits); }

    void set_workspace_limit(size_t bytes) { detail::check(oz2_set_workspace_limit(h_, bytes)); }
    void set_resident_hint(size_t bytes) { detail::check(oz2_set_resident_hint(h_, bytes)); }
    void release_workspace() { detail::check(oz2_release_workspace(h_)); }

    const TileInfo& last_tiling() {
        oz2_tile_info_t t;
        detail::check(oz2_last_tiling(h_, &t));
        tiling_ = TileInfo{t.mb, t.nb, t.pm, t.pn, t.k_chunks, t.workspace_bytes};
        return tiling_;
    }
    void set_tuning(const Tuning& t) {
        const oz2_tuning_t c{t.batched_gemm, t.pad_lda, t.pow2_tiles, t.gemm_tile, t.overlap_blocks};
        detail::check(oz2_set_tuning(h_, &c));
        tune_ = t;
    }
    const Tuning& tuning() const { return tune_; }

    StageTimes stage_times() {
        oz2_stage_times_t t;
        detail::check(oz2_stage_times(h_, &t));
        return StageTimes{t.scale_ms, t.split_ms, t.gemm_ms, t.crt_ms};
    }

    Event gemm(Trans ta, Trans tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int num_moduli,
               const std::vector<Event>& deps = {}) {
        std::vector<void*> d;
        for (const auto& e : deps) d.push_back(e.native());
        Event done = Event::make();
        detail::check(oz2_dgemm(h_, detail::code(ta), detail::code(tb), m, n, k, alpha, A, lda, B, ldb, beta, C, ldc,
                                num_moduli, d.data(), static_cast<int>(d.size()), done.native()));
        return done;
    }

    Event zgemm(Trans ta, Trans tb, int64_t m, int64_t n, int64_t k, std::complex<double> alpha,
                const std::complex<double>* A, int64_t lda, const std::complex<double>* B, int64_t ldb,
                std::complex<double> beta, std::complex<double>* C, int64_t ldc, int num_moduli,
                const std::vector<Event>& deps = {}) {
        std::vector<void*> d;
        for (const auto& e : deps) d.push_back(e.native());
        Event done = Event::make();
        const double al[2] = {alpha.real(), alpha.imag()}, be[2] = {beta.real(), beta.imag()};
        detail::check(oz2_zgemm(h_, detail::code(ta), detail::code(tb), m, n, k, al,
                                reinterpret_cast<const double*>(A), lda, reinterpret_cast<const double*>(B), ldb, be,
                                reinterpret_cast<double*>(C), ldc, num_moduli, d.data(), static_cast<int>(d.size()),
                                done.native()));
        return done;
    }

    oz2_handle native_handle() const { return h_; }

    // New in this library (no reference counterpart): the learned choice of num_moduli for one
    // problem and a bit target (moduli_for_bits when no selector weights are loaded).
    int select_moduli(Trans ta, Trans tb, int64_t m, int64_t n, int64_t k, const double* A, int64_t lda,
                      const double* B, int64_t ldb, int bits = 53) {
        oz2_gemm_desc_t d{detail::code(ta), detail::code(tb), m, n, k, 1.0, A, lda, B, ldb, 0.0, nullptr, 1, 0};
        int s = 0;
        detail::check(oz2_selector_choose(h_, 1, &d, bits, nullptr, &s, nullptr));
        return s;
    }

private:
    oz2_handle h_ = nullptr;
    Tuning tune_{};
    TileInfo tiling_{};
};

// The fastest plan oz2_batch_autotune measured for a list of problems: a macro tile per problem
// and the schedule bits (OZ2_BATCH_PERSISTENT, OZ2_BATCH_PROBLEM_ORDER) to build it with.
struct BatchTuning {
    std::vector<int> tiles;
    int schedule = 0;
    double ms = 0;
};

// A plan for many DGEMMs (oz2_batch_*): one grouped launch sequence, optionally replayed from a
// hipGraph; every problem's C equals Ozaki2::gemm with the same num_moduli.  Descriptors with
// num_moduli <= 0 get the selector's choice (select = true) or moduli_for_bits(k, bits).
// schedule: OZ2_BATCH_PERSISTENT and/or OZ2_BATCH_PROBLEM_ORDER (as autotune() returns them).
class Batch {
public:
    Batch(Ozaki2& oz, const std::vector<oz2_gemm_desc_t>& descs, int bits = 53, bool graph = true,
          bool select = false, const std::vector<int>& tiles = {}, int schedule = 0) {
        const int flags = (graph ? OZ2_BATCH_GRAPH : 0) | (select ? OZ2_BATCH_SELECT : 0) |
                          (schedule & (OZ2_BATCH_PERSISTENT | OZ2_BATCH_PROBLEM_ORDER));
        detail::check(oz2_batch_create(oz.native_handle(), static_cast<int>(descs.size()), descs.data(), bits,
                                       tiles.empty() ? nullptr : tiles.data(), flags, &b_));
        n_ = static_cast<int>(descs.size());
    }
    // measures candidate plans for these problems (runs them, overwriting C) and returns the fastest
    static BatchTuning autotune(Ozaki2& oz, const std::vector<oz2_gemm_desc_t>& descs, int bits = 53,
                                bool select = false) {
        BatchTuning t;
        t.tiles.assign(descs.size(), 0);
        detail::check(oz2_batch_autotune(oz.native_handle(), static_cast<int>(descs.size()), descs.data(), bits,
                                         select ? OZ2_BATCH_SELECT : 0, t.tiles.data(), &t.schedule, &t.ms));
        return t;
    }
    ~Batch() {
        if (b_) (void)oz2_batch_destroy(b_);
    }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void run() { detail::check(oz2_batch_run(b_)); }
    // choose the moduli again from the buffers' current contents (oz2_batch_reselect); returns how
    // many problems changed
    int reselect() {
        int changed = 0;
        detail::check(oz2_batch_reselect(b_, &changed));
        return changed;
    }
    std::vector<int> moduli() const {
        std::vector<int> s(n_);
        detail::check(oz2_batch_moduli(b_, s.data()));
        return s;
    }

private:
    oz2_batch b_ = nullptr;
    int n_ = 0;
};

// Ozaki Scheme I, the reference's comparison baseline (oz1::Ozaki1, benchmarks/ozaki1_gpu.hpp:86-155):
// C = A B (op N, alpha 1, beta 0) with L slices, on an Ozaki2 handle's stream and workspace.
class Ozaki1 {
public:
    explicit Ozaki1(Ozaki2& oz) : oz_(oz) {}
    static int precision_bits(int L) { return 7 * L - 2; }
    static int num_gemms(int L) { return L * (L + 1) / 2; }
    void gemm(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
              int64_t ldc, int L) {
        detail::check(oz2_ozaki1_dgemm(oz_.native_handle(), m, n, k, A, lda, B, ldb, C, ldc, L));
    }

private:
    Ozaki2& oz_;
};

}  // namespace oz2
