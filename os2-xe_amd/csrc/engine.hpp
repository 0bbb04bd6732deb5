// engine.hpp - host orchestrator of the Ozaki Scheme II DGEMM/ZGEMM path.
//
// Counterpart of the reference's class oz2::Ozaki2 (src/ozaki2_gpu.hpp:678-1188):
// argument validation, plan cache, tiling / k-chunking, grow-only workspace,
// stage timing, and the launch sequence scale -> split -> INT8 GEMM -> CRT on
// one HIP stream.  All work is asynchronous on the caller's stream; the host
// blocks only to grow the workspace and in stage_times().
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "kernels.hpp"
#include "plan.hpp"
#include "planner.hpp"

namespace oz2 {

struct StageTimes {
    double scale_ms = 0, split_ms = 0, gemm_ms = 0, crt_ms = 0;
};

struct Tuning {
    bool batched_gemm = true;  // one GEMM launch for all planes; false: one launch per plane
    bool pad_lda = true;       // stagger plane strides that are multiples of 64 KiB by 4 KiB (engine.cpp)
    bool pow2_tiles = true;    // include power-of-two tile candidates
    int gemm_tile = -1;        // -1 auto, else a tile code: 256 / 128 / 64 (square) or 256128 (256 x 128)
    int overlap_blocks = -1;   // column blocks per panel pair overlapped over two streams: -1 auto, 0 or 1 none
};
inline constexpr int kMaxOverlapBlocks = 16;

enum class Op { N = 0, T = 1, C = 2 };

class Engine {
public:
    explicit Engine(hipStream_t stream);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    void gemm(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A, int64_t lda,
              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int s);
    // complex: interleaved (re, im) doubles; alpha/beta as (re, im)
    void zgemm(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double ar, double ai, const double* A, int64_t lda,
               const double* B, int64_t ldb, double br, double bi, double* C, int64_t ldc, int s);

    void set_workspace_limit(size_t b) { ws_limit_ = b; }
    void set_resident_hint(size_t b) { resident_hint_ = b; }
    void release_workspace();
    void reserve_workspace(size_t bytes);
    size_t workspace_size() const { return ws_size_; }
    const TileInfo& last_tiling() const { return tiling_; }
    void set_tuning(const Tuning& t) { tune_ = t; }
    const Tuning& tuning() const { return tune_; }
    void set_profiling(bool on) { profiling_ = on; }
    bool profiling() const { return profiling_; }
    StageTimes stage_times();
    // stage times summed over every call between profile_begin() and profile_end()
    void profile_begin();
    StageTimes profile_end(int64_t* calls);
    hipStream_t stream() const { return stream_; }
    // later work goes to `s`, ordered after everything already queued on the old stream
    void set_stream(hipStream_t s);
    int device() const { return dev_; }

    // Workspace a call would need (no allocation); used to pre-size before graph capture.
    size_t workspace_for(Op ta, Op tb, int64_t m, int64_t n, int64_t k, int s, bool cplx);
    // gemm() with all intermediates in the caller's device buffer ws of ws_bytes: the panels
    // are planned to fit it (std::invalid_argument if even the smallest do not), nothing is
    // allocated or created, and no stage times are recorded
    void gemm_fixed(std::byte* ws, size_t ws_bytes, Op ta, Op tb, int64_t m, int64_t n, int64_t k, double alpha,
                    const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc,
                    int s);

    // Ozaki Scheme I with L slices (the reference's comparison baseline oz1::Ozaki1::gemm,
    // benchmarks/ozaki1_gpu.hpp:100-151): C = A B, op N, alpha 1, beta 0, on this library's kernels
    void ozaki1(int64_t m, int64_t n, int64_t k, const double* A, int64_t lda, const double* B, int64_t ldb, double* C,
                int64_t ldc, int L);

    // shared with batch plans
    const TabSet* device_tables() const { return dtabs_; }
    const Plan& plan_for(int s, int64_t k);
    k::Tile pick_tile(int64_t mb, int64_t nb, int64_t klen, int planes) const;
    size_t budget() const;
    int compute_units() const { return cus_; }

private:
    enum class Stage { Scale, Split, Gemm, Crt };
    struct Mark {
        Stage st;
        size_t ev;  // events_[ev], events_[ev + 1]: begin and end on the stream the stage ran on
    };
    void run(Op ta, Op tb, int64_t m, int64_t n, int64_t k, double ar, double ai, const double* A, int64_t lda,
             const double* B, int64_t ldb, double br, double bi, double* C, int64_t ldc, int s, bool cplx,
             std::byte* fixed_ws, size_t fixed_bytes);
    hipEvent_t timing_event(size_t idx);
    size_t stage_begin(Stage st, hipStream_t on);
    void stage_end(size_t mark, hipStream_t on);
    hipEvent_t sync_event();  // from a pool reused by every call (a wait binds the record before it)
    void order_after(hipStream_t to, hipStream_t from);
    int overlap_blocks() const;  // column blocks per panel pair (Tuning::overlap_blocks; automatic: 1)
    CallShape call_shape(Op ta, Op tb, int64_t m, int64_t n, int64_t k, int s, bool cplx) const;
    // plan_tiling, remembered for the last shape and limit
    TileInfo plan_for_call(const CallShape& c, size_t limit);
    void ensure_workspace(size_t bytes);

    hipStream_t stream_;
    int dev_ = 0;
    hipEvent_t switch_ev_ = nullptr;  // orders a new stream after the old one (set_stream)
    TabSet* dtabs_ = nullptr;  // device copy of make_tabsets(), s = 0..24
    Tuning tune_{};
    bool profiling_ = true;
    bool have_plan_ = false;
    Plan plan_{};
    std::byte* ws_ = nullptr;
    size_t ws_size_ = 0;
    size_t ws_limit_ = 0;
    size_t resident_hint_ = 0;
    TileInfo tiling_{};
    bool have_memo_ = false;
    CallShape memo_shape_{};
    size_t memo_limit_ = 0;
    TileInfo memo_ti_{};
    k::Tile tile_ = k::Tile::T256;
    hipStream_t aux_ = nullptr;  // second stream of the column-block overlap
    std::vector<hipEvent_t> events_;        // timing events
    std::vector<Mark> marks_;
    std::vector<hipEvent_t> sync_events_;   // ordering events between the two streams
    size_t sync_used_ = 0;
    bool accumulate_ = false;
    int64_t calls_ = 0;
    int cus_ = 256;
};

}  // namespace oz2
