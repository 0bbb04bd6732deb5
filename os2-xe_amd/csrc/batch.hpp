// batch.hpp - many DGEMMs of different shapes as one grouped launch sequence.
//
// No counterpart in the reference (it runs one GEMM per call): a BatchPlan
// takes a list of DGEMMs (each with the full gemm() semantics: op, alpha,
// beta, leading dimensions), gives every problem its own number of moduli,
// lays out one workspace for all of them and prepares grouped launches:
// scale -> split -> persistent INT8 GEMM (one per macro-tile shape) -> CRT,
// about ten launches for any number of problems.  run() replays that sequence
// from a captured hipGraph (or launches it directly).  Each problem's result is
// bit-identical to Ozaki2::gemm with the same s: the grouped kernels run the
// same per-block code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "engine.hpp"

namespace oz2 {

struct GemmDesc {
    Op ta = Op::N, tb = Op::N;
    int64_t m = 0, n = 0, k = 0;
    double alpha = 1.0;
    const double* A = nullptr;
    int64_t lda = 1;
    const double* B = nullptr;
    int64_t ldb = 1;
    double beta = 0.0;
    double* C = nullptr;
    int64_t ldc = 1;
    int s = 0;  // number of moduli; <= 0: chosen by the plan (selector, else moduli_for_bits)
};

class BatchPlan {
public:
    // s[i] > 0 fixes problem i's moduli; the engine supplies tables, stream and device facts.
    // tiles[i] in {256, 128, 64} forces a macro tile (empty: engine heuristic per problem).
    BatchPlan(Engine& eng, std::vector<GemmDesc> descs, const std::vector<int>& s, const std::vector<int>& tiles,
              bool use_graph);
    ~BatchPlan();
    BatchPlan(const BatchPlan&) = delete;
    BatchPlan& operator=(const BatchPlan&) = delete;

    void run(hipStream_t st);           // enqueue the whole batch on st
    void run_gemm_only(hipStream_t st); // the grouped INT8 GEMM launches alone (autotuning)
    const std::vector<int>& moduli() const { return s_; }
    const std::vector<int>& tiles() const { return tile_rows_; }
    size_t workspace_bytes() const { return ws_size_; }
    int launches() const;               // kernel launches per run()
    bool uses_graph() const { return exec_ != nullptr; }

private:
    void enqueue(hipStream_t st);
    Engine& eng_;
    std::vector<GemmDesc> d_;
    std::vector<int> s_, tile_rows_;
    std::vector<size_t> special_;  // k == 0 or alpha == 0: C = beta C
    std::vector<k::GroupedLaunch> scale_, split_, gemm_, crt_;
    std::byte* ws_ = nullptr;
    size_t ws_size_ = 0;
    uint32_t* counters_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

}  // namespace oz2
