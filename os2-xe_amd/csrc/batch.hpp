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
#include <string>
#include <utility>
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

// schedule of the grouped INT8 GEMM launches (chosen by measurement, see autotune_batch)
struct BatchOptions {
    bool persistent = false;     // 128/64 tiles: fixed grid pulling tiles from a counter, else one workgroup per tile
    bool largest_first = true;   // GEMM work listed by decreasing size, else in problem order
};

class BatchPlan {
public:
    // s[i] > 0 fixes problem i's moduli; the engine supplies tables, stream and device facts.
    // tiles[i] in {256, 128, 64, 256128} forces a macro tile (k::tile_code; 0 or empty:
    // engine heuristic per problem).
    BatchPlan(Engine& eng, std::vector<GemmDesc> descs, const std::vector<int>& s, const std::vector<int>& tiles,
              bool use_graph, BatchOptions opt = {});
    ~BatchPlan();
    BatchPlan(const BatchPlan&) = delete;
    BatchPlan& operator=(const BatchPlan&) = delete;

    void run(hipStream_t st);           // enqueue the whole batch on st
    void run_gemm_only(hipStream_t st); // the grouped INT8 GEMM launches alone (autotuning)
    const std::vector<int>& moduli() const { return s_; }
    const std::vector<int>& tiles() const { return tile_codes_; }  // k::tile_code per problem (0: no GEMM)
    size_t workspace_bytes() const { return ws_size_; }
    int launches() const;               // kernel launches per run()
    bool uses_graph() const { return exec_ != nullptr; }

private:
    void enqueue(hipStream_t st);
    Engine& eng_;
    std::vector<GemmDesc> d_;
    BatchOptions opt_;
    std::vector<int> s_, tile_codes_;
    std::vector<size_t> special_;  // k == 0 or alpha == 0: C = beta C
    std::vector<k::GroupedLaunch> scale_, split_, gemm_, crt_;
    std::byte* ws_ = nullptr;
    size_t ws_size_ = 0;
    uint32_t* counters_ = nullptr;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

// Tile shapes and schedule for one batch, chosen by measuring whole runs of candidate
// plans (the union of all problems fills the GPU, so the per-problem heuristic is not
// the best choice): tile policies tile(min(m, n)) = 256 if >= a, 128 if >= b, else 64,
// each also with 256 x 128 tiles for the problems whose k loop is at most 1024, plus
// the single-problem heuristic and a per-problem choice by padded work / measured tile
// rate at the problem's k, each with and without the persistent schedule; then
// problem order instead of largest first for the winner.
struct TuneResult {
    std::vector<int> tiles;
    BatchOptions opt;
    double ms = 0;
    std::vector<std::pair<std::string, double>> trials;  // candidate -> best ms
};
TuneResult autotune_batch(Engine& eng, const std::vector<GemmDesc>& d, const std::vector<int>& s, int reps = 3);

}  // namespace oz2
