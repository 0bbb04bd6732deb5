// selector.hpp - learned moduli selector: data-dependent features, a 48-64-64-23
// MLP run and trained with hand-written gfx950 kernels, and the host rule that
// turns its outputs into a number of moduli (moduli_for_bits stays the fallback).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace oz2 {

constexpr int kSelFeatures = 48;
constexpr int kSelHidden = 64;
constexpr int kSelOutputs = 23;  // s = 2 .. 24
constexpr int kSelParams = kSelHidden * kSelFeatures + kSelHidden + kSelHidden * kSelHidden + kSelHidden +
                           kSelOutputs * kSelHidden + kSelOutputs;
constexpr int kSelRows = 128;   // sampled rows of op(A) / columns of op(B)
constexpr int kSelCols = 2048;  // sampled k positions per row

// one problem as the feature kernels read it (ta, tb: 0 = N, else T; column-major device pointers)
struct SelProblem {
    int ta, tb;
    int64_t m, n, k;
    const double* A;
    int64_t lda;
    const double* B;
    int64_t ldb;
    int bits;
};

namespace k {
size_t sel_scratch_bytes(int nprob);
// feat: nprob x kSelFeatures floats (device); probs: device array of nprob SelProblem
void sel_features(hipStream_t st, const SelProblem* probs, int nprob, void* scratch, float* feat);
// X: nb x F; H1, H2: nb x 64 (may be null in inference); Z logits, P probabilities: nb x 23 (either may be null)
void mlp_forward(hipStream_t st, const float* X, int nb, const float* W, float* H1, float* H2, float* Z, float* P);
// Y labels and M mask (null = all ones): nb x 23; loss per (sample, output) if non-null; G: kSelParams
void mlp_backward(hipStream_t st, const float* X, int nb, const float* W, const float* H1, const float* H2,
                  const float* Z, const float* Y, const float* M, float inv_norm, float* D1, float* D2, float* D3,
                  float* G, float* loss);
void mlp_adam(hipStream_t st, float* W, const float* G, float* M, float* V, int step, float lr, float b1, float b2,
              float eps, float wd);
}  // namespace k

// Host side: weights, threshold and the choice rule.
class Selector {
public:
    Selector();  // the weights trained for this library (selector_weights.inc), if any
    ~Selector();
    Selector(const Selector&) = delete;
    Selector& operator=(const Selector&) = delete;

    bool loaded() const { return !w_.empty(); }
    void set(const float* w, size_t n, float threshold);
    const std::vector<float>& weights() const { return w_; }
    float threshold() const { return tau_; }

    // features (nprob x 48) and output probabilities (nprob x 23) on the host; synchronizes `st`
    void evaluate(hipStream_t st, const std::vector<SelProblem>& probs, std::vector<float>* feat,
                  std::vector<float>* prob);
    // smallest admissible s whose probability reaches the threshold, else the largest
    // admissible s; moduli_for_bits(k, bits) when no weights are loaded
    std::vector<int> choose(hipStream_t st, const std::vector<SelProblem>& probs, std::vector<bool>* learned = nullptr);

private:
    void ensure(size_t bytes);
    void ensure_host(size_t bytes);
    std::vector<float> w_;
    float tau_ = 0.9f;
    float* dw_ = nullptr;  // device copy of the weights
    bool dirty_ = true;
    std::byte* buf_ = nullptr;
    size_t buf_size_ = 0;
    std::byte* host_ = nullptr;  // pinned staging for the problem list and the results
    size_t host_size_ = 0;
    bool pending_ = false;       // a call queued work and did not wait for it (it threw)
};

}  // namespace oz2
