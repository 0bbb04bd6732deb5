// selector.cpp - host side of the learned moduli selector (see selector.hpp).
#include "selector.hpp"

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>

#include "plan.hpp"

namespace oz2 {

namespace {
void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("oz2: ") + what + ": " + hipGetErrorString(e));
}
inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

#include "selector_weights.inc"  // kTrainedWeights[kSelParams] (or empty), kTrainedThreshold
}  // namespace

Selector::Selector() {
    if (sizeof(kTrainedWeights) / sizeof(float) == static_cast<size_t>(kSelParams))
        set(kTrainedWeights, kSelParams, kTrainedThreshold);
}

Selector::~Selector() {
    if (dw_) (void)hipFree(dw_);
    if (buf_) (void)hipFree(buf_);
}

void Selector::set(const float* w, size_t n, float threshold) {
    if (n == 0) {
        w_.clear();
        return;
    }
    if (n != static_cast<size_t>(kSelParams))
        throw std::invalid_argument("oz2: selector weights must have " + std::to_string(kSelParams) + " floats");
    if (!(threshold > 0.f && threshold < 1.f)) throw std::invalid_argument("oz2: selector threshold must be in (0, 1)");
    w_.assign(w, w + n);
    tau_ = threshold;
    dirty_ = true;
}

void Selector::ensure(size_t bytes) {
    if (buf_size_ >= bytes) return;
    if (buf_) (void)hipFree(buf_);
    buf_ = nullptr;
    buf_size_ = 0;
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "selector buffer allocation");
    buf_ = static_cast<std::byte*>(p);
    buf_size_ = bytes;
}

void Selector::evaluate(hipStream_t st, const std::vector<SelProblem>& probs, std::vector<float>* feat,
                        std::vector<float>* prob) {
    const int n = static_cast<int>(probs.size());
    if (n == 0) return;
    const size_t oP = 0, oS = align_up(sizeof(SelProblem) * n, 256);
    const size_t oF = oS + align_up(k::sel_scratch_bytes(n), 256);
    const size_t oO = oF + align_up(sizeof(float) * kSelFeatures * n, 256);
    const size_t total = oO + align_up(sizeof(float) * kSelOutputs * n, 256);
    hip_check(hipStreamSynchronize(st), "stream synchronize");  // buf_ may be in use by earlier work
    ensure(total);
    hip_check(hipMemcpy(buf_ + oP, probs.data(), sizeof(SelProblem) * n, hipMemcpyHostToDevice), "selector upload");
    float* F = reinterpret_cast<float*>(buf_ + oF);
    k::sel_features(st, reinterpret_cast<const SelProblem*>(buf_ + oP), n, buf_ + oS, F);
    if (prob && loaded()) {
        if (!dw_) {
            void* p = nullptr;
            hip_check(hipMalloc(&p, sizeof(float) * kSelParams), "selector weights allocation");
            dw_ = static_cast<float*>(p);
        }
        if (dirty_) {
            hip_check(hipMemcpyAsync(dw_, w_.data(), sizeof(float) * kSelParams, hipMemcpyHostToDevice, st),
                      "selector weights upload");
            dirty_ = false;
        }
        k::mlp_forward(st, F, n, dw_, nullptr, nullptr, nullptr, reinterpret_cast<float*>(buf_ + oO));
    }
    hip_check(hipGetLastError(), "selector launch");
    hip_check(hipStreamSynchronize(st), "stream synchronize");
    if (feat) {
        feat->resize(static_cast<size_t>(n) * kSelFeatures);
        hip_check(hipMemcpy(feat->data(), F, sizeof(float) * feat->size(), hipMemcpyDeviceToHost), "selector read");
    }
    if (prob) {
        prob->assign(static_cast<size_t>(n) * kSelOutputs, 0.f);
        if (loaded())
            hip_check(hipMemcpy(prob->data(), buf_ + oO, sizeof(float) * prob->size(), hipMemcpyDeviceToHost),
                      "selector read");
    }
}

std::vector<int> Selector::choose(hipStream_t st, const std::vector<SelProblem>& probs, std::vector<bool>* learned) {
    const size_t n = probs.size();
    std::vector<int> s(n);
    if (learned) learned->assign(n, false);
    std::vector<float> pr;
    if (loaded()) evaluate(st, probs, nullptr, &pr);
    for (size_t i = 0; i < n; ++i) {
        const int64_t k = std::max<int64_t>(probs[i].k, 1);
        s[i] = moduli_for_bits(k, probs[i].bits);
        if (!loaded()) continue;
        // smallest admissible s the network is confident about; when it is confident about
        // none, the most accurate admissible one
        int pick = 0, largest = 0;
        for (int t = 0; t < kSelOutputs; ++t) {
            const int si = t + kMinModuli;
            if (scale_bits(si, k) < 1) continue;  // s must admit this k
            largest = si;
            if (!pick && pr[i * kSelOutputs + t] >= tau_) pick = si;
        }
        if (largest) {
            s[i] = pick ? pick : largest;
            if (learned) (*learned)[i] = true;
        }
    }
    return s;
}

}  // namespace oz2
