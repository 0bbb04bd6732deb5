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
    if (host_) (void)hipHostFree(host_);
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

void Selector::ensure_host(size_t bytes) {
    if (host_size_ >= bytes) return;
    if (host_) (void)hipHostFree(host_);
    host_ = nullptr;
    host_size_ = 0;
    void* p = nullptr;
    hip_check(hipHostMalloc(&p, bytes, hipHostMallocDefault), "selector staging allocation");
    host_ = static_cast<std::byte*>(p);
    host_size_ = bytes;
}

// The problem list goes up and the results come back through pinned staging memory on
// `st`, behind the work already queued there (which may still be writing A and B); one
// synchronization at the end.  Both buffers are idle on entry, since a call that returns
// normally ends with that synchronization; one that threw after queueing work is waited
// for first (pending_).
void Selector::evaluate(hipStream_t st, const std::vector<SelProblem>& probs, std::vector<float>* feat,
                        std::vector<float>* prob) {
    const int n = static_cast<int>(probs.size());
    if (n == 0) return;
    const size_t oP = 0, oS = align_up(sizeof(SelProblem) * n, 256);
    const size_t oF = oS + align_up(k::sel_scratch_bytes(n), 256);
    const size_t oO = oF + align_up(sizeof(float) * kSelFeatures * n, 256);
    const size_t total = oO + align_up(sizeof(float) * kSelOutputs * n, 256);
    const size_t nfeat = static_cast<size_t>(n) * kSelFeatures, nout = static_cast<size_t>(n) * kSelOutputs;
    // host staging: the problems at 0, features at hF, outputs at hO
    const size_t hF = align_up(sizeof(SelProblem) * n, 256), hO = hF + align_up(sizeof(float) * nfeat, 256);
    if (pending_) {  // the stream of that call may be gone: wait for the device
        hip_check(hipDeviceSynchronize(), "device synchronize");
        pending_ = false;
    }
    ensure(total);
    ensure_host(hO + sizeof(float) * nout);
    std::memcpy(host_, probs.data(), sizeof(SelProblem) * n);
    pending_ = true;
    hip_check(hipMemcpyAsync(buf_ + oP, host_, sizeof(SelProblem) * n, hipMemcpyHostToDevice, st), "selector upload");
    float* F = reinterpret_cast<float*>(buf_ + oF);
    k::sel_features(st, reinterpret_cast<const SelProblem*>(buf_ + oP), n, buf_ + oS, F);
    const bool run_mlp = prob && loaded();
    if (run_mlp) {
        if (!dw_) {
            void* p = nullptr;
            hip_check(hipMalloc(&p, sizeof(float) * kSelParams), "selector weights allocation");
            dw_ = static_cast<float*>(p);
        }
        if (dirty_) {
            hip_check(hipMemcpy(dw_, w_.data(), sizeof(float) * kSelParams, hipMemcpyHostToDevice),
                      "selector weights upload");
            dirty_ = false;
        }
        k::mlp_forward(st, F, n, dw_, nullptr, nullptr, nullptr, reinterpret_cast<float*>(buf_ + oO));
    }
    hip_check(hipGetLastError(), "selector launch");
    if (feat)
        hip_check(hipMemcpyAsync(host_ + hF, F, sizeof(float) * nfeat, hipMemcpyDeviceToHost, st), "selector read");
    if (run_mlp)
        hip_check(hipMemcpyAsync(host_ + hO, buf_ + oO, sizeof(float) * nout, hipMemcpyDeviceToHost, st),
                  "selector read");
    hip_check(hipStreamSynchronize(st), "stream synchronize");
    pending_ = false;
    if (feat) {
        const float* f = reinterpret_cast<const float*>(host_ + hF);
        feat->assign(f, f + nfeat);
    }
    if (prob) {
        prob->assign(nout, 0.f);
        if (run_mlp) std::memcpy(prob->data(), host_ + hO, sizeof(float) * nout);
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
