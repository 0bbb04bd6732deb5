// selector_kernels.hip - learned choice of the number of moduli: feature extraction and
// a small MLP (48 -> 64 -> 64 -> 23, FP32) with its forward, backward and Adam
// kernels, all on gfx950.
//
// No counterpart in the reference, whose moduli_for_bits (src/ozaki2_gpu.hpp:
// 693-700) picks s from k alone.  What s a product really needs depends on the
// data: Ozaki II keeps tk bits of every element relative to the largest
// magnitude of its row of op(A) (column of op(B)), so elements 2^g below that
// maximum keep only tk - g bits.  The features describe exactly that: for
// sampled rows of op(A) and columns of op(B), the histogram of the exponent gap
// g between each element and its row maximum, plus log2 k and the bit target.
// Output t of the network is the probability that s = t + 2 moduli reach the
// target, where "b bits" means every entry within 2^(53-b) ulp of the exact
// product.
//
// Feature layout (kSelFeatures = 48), per side X in {A, B} (A at 4, B at 18):
//   [0] log2 k / 16   [1] bits / 64   [2] log2 m / 16   [3] log2 n / 16
//   [X+0 .. X+12] fraction of sampled nonzero elements with gap in
//       {0},{1},{2},{3},[4,6),[6,8),[8,12),[12,16),[16,24),[24,32),[32,48),[48,64),[64,inf)
//   [X+13] fraction of sampled elements that are zero
//   [32+4s .. 35+4s] side s: mean gap / 32, max gap / 64, 90% and 99% gap quantiles / 64
//   [40+2s, 41+2s] side s: (max - min) of the sampled row-maximum exponents / 64, and their
//       mean absolute deviation / 16
//   [44+s] side s: fraction of sampled rows with some gap >= 24
//   [46] log2 k * bits / 1024   [47] 1
#include "grouped.hpp"
#include "selector.hpp"

namespace oz2 {
namespace k {

namespace {

constexpr int kBins = 13;
__device__ __constant__ int kBinLo[kBins + 1] = {0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 1 << 30};

__device__ __forceinline__ int gap_bin(int g) {
    int b = 0;
#pragma unroll
    for (int i = 1; i < kBins; ++i) b += g >= kBinLo[i] ? 1 : 0;
    return b;
}

// per (problem, side): counts[kBins] nonzero by gap bin, zeros, gap sum, max gap, rows with gap >= 24,
// and the row-maximum exponents of the sampled rows
struct SideAcc {
    unsigned long long cnt[kBins];
    unsigned long long zeros;
    unsigned long long gap_sum;
    unsigned int gap_max;
    unsigned int wide_rows;
    int rowexp[kSelRows];
};

// Sampled row r of nr (of `rows`) is row r rows / nr; sampled position j of np = min(k, 2^11) is
// p = j k / np, which is j when np = k and j k / 2^11 when np = 2^11.
static_assert(kSelCols == 1 << 11, "the sampled positions are a shift of j k");
__device__ __forceinline__ int64_t sampled_pos(int j, int64_t k) {
    return k <= kSelCols ? j : (static_cast<int64_t>(j) * k) >> 11;
}

// One block per group of 16 consecutive sampled rows of op(A) or columns of op(B), all
// their sampled positions.  Lanes run along memory: along k for rows contiguous in memory
// (op(A) = T, op(B) = N; thread t: row t >> 5, positions (t & 31) + 32 i), along the rows
// for strided ones (op(A) = N, op(B) = T; row t & 15, positions (t >> 4) + 32 i), where
// sampled rows less than 16 apart share the lines a load fetches.  Pass 1 keeps the
// exponent field of every sampled element (0xFFFF for a zero) in LDS, position-major,
// and finds each row's largest; pass 2 bins the gaps to it.  The block's counts go to
// the (problem, side) accumulator in one set of atomics (one set per row took as long as
// all the loads).
constexpr int kSelGroup = 16;
constexpr int kSelThreads = 512;                    // 32 position lanes per row
constexpr int kSelLanes = kSelThreads / kSelGroup;
__global__ void __launch_bounds__(kSelThreads) sel_stats_kernel(const SelProblem* __restrict__ probs, int nprob,
                                                        SideAcc* __restrict__ acc) {
    constexpr int groups = kSelRows / kSelGroup;
    const int p = blockIdx.x / (2 * groups);
    const int rem = blockIdx.x - p * 2 * groups;
    const int side = rem / groups, g = rem - side * groups;
    if (p >= nprob) return;
    const SelProblem& q = probs[p];
    const int64_t rows = side == 0 ? q.m : q.n;
    const int nr = static_cast<int>(min<int64_t>(rows, kSelRows));
    const int np = static_cast<int>(min<int64_t>(q.k, kSelCols));
    if (g * kSelGroup >= nr || np <= 0) return;
    const bool contig = side == 0 ? q.ta != 0 : q.tb == 0;  // k runs along memory
    const int t = static_cast<int>(threadIdx.x);
    const int rl = contig ? t / kSelLanes : t % kSelGroup, kl = contig ? t % kSelLanes : t / kSelGroup;
    const int r = g * kSelGroup + rl;
    const bool rowok = r < nr;
    const int64_t i = rowok ? static_cast<int64_t>(r) * rows / nr : 0;  // absent rows read row 0, unused
    // element (i, p) of op(A) (side 0) or (p, i) of op(B) (side 1), column-major storage
    const double* X = side == 0 ? q.A : q.B;
    const int64_t ld = side == 0 ? q.lda : q.ldb;
    const int64_t step_k = contig ? 1 : ld, step_i = contig ? ld : 1;
    const double* row = X + i * step_i;
    const int64_t kk = q.k;

    __shared__ uint16_t ex[kSelCols][kSelGroup];  // 64 KB
    __shared__ unsigned int rmax[kSelGroup], rgmax[kSelGroup];
    __shared__ unsigned long long bcnt[kBins + 2];
    if (t < kSelGroup) {
        rmax[t] = 0;
        rgmax[t] = 0;
    }
    if (t < kBins + 2) bcnt[t] = 0;

    // pass 1: exponents into LDS and row maxima; 16 loads in flight per step, all at j < np
    constexpr int U = 16;
    unsigned int emax = 0;
    for (int j0 = kl; j0 < np; j0 += kSelLanes * U) {
        uint64_t v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = min(j0 + kSelLanes * u, np - 1);
            v[u] = static_cast<uint64_t>(__double_as_longlong(row[sampled_pos(j, kk) * step_k])) &
                   0x7FFFFFFFFFFFFFFFull;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int j = j0 + kSelLanes * u;
            if (j < np) {
                const unsigned int e = static_cast<unsigned int>(v[u] >> 52);
                ex[j][rl] = v[u] == 0 ? uint16_t{0xFFFF} : static_cast<uint16_t>(e);
                emax = max(emax, e);
            }
        }
    }
    __syncthreads();
    atomicMax(&rmax[rl], emax);
    __syncthreads();

    // pass 2: gaps against the row maximum (absent rows count nothing)
    const unsigned int e0 = rmax[rl];
    unsigned int gmax = 0;
    unsigned long long gsum = 0;
    unsigned int lc[kBins + 1];
#pragma unroll
    for (int b = 0; b <= kBins; ++b) lc[b] = 0;
    for (int j = kl; rowok && j < np; j += kSelLanes) {
        const unsigned int e = ex[j][rl];
        if (e == 0xFFFF) {
            ++lc[kBins];
            continue;
        }
        const int gp = e >= 0x7FF || e0 >= 0x7FF ? 0 : static_cast<int>(e0 - e);
        gmax = max(gmax, static_cast<unsigned int>(gp));
        gsum += static_cast<unsigned long long>(gp);
        const int bin = gap_bin(gp);
#pragma unroll
        for (int b = 0; b < kBins; ++b) lc[b] += b == bin ? 1u : 0u;
    }
    // the wave's counts summed across its lanes (each at most 64 x 64 elements; the gap sum
    // at most 64 x 64 x 2047 < 2^32), then one LDS atomic per count and wave
    uint32_t gs = static_cast<uint32_t>(gsum);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
#pragma unroll
        for (int b = 0; b <= kBins; ++b) lc[b] += __shfl_xor(lc[b], o, 64);
        gs += __shfl_xor(gs, o, 64);
    }
    if ((t & 63) == 0) {
#pragma unroll
        for (int b = 0; b <= kBins; ++b)
            if (lc[b]) atomicAdd(&bcnt[b], static_cast<unsigned long long>(lc[b]));
        if (gs) atomicAdd(&bcnt[kBins + 1], static_cast<unsigned long long>(gs));
    }
    atomicMax(&rgmax[rl], gmax);
    __syncthreads();
    SideAcc& a = acc[2 * p + side];
    if (t < kSelGroup && g * kSelGroup + t < nr) a.rowexp[g * kSelGroup + t] = static_cast<int>(rmax[t]);
    if (t == 0) {
#pragma unroll
        for (int b = 0; b < kBins; ++b)
            if (bcnt[b]) atomicAdd(&a.cnt[b], bcnt[b]);
        if (bcnt[kBins]) atomicAdd(&a.zeros, bcnt[kBins]);
        if (bcnt[kBins + 1]) atomicAdd(&a.gap_sum, bcnt[kBins + 1]);
        unsigned int gm = 0, wide = 0;
        for (int w = 0; w < kSelGroup && g * kSelGroup + w < nr; ++w) {
            gm = max(gm, rgmax[w]);
            wide += rgmax[w] >= 24 ? 1u : 0u;
        }
        atomicMax(&a.gap_max, gm);
        if (wide) atomicAdd(&a.wide_rows, wide);
    }
}

__global__ void sel_clear_kernel(SideAcc* acc, int n) {
    const int64_t words = static_cast<int64_t>(n) * sizeof(SideAcc) / 4;
    uint32_t* w = reinterpret_cast<uint32_t*>(acc);
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < words;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        w[i] = 0;
}

// quantile of the gap distribution from the histogram: lower edge of the bin holding fraction q
__device__ float hist_quantile(const unsigned long long* cnt, unsigned long long total, float q) {
    if (total == 0) return 0.f;
    const double want = q * static_cast<double>(total);
    double run = 0;
    for (int b = 0; b < kBins; ++b) {
        run += static_cast<double>(cnt[b]);
        if (run >= want) return static_cast<float>(kBinLo[b]);
    }
    return static_cast<float>(kBinLo[kBins - 1]);
}

// One thread per problem: 48 features.
// One 128-thread block per problem: the row exponents of both sides come into LDS with one
// load per thread, then thread `side` reduces its side in row order and writes its
// features (thread 0 also the shape features).
__global__ void __launch_bounds__(kSelRows) sel_finalize_kernel(const SelProblem* __restrict__ probs, int nprob,
                                                                const SideAcc* __restrict__ acc,
                                                                float* __restrict__ feat) {
    const int p = blockIdx.x, t = static_cast<int>(threadIdx.x);
    if (p >= nprob) return;
    const SelProblem& q = probs[p];
    __shared__ int rowexp[2][kSelRows];
    rowexp[0][t] = acc[2 * p].rowexp[t];
    rowexp[1][t] = acc[2 * p + 1].rowexp[t];
    __syncthreads();
    if (t >= 2) return;
    float* f = feat + static_cast<int64_t>(p) * kSelFeatures;
    const float lk = log2f(static_cast<float>(max<int64_t>(q.k, 1)));
    if (t == 0) {
        f[0] = lk / 16.f;
        f[1] = static_cast<float>(q.bits) / 64.f;
        f[2] = log2f(static_cast<float>(max<int64_t>(q.m, 1))) / 16.f;
        f[3] = log2f(static_cast<float>(max<int64_t>(q.n, 1))) / 16.f;
        f[46] = lk * static_cast<float>(q.bits) / 1024.f;
        f[47] = 1.f;
    }
    const int side = t;
    const SideAcc& a = acc[2 * p + side];
    unsigned long long nz = 0;
    for (int b = 0; b < kBins; ++b) nz += a.cnt[b];
    const unsigned long long all = nz + a.zeros;
    float* h = f + 4 + 14 * side;
    for (int b = 0; b < kBins; ++b) h[b] = nz ? static_cast<float>(static_cast<double>(a.cnt[b]) / nz) : 0.f;
    h[kBins] = all ? static_cast<float>(static_cast<double>(a.zeros) / all) : 0.f;
    float* g = f + 32 + 4 * side;
    g[0] = nz ? static_cast<float>(static_cast<double>(a.gap_sum) / nz) / 32.f : 0.f;
    g[1] = static_cast<float>(min(a.gap_max, 128u)) / 64.f;
    g[2] = hist_quantile(a.cnt, nz, 0.90f) / 64.f;
    g[3] = hist_quantile(a.cnt, nz, 0.99f) / 64.f;
    const int64_t rows = side == 0 ? q.m : q.n;
    const int nr = static_cast<int>(min<int64_t>(rows, kSelRows));
    const int* re = rowexp[side];
    int lo = 1 << 30, hi = 0;
    double mean = 0;
    for (int r = 0; r < nr; ++r) {
        lo = min(lo, re[r]);
        hi = max(hi, re[r]);
        mean += re[r];
    }
    mean = nr ? mean / nr : 0;
    double mad = 0;
    for (int r = 0; r < nr; ++r) mad += fabs(re[r] - mean);
    f[40 + 2 * side] = nr ? static_cast<float>(hi - lo) / 64.f : 0.f;
    f[41 + 2 * side] = nr ? static_cast<float>(mad / nr) / 16.f : 0.f;
    f[44 + side] = nr ? static_cast<float>(a.wide_rows) / nr : 0.f;
}

// ---- MLP ------------------------------------------------------------------------
// parameters: W1[64][48] b1[64] W2[64][64] b2[64] W3[23][64] b3[23] (row-major, out x in)
constexpr int H = kSelHidden, F = kSelFeatures, O = kSelOutputs;
constexpr int oW1 = 0, ob1 = oW1 + H * F, oW2 = ob1 + H, ob2 = oW2 + H * H, oW3 = ob2 + H, ob3 = oW3 + O * H;
static_assert(ob3 + O == kSelParams, "parameter layout");

// One 64-thread block per sample: activations H1, H2 (post-ReLU) and logits Z; P = sigmoid(Z).
__global__ void __launch_bounds__(64) mlp_forward_kernel(const float* __restrict__ X, int nb,
                                                         const float* __restrict__ W, float* __restrict__ H1,
                                                         float* __restrict__ H2, float* __restrict__ Z,
                                                         float* __restrict__ P) {
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= nb) return;
    __shared__ float x[F], h1[H], h2[H];
    if (t < F) x[t] = X[static_cast<int64_t>(b) * F + t];
    __syncthreads();
    float a = W[ob1 + t];
    for (int i = 0; i < F; ++i) a = fmaf(W[oW1 + t * F + i], x[i], a);
    h1[t] = fmaxf(a, 0.f);
    __syncthreads();
    a = W[ob2 + t];
    for (int i = 0; i < H; ++i) a = fmaf(W[oW2 + t * H + i], h1[i], a);
    h2[t] = fmaxf(a, 0.f);
    __syncthreads();
    if (H1) H1[static_cast<int64_t>(b) * H + t] = h1[t];
    if (H2) H2[static_cast<int64_t>(b) * H + t] = h2[t];
    if (t < O) {
        a = W[ob3 + t];
        for (int i = 0; i < H; ++i) a = fmaf(W[oW3 + t * H + i], h2[i], a);
        if (Z) Z[static_cast<int64_t>(b) * O + t] = a;
        if (P) P[static_cast<int64_t>(b) * O + t] = 1.f / (1.f + expf(-a));
    }
}

// Loss L = sum over samples and outputs of mask * BCE(sigmoid(z), y) / norm.
// Deltas per sample: D3 = dL/dz, D2 = dL/d(pre-ReLU layer 2), D1 = dL/d(pre-ReLU layer 1).
__global__ void __launch_bounds__(64) mlp_delta_kernel(int nb, const float* __restrict__ W,
                                                       const float* __restrict__ H1, const float* __restrict__ H2,
                                                       const float* __restrict__ Z, const float* __restrict__ Y,
                                                       const float* __restrict__ M, float inv_norm,
                                                       float* __restrict__ D1, float* __restrict__ D2,
                                                       float* __restrict__ D3, float* __restrict__ loss) {
    const int b = blockIdx.x, t = threadIdx.x;
    if (b >= nb) return;
    __shared__ float d3[O], d2[H];
    const int64_t bo = static_cast<int64_t>(b) * O, bh = static_cast<int64_t>(b) * H;
    if (t < O) {
        const float z = Z[bo + t], y = Y[bo + t], m = M ? M[bo + t] : 1.f;
        const float p = 1.f / (1.f + expf(-z));
        d3[t] = m * (p - y) * inv_norm;
        D3[bo + t] = d3[t];
        if (loss) {
            // stable BCE with logits: max(z,0) - z y + log(1 + exp(-|z|))
            loss[bo + t] = m * (fmaxf(z, 0.f) - z * y + log1pf(expf(-fabsf(z)))) * inv_norm;
        }
    }
    __syncthreads();
    float a = 0.f;
    for (int o = 0; o < O; ++o) a = fmaf(W[oW3 + o * H + t], d3[o], a);
    d2[t] = H2[bh + t] > 0.f ? a : 0.f;
    D2[bh + t] = d2[t];
    __syncthreads();
    a = 0.f;
    for (int j = 0; j < H; ++j) a = fmaf(W[oW2 + j * H + t], d2[j], a);
    D1[bh + t] = H1[bh + t] > 0.f ? a : 0.f;
}

// One thread per parameter: gradient summed over the batch in sample order (deterministic).
__global__ void __launch_bounds__(256) mlp_grad_kernel(const float* __restrict__ X, int nb,
                                                       const float* __restrict__ H1, const float* __restrict__ H2,
                                                       const float* __restrict__ D1, const float* __restrict__ D2,
                                                       const float* __restrict__ D3, float* __restrict__ G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kSelParams) return;
    float g = 0.f;
    if (i < ob1) {
        const int o = i / F, c = i % F;
        for (int b = 0; b < nb; ++b) g = fmaf(D1[static_cast<int64_t>(b) * H + o], X[static_cast<int64_t>(b) * F + c], g);
    } else if (i < oW2) {
        const int o = i - ob1;
        for (int b = 0; b < nb; ++b) g += D1[static_cast<int64_t>(b) * H + o];
    } else if (i < ob2) {
        const int o = (i - oW2) / H, c = (i - oW2) % H;
        for (int b = 0; b < nb; ++b) g = fmaf(D2[static_cast<int64_t>(b) * H + o], H1[static_cast<int64_t>(b) * H + c], g);
    } else if (i < oW3) {
        const int o = i - ob2;
        for (int b = 0; b < nb; ++b) g += D2[static_cast<int64_t>(b) * H + o];
    } else if (i < ob3) {
        const int o = (i - oW3) / H, c = (i - oW3) % H;
        for (int b = 0; b < nb; ++b) g = fmaf(D3[static_cast<int64_t>(b) * O + o], H2[static_cast<int64_t>(b) * H + c], g);
    } else {
        const int o = i - ob3;
        for (int b = 0; b < nb; ++b) g += D3[static_cast<int64_t>(b) * O + o];
    }
    G[i] = g;
}

// Adam (Kingma & Ba 2015) with bias correction; step >= 1.
__global__ void __launch_bounds__(256) mlp_adam_kernel(float* __restrict__ W, const float* __restrict__ G,
                                                       float* __restrict__ Mm, float* __restrict__ Vv, int step,
                                                       float lr, float b1, float b2, float eps, float wd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kSelParams) return;
    const float g = G[i] + wd * W[i];
    const float m = b1 * Mm[i] + (1.f - b1) * g;
    const float v = b2 * Vv[i] + (1.f - b2) * g * g;
    Mm[i] = m;
    Vv[i] = v;
    const float mh = m / (1.f - powf(b1, static_cast<float>(step)));
    const float vh = v / (1.f - powf(b2, static_cast<float>(step)));
    W[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

size_t sel_scratch_bytes(int nprob) { return static_cast<size_t>(2 * nprob) * sizeof(SideAcc); }

void sel_features(hipStream_t st, const SelProblem* probs, int nprob, void* scratch, float* feat) {
    if (nprob <= 0) return;
    SideAcc* acc = static_cast<SideAcc*>(scratch);
    hipLaunchKernelGGL(sel_clear_kernel, dim3(64), dim3(256), 0, st, acc, 2 * nprob);
    hipLaunchKernelGGL(sel_stats_kernel, dim3(static_cast<unsigned>(nprob) * 2 * (kSelRows / kSelGroup)),
                       dim3(kSelThreads), 0, st, probs, nprob, acc);
    hipLaunchKernelGGL(sel_finalize_kernel, dim3(static_cast<unsigned>(nprob)), dim3(kSelRows), 0, st, probs, nprob,
                       acc, feat);
}

void mlp_forward(hipStream_t st, const float* X, int nb, const float* W, float* H1, float* H2, float* Z, float* P) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(nb), dim3(64), 0, st, X, nb, W, H1, H2, Z, P);
}

void mlp_backward(hipStream_t st, const float* X, int nb, const float* W, const float* H1, const float* H2,
                  const float* Z, const float* Y, const float* M, float inv_norm, float* D1, float* D2, float* D3,
                  float* G, float* loss) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(mlp_delta_kernel, dim3(nb), dim3(64), 0, st, nb, W, H1, H2, Z, Y, M, inv_norm, D1, D2, D3, loss);
    hipLaunchKernelGGL(mlp_grad_kernel, dim3((kSelParams + 255) / 256), dim3(256), 0, st, X, nb, H1, H2, D1, D2, D3, G);
}

void mlp_adam(hipStream_t st, float* W, const float* G, float* M, float* V, int step, float lr, float b1, float b2,
              float eps, float wd) {
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kSelParams + 255) / 256), dim3(256), 0, st, W, G, M, V, step, lr, b1, b2,
                       eps, wd);
}

}  // namespace k
}  // namespace oz2
