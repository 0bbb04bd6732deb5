// zgemm.cpp - C = alpha * A^H * B + beta * C with oz2::Ozaki2 on one MI355X: the
// calls a user of the reference's complex example makes (examples/zgemm.cpp:49:
// precision_bits_complex, zgemm(...).wait()), with hipStream_t in place of
// sycl::queue.  The result is compared with a plain FP64 complex product on the
// host (all entries up to 2^27 multiply-adds, else 4096 sampled entries).
//   zgemm [m [n [k [s]]]] [--dump DIR]: DIR receives A, B, C0 and C as raw
//   interleaved complex doubles (A: k x m, B: k x n, C: m x n, column-major).
// Build: make -C os2-xe_amd examples
#include <hip/hip_runtime.h>

#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include <oz2/ozaki2.hpp>

using cd = std::complex<double>;

static bool dump(const std::string& dir, const char* name, const std::vector<cd>& v) {
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(v.data(), sizeof(cd), v.size(), f) == v.size();
    return std::fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
    std::vector<const char*> pos;
    std::string dir;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dir = argv[++i];
        else pos.push_back(argv[i]);
    }
    const int64_t m = pos.size() > 0 ? std::atoll(pos[0]) : 1024;
    const int64_t n = pos.size() > 1 ? std::atoll(pos[1]) : m;
    const int64_t k = pos.size() > 2 ? std::atoll(pos[2]) : m;
    const int s = pos.size() > 3 ? std::atoi(pos[3]) : 16;

    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    std::printf("device: %s\n", prop.name);

    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<cd> hA(k * m), hB(k * n), hC(m * n);
    for (auto& v : hA) v = {U(rng), U(rng)};
    for (auto& v : hB) v = {U(rng), U(rng)};
    for (auto& v : hC) v = {U(rng), U(rng)};
    const std::vector<cd> hC0 = hC;

    cd *dA, *dB, *dC;
    if (hipMalloc(&dA, hA.size() * sizeof(cd)) || hipMalloc(&dB, hB.size() * sizeof(cd)) ||
        hipMalloc(&dC, hC.size() * sizeof(cd)))
        return 1;
    if (hipMemcpy(dA, hA.data(), hA.size() * sizeof(cd), hipMemcpyHostToDevice) ||
        hipMemcpy(dB, hB.data(), hB.size() * sizeof(cd), hipMemcpyHostToDevice) ||
        hipMemcpy(dC, hC.data(), hC.size() * sizeof(cd), hipMemcpyHostToDevice))
        return 1;

    const cd alpha(1.0, 0.5), beta(0.25, -1.0);
    hipStream_t stream;
    if (hipStreamCreate(&stream)) return 1;
    {
        oz2::Ozaki2 oz(stream);
        std::printf("m=%lld n=%lld k=%lld, moduli s=%d, bits per element tk=%d\n", (long long)m, (long long)n,
                    (long long)k, s, oz2::Ozaki2::precision_bits_complex(k, s));
        oz.zgemm(oz2::Trans::C, oz2::Trans::N, m, n, k, alpha, dA, k, dB, k, beta, dC, m, s).wait();
    }
    if (hipMemcpy(hC.data(), dC, hC.size() * sizeof(cd), hipMemcpyDeviceToHost)) return 1;

    // plain FP64 reference: alpha * sum conj(A[p, i]) B[p, j] + beta * C0
    const bool all = static_cast<double>(m) * n * k <= double(1 << 27);
    std::mt19937_64 pick(11);
    double max_diff = 0, max_ref = 0;
    const int64_t count = all ? m * n : 4096;
    for (int64_t e = 0; e < count; ++e) {
        const int64_t i = all ? e % m : static_cast<int64_t>(pick() % m);
        const int64_t j = all ? e / m : static_cast<int64_t>(pick() % n);
        cd acc = 0;
        for (int64_t p = 0; p < k; ++p) acc += std::conj(hA[p + i * k]) * hB[p + j * k];
        const cd ref = alpha * acc + beta * hC0[i + j * m];
        max_diff = std::max(max_diff, std::abs(hC[i + j * m] - ref));
        max_ref = std::max(max_ref, std::abs(ref));
    }
    std::printf("max |C_oz2 - C_fp64| / max |C_fp64| = %.3e over %lld entries\n", max_diff / max_ref,
                (long long)count);

    if (!dir.empty() && !(dump(dir, "A.bin", hA) && dump(dir, "B.bin", hB) && dump(dir, "C0.bin", hC0) &&
                          dump(dir, "C.bin", hC)))
        return 1;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipStreamDestroy(stream);
    return 0;
}
