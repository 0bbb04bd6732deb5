// dgemm.cpp - C = A*B with oz2::Ozaki2 on one MI355X, the same calls a user of
// the reference makes (moduli_for_bits, set_resident_hint, gemm(...).wait(),
// stage_times, last_tiling), with hipStream_t in place of sycl::queue.
// Build: hipcc -std=c++17 -I../../include dgemm.cpp -L../lib -loz2 -Wl,-rpath,../lib
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include <oz2/ozaki2.hpp>

int main(int argc, char** argv) {
    const int64_t m = argc > 1 ? std::atoll(argv[1]) : 2048;
    const int64_t n = argc > 2 ? std::atoll(argv[2]) : m;
    const int64_t k = argc > 3 ? std::atoll(argv[3]) : m;

    std::mt19937_64 rng(42);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    std::vector<double> hA(m * k), hB(k * n), hC(m * n);
    for (auto& v : hA) v = U(rng);
    for (auto& v : hB) v = U(rng);

    double *dA, *dB, *dC;
    if (hipMalloc(&dA, hA.size() * 8) || hipMalloc(&dB, hB.size() * 8) || hipMalloc(&dC, hC.size() * 8)) return 1;
    if (hipMemcpy(dA, hA.data(), hA.size() * 8, hipMemcpyHostToDevice) ||
        hipMemcpy(dB, hB.data(), hB.size() * 8, hipMemcpyHostToDevice))
        return 1;

    hipStream_t stream;
    if (hipStreamCreate(&stream)) return 1;
    oz2::Ozaki2 oz(stream);
    oz.set_resident_hint(sizeof(double) * (m * k + k * n + m * n));

    const int s = oz2::Ozaki2::moduli_for_bits(k, 53);
    std::printf("m=%lld n=%lld k=%lld, moduli s=%d, bits per element tk=%d\n", (long long)m, (long long)n,
                (long long)k, s, oz2::Ozaki2::precision_bits(k, s));
    oz.gemm(oz2::Trans::N, oz2::Trans::N, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s).wait();

    const oz2::StageTimes st = oz.stage_times();
    const oz2::TileInfo& ti = oz.last_tiling();
    std::printf("scale %.3f split %.3f gemm %.3f crt %.3f ms, tiles %lldx%lld\n", st.scale_ms, st.split_ms,
                st.gemm_ms, st.crt_ms, (long long)ti.pm, (long long)ti.pn);

    if (hipMemcpy(hC.data(), dC, hC.size() * 8, hipMemcpyDeviceToHost)) return 1;
    std::printf("C[0] = %.17g\n", hC[0]);

    // the same product and one over the first half of k as a batch: num_moduli = 0 lets the
    // learned selector choose from the data, the autotuner picks the tiles and the GEMM
    // schedule, and the plan is replayed from a hipGraph
    double* dC2 = nullptr;
    if (hipMalloc(&dC2, hC.size() * 8)) return 1;
    const std::vector<oz2_gemm_desc_t> d = {
        {OZ2_TRANS_N, OZ2_TRANS_N, m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, 0},
        {OZ2_TRANS_N, OZ2_TRANS_N, m, n, k / 2, 1.0, dA, m, dB, k, 0.0, dC2, m, 0}};
    const oz2::BatchTuning tune = oz2::Batch::autotune(oz, d, 53, /*select=*/true);
    oz2::Batch batch(oz, d, 53, /*graph=*/true, /*select=*/true, tune.tiles, tune.schedule);
    batch.run();
    if (hipStreamSynchronize(stream)) return 1;
    std::printf("batch: learned s = %d, %d (moduli_for_bits: %d); tiles %d, %d, %s schedule, %.3f ms\n",
                batch.moduli()[0], batch.moduli()[1], s, tune.tiles[0], tune.tiles[1],
                (tune.schedule & OZ2_BATCH_PERSISTENT) ? "persistent" : "one-tile-per-workgroup", tune.ms);
    (void)hipFree(dC2);

    // the reference's comparison baseline, Ozaki Scheme I with 9 slices (45 INT8 GEMMs, 61 bits)
    oz2::Ozaki1 oz1(oz);
    oz1.gemm(m, n, k, dA, m, dB, k, dC, m, 9);
    if (hipStreamSynchronize(stream)) return 1;
    std::printf("ozaki1: L=9, %d INT8 GEMMs, %d bits per element\n", oz2::Ozaki1::num_gemms(9),
                oz2::Ozaki1::precision_bits(9));
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipStreamDestroy(stream);
    return 0;
}
