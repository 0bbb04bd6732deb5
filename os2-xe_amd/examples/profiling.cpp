// profiling.cpp - accuracy/speed trade-off: one DGEMM N x N x N for several numbers
// of moduli, with the per-stage times, the tiling and the achieved TFLOP/s: the
// calls of the reference's profiling example (examples/profiling.cpp:30-42:
// set_resident_hint, gemm(...).wait(), stage_times, last_tiling), with
// hipStream_t in place of a profiling sycl::queue.
//   profiling [N] [--dump DIR]: DIR receives A, B (sin(i), cos(i) of the linear
//   index, computed on the GPU) and C of the last s as raw doubles.
// Build: make -C os2-xe_amd examples
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <oz2/ozaki2.hpp>

__global__ void fill_sincos(double* A, double* B, int64_t count) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < count) {
        A[i] = sin(static_cast<double>(i));
        B[i] = cos(static_cast<double>(i));
    }
}

static bool dump(const std::string& dir, const char* name, const double* d, size_t count) {
    std::vector<double> h(count);
    if (hipMemcpy(h.data(), d, count * 8, hipMemcpyDeviceToHost) != hipSuccess) return false;
    FILE* f = std::fopen((dir + "/" + name).c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(h.data(), 8, count, f) == count;
    return std::fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
    int64_t N = 8192;
    std::string dir;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--dump") && i + 1 < argc) dir = argv[++i];
        else N = std::atoll(argv[i]);
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 1;
    std::printf("device: %s, N=%lld\n", prop.name, (long long)N);

    const int64_t count = N * N;
    double *dA, *dB, *dC;
    if (hipMalloc(&dA, count * 8) || hipMalloc(&dB, count * 8) || hipMalloc(&dC, count * 8)) return 1;
    hipStream_t stream;
    if (hipStreamCreate(&stream)) return 1;
    hipLaunchKernelGGL(fill_sincos, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, stream, dA, dB,
                       count);
    if (hipStreamSynchronize(stream)) return 1;

    {
        oz2::Ozaki2 oz(stream);
        oz.set_resident_hint(3 * sizeof(double) * count);
        std::printf("%4s %4s %10s %10s %10s %10s %10s %10s %8s\n", "s", "tk", "scale_ms", "split_ms", "gemm_ms",
                    "crt_ms", "wall_ms", "TFLOP/s", "tiles");
        for (int s : {8, 12, 14, 16, 20}) {
            oz.gemm(oz2::Trans::N, oz2::Trans::N, N, N, N, 1.0, dA, N, dB, N, 0.0, dC, N, s).wait();
            const auto t0 = std::chrono::steady_clock::now();
            oz.gemm(oz2::Trans::N, oz2::Trans::N, N, N, N, 1.0, dA, N, dB, N, 0.0, dC, N, s).wait();
            const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            const oz2::StageTimes st = oz.stage_times();
            const oz2::TileInfo& ti = oz.last_tiling();
            std::printf("%4d %4d %10.3f %10.3f %10.3f %10.3f %10.3f %10.2f %4lldx%lld\n", s,
                        oz2::Ozaki2::precision_bits(N, s), st.scale_ms, st.split_ms, st.gemm_ms, st.crt_ms, ms,
                        2.0 * N * N * N / (ms * 1e9), (long long)ti.pm, (long long)ti.pn);
        }
    }
    if (!dir.empty() && !(dump(dir, "A.bin", dA, count) && dump(dir, "B.bin", dB, count) &&
                          dump(dir, "C.bin", dC, count)))
        return 1;
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dC);
    (void)hipStreamDestroy(stream);
    return 0;
}
