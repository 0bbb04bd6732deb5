// fill.hip - device inputs of the throughput benchmark (oz2/bench_ozaki2.py), the
// counterpart of the fill of the reference's benchmarks/bench_ozaki2.cpp:35-56.
// Measurement harness only, not part of liboz2.so.
//
// Column-major rows x cols, entry i = r + c*rows:
//   phi < 0:  A: sin(r + c), B: cos(r * c)
//   phi >= 0: three uniforms from the splitmix64 finalizer of 3i + seed, 3i + seed + 1,
//             3i + seed + 2 (top 53 bits, centred in their interval), a Box-Muller normal
//             g from the first two and x = (u3 - 1/2) exp(phi g).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

// splitmix64 output function (Steele, Lea, Flood 2014): golden-ratio increment, then
// two xor-shift-multiply rounds and a final xor-shift
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit53(uint64_t h) {
    return (static_cast<double>(h >> 11) + 0.5) * 0x1.0p-53;
}

__global__ void __launch_bounds__(256) bench_fill_kernel(double* __restrict__ X, int64_t rows, int64_t total,
                                                         int is_a, double phi, uint64_t seed) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int64_t c = i / rows, r = i - c * rows;
        double v;
        if (phi < 0) {
            v = is_a ? sin(static_cast<double>(r + c)) : cos(static_cast<double>(r) * static_cast<double>(c));
        } else {
            const uint64_t base = static_cast<uint64_t>(i) * 3u + seed;
            const double u1 = unit53(splitmix64(base));
            const double u2 = unit53(splitmix64(base + 1));
            const double u3 = unit53(splitmix64(base + 2)) - 0.5;
            const double g = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
            v = u3 * exp(phi * g);
        }
        X[i] = v;
    }
}

}  // namespace

extern "C" int oz2chk_bench_fill(void* stream, double* X, int64_t rows, int64_t cols, int is_a, double phi,
                                 uint64_t seed) {
    if (rows <= 0 || cols <= 0) return 0;
    if (!X) return 1;
    const int64_t total = rows * cols;
    const int64_t want = (total + 255) / 256;
    const unsigned blocks = static_cast<unsigned>(want < 65536 ? want : 65536);
    hipLaunchKernelGGL(bench_fill_kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream), X, rows,
                       total, is_a, phi, seed);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
