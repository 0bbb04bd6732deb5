// mfma_rate.hip - sustained INT8 MFMA rate of the GPU with no memory traffic: every
// wave runs 16 independent v_mfma_i32_32x32x32_i8 accumulator chains held in
// registers (the instruction and wave tile of the library's INT8 GEMM), one wave per
// SIMD, and stores one word per lane at the end.  Measurement harness only, not part
// of liboz2.so; it gives the practical MFMA ceiling that the GEMM's rate is read
// against (bench.py --mfma-rate).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__global__ void __launch_bounds__(256, 1) mfma_loop_kernel(int iters, int* __restrict__ out) {
    const v4i a = {static_cast<int>(threadIdx.x), 3, 5, 7};
    const v4i b = {1, static_cast<int>(blockIdx.x), 2, 9};
    v16i acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) acc[j][v] = 0;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc[j], 0, 0, 0);
    }
    int s = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
#pragma unroll
        for (int v = 0; v < 16; ++v) s += acc[j][v];
    out[static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x] = s;
}

}  // namespace

// blocks workgroups of 256 threads, each running iters x 16 MFMAs per wave (4 waves):
// 2 * 32768 * 16 * iters * 4 * blocks int8 operations; out needs blocks * 256 ints
extern "C" int oz2chk_mfma_rate(void* stream, int blocks, int iters, int* out) {
    if (blocks <= 0 || iters <= 0) return 1;
    if (!out) return 1;
    hipLaunchKernelGGL(mfma_loop_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(stream), iters, out);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
