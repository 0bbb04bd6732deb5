// device.hpp - device-side integer helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace oz2 {
namespace dev {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint32_t mulhi24(uint32_t a, uint32_t b) {
    // v_mul_hi_u32_u24: bits 47..32 of a 24 x 24-bit product
    return static_cast<uint32_t>((static_cast<uint64_t>(a & 0xFFFFFFu) * static_cast<uint64_t>(b & 0xFFFFFFu)) >> 32);
}
__device__ __forceinline__ uint32_t mul24(uint32_t a, uint32_t b) { return __umul24(a, b); }

// x mod m for x < 2^24 and 2 <= m <= 255, with mu = floor(2^32 / m) (< 2^24).
// Barrett: q in {floor(x/m) - 1, floor(x/m)}, so r = x - q m lies in [0, 2m).
__device__ __forceinline__ uint32_t mod24(uint32_t x, uint32_t m, uint32_t mu) {
    const uint32_t q = mulhi24(x, mu);
    const uint32_t r = x - mul24(q, m);
    return min(r, r - m);
}

// x mod m for any 32-bit x and m <= 255: fold the high half with c16 = 2^16 mod m,
// (x >> 16) * c16 + (x & 0xffff) <= 65535 * 254 + 65535 < 2^24, then Barrett.
__device__ __forceinline__ uint32_t mod32(uint32_t x, uint32_t m, uint32_t mu, uint32_t c16) {
    return mod24(mul24(x >> 16, c16) + (x & 0xFFFFu), m, mu);
}

// frexp exponent of a positive finite double given by its bit pattern (bits != 0)
__device__ __forceinline__ int frexp_exp_bits(uint64_t b) {
    const int bexp = static_cast<int>(b >> 52);
    if (bexp) return bexp - 1022;
    const uint64_t mant = b & 0xFFFFFFFFFFFFFull;
    return (64 - __clzll(static_cast<long long>(mant))) - 1074;
}

// Exponent rule of the scaling step: 0 for an all-zero row, kNonFinite for Inf/NaN
__device__ __forceinline__ int32_t exponent_from_maxbits(uint64_t b, int tk) {
    if (b >= 0x7FF0000000000000ull) return INT32_MIN;
    if (b == 0) return 0;
    return tk - frexp_exp_bits(b);
}

// Exact RNE(|v| * 2^e) for |v * 2^e| < 2^101, as a 128-bit magnitude (hi, lo).
__device__ __forceinline__ void scaled_mag(uint64_t bits, int e, uint64_t& hi, uint64_t& lo) {
    const int bexp = static_cast<int>((bits >> 52) & 0x7FF);
    uint64_t mant = bits & 0xFFFFFFFFFFFFFull;
    if (bexp) mant |= 0x10000000000000ull;
    const int sh = (bexp ? bexp : 1) - 1075 + e;
    if (sh >= 0) {
        if (sh >= 64) {
            hi = mant << (sh - 64);
            lo = 0;
        } else {
            lo = mant << sh;
            hi = sh == 0 ? 0 : (mant >> (64 - sh));
        }
    } else {
        const int rs = min(-sh, 63);
        const uint64_t q = mant >> rs;
        const uint64_t half = 1ull << (rs - 1);
        const uint64_t rem = mant & ((half << 1) - 1);
        lo = q + ((rem > half || (rem == half && (q & 1))) ? 1 : 0);
        hi = 0;
    }
}

}  // namespace dev
}  // namespace oz2
