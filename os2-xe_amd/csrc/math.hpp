// math.hpp - per-element integer arithmetic of the Ozaki Scheme II path,
// shared by the gfx950 kernels and by the host build used to test it on the
// CPU (csrc/test/emu.cpp).  Everything here is exact integer work except the
// fp32 CRT quotient estimate, whose error is bounded far below its margin.
#pragma once
#include <cstdint>

#include "plan.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OZ2_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define OZ2_HD inline
#endif

namespace oz2 {
namespace math {

// The functions below take the constant tables as a pointer of any type with
// TabSet's members.  Kernels pass const_tab(p): loads through an address-space-4
// (constant) pointer compile to scalar loads even where the kernel also stores
// through pointers that could alias the tables (int8 planes, byte residues).
#if defined(__HIPCC__)
using ConstTab = const __attribute__((address_space(4))) TabSet*;
__device__ __forceinline__ ConstTab const_tab(const TabSet* p) { return (ConstTab)p; }
#endif

#if defined(__HIP_DEVICE_COMPILE__)
// The LLVM intrinsics behind v_mul_hi_u32_u24 and v_mul_i32_i24 (no clang builtin names
// them).  Called directly they take the low 24 bits of each operand themselves, so no
// masking instructions are emitted, and the optimizer cannot turn a 24-bit product whose
// low bits alone are used into a full 32- or 64-bit multiply (quarter rate).
extern "C" __device__ uint32_t oz2_llvm_mulhi_u24(uint32_t, uint32_t) __asm("llvm.amdgcn.mulhi.u24");
extern "C" __device__ int32_t oz2_llvm_mul_i24(int32_t, int32_t) __asm("llvm.amdgcn.mul.i24");
#endif

// v_mul_u32_u24 / v_mul_hi_u32_u24: low 32 and high 16 bits of a 24 x 24-bit product
OZ2_HD uint32_t mul24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return static_cast<uint32_t>(static_cast<uint64_t>(a & 0xFFFFFFu) * static_cast<uint64_t>(b & 0xFFFFFFu));
#endif
}
// v_mad_i32_i24: a * b + c with a, b sign-extended from 24 bits (low 32 bits of the result)
OZ2_HD uint32_t mad_i24(int32_t a, int32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return static_cast<uint32_t>(oz2_llvm_mul_i24(a, b)) + c;
#else
    const int64_t sa = static_cast<int64_t>(static_cast<uint64_t>(a) << 40) >> 40;
    const int64_t sb = static_cast<int64_t>(static_cast<uint64_t>(b) << 40) >> 40;
    return static_cast<uint32_t>(static_cast<uint64_t>(sa * sb)) + c;
#endif
}
OZ2_HD uint32_t mulhi24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return oz2_llvm_mulhi_u24(a, b);
#else
    return static_cast<uint32_t>((static_cast<uint64_t>(a & 0xFFFFFFu) * static_cast<uint64_t>(b & 0xFFFFFFu)) >> 32);
#endif
}

// v_dot2_u32_u16: c + a.lo * b.lo + a.hi * b.hi (16-bit halves, no clamping)
OZ2_HD uint32_t udot2(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
    return __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, a), __builtin_bit_cast(u16x2, b), c, false);
#else
    return c + (a & 0xFFFFu) * (b & 0xFFFFu) + (a >> 16) * (b >> 16);
#endif
}

// v_dot4_u32_u8: c + the sum of the four byte products (no clamping)
OZ2_HD uint32_t udot4(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_udot4(a, b, c, false);
#else
    for (int i = 0; i < 32; i += 8) c += ((a >> i) & 0xFFu) * ((b >> i) & 0xFFu);
    return c;
#endif
}

OZ2_HD int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }
OZ2_HD int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

// frexp exponent of a positive finite double given by its bit pattern (bits != 0)
OZ2_HD int frexp_exp_bits(uint64_t b) {
    const int bexp = static_cast<int>(b >> 52);
    if (bexp) return bexp - 1022;
    return (64 - clz64(b & 0xFFFFFFFFFFFFFull)) - 1074;
}

// exponent rule of the scaling step (hpp:360-377): 0 for an all-zero row,
// kNonFinite for Inf/NaN (bit patterns >= the infinity pattern), else tk - frexp_exp(max)
OZ2_HD int32_t exponent_from_maxbits(uint64_t b, int tk) {
    if (b >= 0x7FF0000000000000ull) return INT32_MIN;
    if (b == 0) return 0;
    return tk - frexp_exp_bits(b);
}

// Exact RNE(|v| * 2^e) for |v * 2^e| < 2^101 as a 128-bit magnitude (hpp:251-289),
// in straight-line code: with sh < 0 the 53-bit mantissa moves right by
// min(-sh, 63) and rounds to nearest even (all in lo); with sh >= 0 it moves
// left by sh <= 100 across the two words.  Shift counts stay below 64.
OZ2_HD void scaled_mag(uint64_t bits, int e, uint64_t& hi, uint64_t& lo) {
    const int bexp = static_cast<int>((bits >> 52) & 0x7FF);
    const uint64_t mant = (bits & 0xFFFFFFFFFFFFFull) | (bexp ? 0x10000000000000ull : 0);
    const int sh = (bexp ? bexp : 1) - 1075 + e;
    const int ls = sh > 0 ? sh : 0;
    const int rs = sh < 0 ? (-sh < 63 ? -sh : 63) : 0;
    const uint64_t q = mant >> rs;
    const uint64_t rem = mant & ((1ull << rs) - 1);
    const uint64_t half = rs ? (1ull << (rs - 1)) : ~0ull;
    const uint64_t down = q + (rem + (q & 1) > half ? 1 : 0);
    const uint64_t l0 = mant << (ls & 63);                       // ls < 64: low word
    const uint64_t h0 = (mant >> 1) >> ((63 - ls) & 63);         // ls < 64: bits shifted out of it
    lo = sh < 0 ? down : (ls < 64 ? l0 : 0);
    hi = sh < 0 ? 0 : (ls < 64 ? h0 : l0);                       // ls >= 64: mant << (ls - 64)
}

// RNE(|v| * 2^e) for |v * 2^e| < 2^62 (narrow values) in straight-line code: the
// mantissa moves left by max(sh, 0) and right by min(max(-sh, 0), 63) with the
// round-to-nearest-even increment rem + lsb > half (half = 2^(rs-1); none if rs = 0).
OZ2_HD uint64_t scaled_mag62(uint64_t bits, int e) {
    const int bexp = static_cast<int>((bits >> 52) & 0x7FF);
    const uint64_t mant = (bits & 0xFFFFFFFFFFFFFull) | (bexp ? 0x10000000000000ull : 0);
    const int sh = (bexp ? bexp : 1) - 1075 + e;
    const int ls = sh > 0 ? sh : 0;
    const int rs = sh < 0 ? (-sh < 63 ? -sh : 63) : 0;
    const uint64_t up = mant << ls;
    const uint64_t q = up >> rs;
    const uint64_t rem = up & ((1ull << rs) - 1);
    const uint64_t half = rs ? (1ull << (rs - 1)) : ~0ull;
    return q + (rem + (q & 1) > half ? 1 : 0);
}

// Octet forms of the offset-binary value u = 2^OFF + x (x = +-RNE(|v| 2^e), |x| <= 2^tk),
// kept as little-endian 32-bit words; the residue chain takes one v_dot4 per word:
//   Octets8  OFF = 62  (tk <= 62): the 8 bytes of u <= 2^63 in a[0..1];
//   Octets12 OFF = 95  (tk <= 94): the 12 bytes of u < 2^96 in a[0..2] (every plan: tk <= 91);
//   Octets14 OFF = 104 (tk <= 100): the 14 bytes of u <= 2^105 in a[0..3].
enum class Limbs { Octets8, Octets12, Octets14 };
template <Limbs F>
inline constexpr int kLimbWords = F == Limbs::Octets8 ? 2 : F == Limbs::Octets12 ? 3 : 4;

// zero: the element is outside the matrix or its row is non-finite, giving x = 0.
template <Limbs F>
OZ2_HD void to_limbs(uint64_t bits, int e, bool zero, uint32_t (&a)[4]) {
    const bool neg = (bits >> 63) != 0 && !zero;
    if (F != Limbs::Octets8) {  // mag <= 2^100; the high word of u is 2^(OFF - 64) +- hi
        constexpr int kTop = F == Limbs::Octets12 ? 31 : 40;
        uint64_t hi = 0, lo = 0;
        if (!zero) scaled_mag(bits, e, hi, lo);
        uint64_t uh, ul;
        if (neg) {
            ul = 0 - lo;
            uh = (1ull << kTop) - hi - (lo != 0 ? 1 : 0);
        } else {
            ul = lo;
            uh = (1ull << kTop) + hi;
        }
        a[0] = static_cast<uint32_t>(ul);
        a[1] = static_cast<uint32_t>(ul >> 32);
        a[2] = static_cast<uint32_t>(uh);
        a[3] = static_cast<uint32_t>(uh >> 32);
    } else {  // mag <= 2^62
        const uint64_t lo = zero ? 0 : scaled_mag62(bits, e);
        const uint64_t ul = neg ? (1ull << 62) - lo : (1ull << 62) + lo;
        a[0] = static_cast<uint32_t>(ul);
        a[1] = static_cast<uint32_t>(ul >> 32);
        a[2] = a[3] = 0;
    }
}

// Exact quotient by a modulus 129 <= m <= 255 for 0 <= v < 2^23: with mu = ceil(2^31 / m)
// (< 2^24) and e = mu m - 2^31 in [0, m), v mu / 2^31 = v/m + v e / (m 2^31), and
// v e < 2^23 * 255 < 2^31 keeps the excess below 1/m, so q = mulhi24(2v, mu) =
// floor(v mu / 2^31) = floor(v/m) and v - q m (one v_mad_i32_i24) is v mod m.
OZ2_HD uint32_t mod_m23(uint32_t v, uint32_t m, uint32_t mu) {
    const uint32_t q = mulhi24(v + v, mu);
    return mad_i24(static_cast<int32_t>(q), -static_cast<int32_t>(m), v);
}

// (v mod m) - h for w = v - h (mod 2^32), v < 2^23: the exact quotient of v = w + h
// takes (w << 1) + 2h (one v_lshl_add_u32) and w - q m = (v mod m) - h needs no
// separate subtraction of h.
OZ2_HD uint32_t sym_m23(uint32_t w, uint32_t m, uint32_t mu, uint32_t h2) {
    const uint32_t q = mulhi24((w << 1) + h2, mu);
    return mad_i24(static_cast<int32_t>(q), -static_cast<int32_t>(m), w);
}

// Symmetric residue r - h of x for modulus index t >= 1 (m <= 255), r = (x + h) mod m,
// as a 32-bit value whose low byte is the int8 plane entry, from the octets of
// u = 2^OFF + x: v = (h - 2^OFF mod m) + sum_j byte_j(u) (2^(8 j) mod m) <= 16 * 255 * 254
// + 254 < 2^20, one v_dot4 per word.  The chain starts at off - h (mod 2^32), so it
// yields w = v - h for sym_m23.
template <Limbs F, typename Tab>
OZ2_HD uint32_t sym_residue(const uint32_t (&a)[4], Tab tb, int t) {
    constexpr int f = static_cast<int>(F);
    uint32_t w = udot4(a[0], tb->c8[t][0], tb->offh[f][t]);
    w = udot4(a[1], tb->c8[t][1], w);
    if (kLimbWords<F> > 2) w = udot4(a[2], tb->c8[t][2], w);
    if (kLimbWords<F> > 3) w = udot4(a[3], tb->c8[t][3], w);
    return sym_m23(w, tb->m[t], tb->mu[t], tb->h2[t]);
}

// int8 byte of the symmetric residue of x for modulus index t (m = 256 only at t = 0)
template <Limbs F, typename Tab>
OZ2_HD uint32_t sym_byte(const uint32_t (&a)[4], Tab tb, int t) {
    if (t == 0) return a[0] & 0xFFu;  // m = 256: x mod 256 = u mod 256
    return sym_residue<F>(a, tb, t) & 0xFFu;
}

// the low bytes of b0..b3 packed into one word (b0 lowest): two v_perm_b32 (each
// takes byte 0 of two registers) and an or
OZ2_HD uint32_t pack_lo_bytes(uint32_t b0, uint32_t b1, uint32_t b2, uint32_t b3) {
#if defined(__HIP_DEVICE_COMPILE__)
    // v_perm_b32 D, S0, S1, sel: selector byte values 0-3 pick bytes of S1, 4-7 bytes of S0, 0x0c gives 0
    return __builtin_amdgcn_perm(b1, b0, 0x0c0c0400u) | __builtin_amdgcn_perm(b3, b2, 0x04000c0cu);
#else
    return (b0 & 0xFFu) | (b1 & 0xFFu) << 8 | (b2 & 0xFFu) << 16 | (b3 & 0xFFu) << 24;
#endif
}

// symmetric residue value in [-h, m-1-h] (complex split needs the value, not the byte)
template <Limbs F, typename Tab>
OZ2_HD int32_t sym_value(const uint32_t (&a)[4], Tab tb, int t) {
    if (t == 0) return static_cast<int32_t>((a[0] + 127u) & 0xFFu) - 127;
    return static_cast<int32_t>(sym_residue<F>(a, tb, t));
}

// GEMM epilogue: int32 dot product x (as bits) of one k-chunk -> x mod m in [0, m),
// m <= 255.  A chunk holds at most kMaxKChunk = 65408 k, so |x| <= 65408 * 2^14 <
// 2^30 - 255, and with c30 = m ceil(2^30 / m) in [2^30, 2^30 + m) the value u = x + c30
// lies in [0, 2^31) and keeps x's residue.  One v_mad_i32_i24 folds the high half:
// y = u + (u >> 16)(c16 - 2^16) = (u >> 16) c16 + (u & 0xffff) <= (2^15 - 1) 254 + 65535
// < 2^23, then y mod m by the exact quotient.
OZ2_HD uint32_t residue_m30(uint32_t x, uint32_t m, uint32_t mu, uint32_t c16, uint32_t c30) {
    const uint32_t u = x + c30;
    const uint32_t y = mad_i24(static_cast<int32_t>(u >> 16), static_cast<int32_t>(c16) - 65536, u);
    return mod_m23(y, m, mu);
}
template <typename Tab>
OZ2_HD uint32_t acc_residue(uint32_t x, Tab tb, int t) {
    const uint32_t m = tb->m[t];
    if (m == 256) return x & 0xFFu;
    return residue_m30(x, m, tb->mu[t], tb->c16[t], tb->c30[t]);
}

OZ2_HD double ldexp_(double v, int e) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ldexp(v, e);
#else
    return std::ldexp(v, e);
#endif
}

// Exact reconstruction of one entry from residues r_t in [0, m_t); NL = 32-bit words of M.
// The quotient q = round(sum_t r_t y_t / m_t) is taken in 16.16 fixed point: with
// F_t = round(2^16 y_t / m_t) the sum sum_t r_t F_t / 2^16 is within s * 255 * 2^-17 < 0.05 of
// the exact one, which lies within 1/4 of q whenever |C'| <= M/4 (the scaling guarantees it),
// so rounding it gives q (the reference's fp32 estimate gives the same q for the same reason).
// The fixed-point sum is one more dot2 per modulus pair in the limb sums.
template <int NL>
struct CrtAcc {
    static constexpr int L = 2 * NL;
    uint32_t uq;  // sum r_t F_t < 24 * 255 * 2^16 < 2^29
    uint32_t acc[L];
    OZ2_HD void init() {
        uq = 0;
        for (int l = 0; l < L; ++l) acc[l] = 0;
    }
    // uq += r F_t; acc_l += r W16[t][l]  (r < 256, W16 < 2^16: products < 2^24, sums < 2^29)
    template <typename Tab>
    OZ2_HD void add(Tab tb, int t, uint32_t r) {
        uq += mul24(r, tb->F16[t]);
        for (int l = 0; l < L; ++l) acc[l] += mul24(r, tb->W16[t][l]);
    }
    // moduli 2i and 2i+1 at once (r1 = 0 past s): one dot2 per limb with the paired table
    template <typename Tab>
    OZ2_HD void add2(Tab tb, int i, uint32_t r0, uint32_t r1) {
        const uint32_t rr = r0 | (r1 << 16);
        uq = udot2(rr, tb->F16p[i], uq);
        for (int l = 0; l < L; ++l) acc[l] = udot2(rr, tb->W16p[i][l], acc[l]);
    }
    // RNE(C') * 2^-shift with C' = sum r_t W_t - q M
    template <typename Tab>
    OZ2_HD double finish(Tab tb, int shift) const {
        const int32_t q = static_cast<int32_t>((uq + 0x8000u) >> 16);
        uint32_t w[NL];
        int32_t carry = 0;
        for (int l = 0; l < L; ++l) {
            // 0 <= q <= s * 255 and M16 < 2^16: one v_mad_i32_i24
            const int32_t v = static_cast<int32_t>(mad_i24(-q, static_cast<int32_t>(tb->M16[l]), acc[l] + static_cast<uint32_t>(carry)));
            const uint32_t lo = static_cast<uint32_t>(v) & 0xFFFFu;
            if (l & 1) w[l >> 1] |= lo << 16;
            else w[l >> 1] = lo;
            carry = v >> 16;
        }
        const bool neg = (w[NL - 1] >> 31) != 0;
        if (neg) {
            uint32_t cy = 1;
            for (int l = 0; l < NL; ++l) {
                const uint64_t s = static_cast<uint64_t>(~w[l]) + cy;
                w[l] = static_cast<uint32_t>(s);
                cy = static_cast<uint32_t>(s >> 32);
            }
        }
        // |C'| as G 64-bit groups; the window is the top 64 bits of the leading nonzero group
        // and the one below it, every lower bit folded into a sticky bit
        constexpr int G = (NL + 1) / 2;
        static_assert(G <= 3, "at most 6 words");
        uint64_t g[G];
        for (int i = 0; i < G; ++i)
            g[i] = w[2 * i] | (2 * i + 1 < NL ? static_cast<uint64_t>(w[2 * i + 1]) << 32 : 0);
        uint64_t hi = g[G - 1], lo = G >= 2 ? g[G - 2] : 0;
        bool below = G == 3 && g[0] != 0;
        int base = 64 * (G - 1);  // log2 weight of hi's lowest bit
        for (int step = 0; step + 1 < G; ++step) {  // step down past zero groups
            const bool z = hi == 0;
            const int idx = G - 2 - step;            // lo is group idx
            hi = z ? lo : hi;
            lo = z ? (idx >= 1 ? g[idx - 1] : 0) : lo;
            base -= z ? 64 : 0;
            below = z ? false : below;               // with G <= 3 nothing lies below the new lo
        }
        if (hi == 0) return 0.0;
        const int z = clz64(hi);
        uint64_t top = z ? (hi << z) | (lo >> (64 - z)) : hi;
        const uint64_t rest = z ? (lo << z) : lo;
        top |= (below || rest != 0) ? 1u : 0u;
        // the window's lsb weighs 2^(base - z); one RNE here, a second only if subnormal
        const double v = ldexp_(static_cast<double>(top), base - z - shift);
        return neg ? -v : v;
    }
};

}  // namespace math
}  // namespace oz2
