// plan.hpp - host-side number theory of Ozaki Scheme II: moduli sets, the
// scaling width tk, CRT weights and the per-launch kernel tables.
//
// Follows the reference's make_plan / make_split_tab / make_crt_tab
// (src/ozaki2_gpu.hpp:55-236); the moduli are the sets of
// src/coprime_tables.hpp:10-111 for s = 2..24.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

namespace oz2 {

inline constexpr int kMaxModuli = 24;
inline constexpr int kMinModuli = 2;
inline constexpr int kMaxLimbs = 6;           // 32-bit words of M for s = 24
inline constexpr int kMaxScaleBits = 100;     // tk cap (hpp:29)
inline constexpr int kNarrowScaleBits = 62;   // tk <= 62: |x| < 2^63 (hpp:30)
inline constexpr int kWideScaleBits = 94;     // tk <= 94: 2^95 + x fits 96 bits (every plan has tk <= 91)
// octet form of the split for a scaling width tk: 0 (8 octets), 1 (12), 2 (14)
constexpr int limb_form(int tk) { return tk <= kNarrowScaleBits ? 0 : tk <= kWideScaleBits ? 1 : 2; }
inline constexpr int32_t kNonFinite = INT32_MIN;
// k-chunks of the INT8 GEMM.  The reference splits k at 131071 (hpp:32), the
// largest chunk whose int32 accumulation of int8 products is exact
// (|sum| <= kc * 2^14 < 2^31).  Here a chunk holds at most 65408 k (a multiple of
// the 128-byte k-step with |sum| < 2^30), which lets the GEMM's residue epilogue use
// its shorter reduction, and the chunks of one product are of equal length.
inline constexpr int64_t kRefMaxKChunk = 131071;
inline constexpr int64_t kKStep = 128;
inline constexpr int64_t kMaxKChunk = 65408;
// chunk length for kp (a multiple of kKStep): kp split into ceil(kp / kMaxKChunk)
// near-equal parts, rounded up to the k-step (<= kMaxKChunk, which is a multiple of it)
inline constexpr int64_t k_chunk_len(int64_t kp) {
    const int64_t parts = kp > 0 ? (kp + kMaxKChunk - 1) / kMaxKChunk : 1;
    return ((kp + parts - 1) / parts + kKStep - 1) / kKStep * kKStep;
}
// chunks of that length covering kp, none empty
inline constexpr int64_t k_chunks(int64_t kp) {
    return kp > 0 ? (kp + k_chunk_len(kp) - 1) / k_chunk_len(kp) : 1;
}

// 32-bit words of M per s: ceil-free form of the reference's limbs_for (hpp:55-59)
constexpr int limbs_for(int s) { return s <= 8 ? 2 : s <= 12 ? 3 : s <= 16 ? 4 : s <= 20 ? 5 : 6; }

const uint16_t* moduli(int s);  // s entries, or nullptr

// Unsigned fixed-width integer, 5 x 64-bit little-endian limbs (320 bits): the
// products of up to 24 moduli below 2^8 (M < 2^192) and their multiples by y_t < 2^8.
struct Wide {
    static constexpr int N = 5;
    uint64_t v[N] = {};
    static Wide of(uint64_t x) { Wide w; w.v[0] = x; return w; }
    void mul(uint32_t f);                 // *this *= f; throws std::overflow_error
    uint64_t divmod(uint64_t d);          // *this /= d (d > 0), returns the remainder
    int bits() const;                     // bit length (0 for zero)
    uint32_t word32(int i) const { return static_cast<uint32_t>(v[i / 2] >> (32 * (i % 2))); }
};

struct Plan {
    int s = 0;
    int64_t k = 0;
    int tk = 0;
    int NL = 0;
    uint32_t m[kMaxModuli] = {};
    uint32_t y[kMaxModuli] = {};
    Wide W[kMaxModuli];
    Wide M;
};

// Throws std::invalid_argument (s out of range, k < 1, s too small for k),
// std::runtime_error (bad table) or std::logic_error (limb count).
Plan make_plan(int s, int64_t k);
// make_plan(s, k).tk without building the plan or throwing: the largest t <= kMaxScaleBits with
// k 2^(2t+2) <= M, or 0 when s moduli do not admit k (s in [kMinModuli, kMaxModuli], k >= 1)
int scale_bits(int s, int64_t k);
// smallest s whose plan keeps >= bits per element; kMaxModuli if none does (hpp:693-700)
int moduli_for_bits(int64_t k, int bits);

// ---- constant tables read by the kernels -----------------------------------
// One TabSet per s (2..24), built on the host and copied to device memory once
// per engine; kernels read it through a const __restrict__ pointer, so the
// per-modulus constants come in through scalar loads.
struct TabSet {
    int32_t s;
    int32_t L16;                            // 16-bit limbs of M: 2 * NL
    // split (hpp:202-220): Barrett constants and symmetric offsets
    uint32_t m[kMaxModuli];
    uint32_t mu[kMaxModuli];                // ceil(2^31 / m): exact quotients of values below 2^23 (mod_m23)
    uint32_t h[kMaxModuli];                 // (m - 1) / 2
    uint32_t c16[kMaxModuli];               // 2^16 mod m
    uint32_t off[3][kMaxModuli];            // (m - 2^OFF mod m + h) mod m, OFF = 62, 95, 104 (limb_form 0, 1, 2)
    uint32_t offh[3][kMaxModuli];           // off - h (mod 2^32): start of the symmetric split's dot chain
    uint32_t h2[kMaxModuli];                // 2 h
    uint32_t c8[kMaxModuli][4];             // bytes 2^(8 j) mod m, j = 0..15, four to a word (low first)
    uint32_t c30[kMaxModuli];               // m * ceil(2^30 / m) (GEMM residue epilogue offset; 0 for m = 256)
    // CRT (hpp:222-236)
    uint32_t F16[kMaxModuli];               // round(2^16 y_t / m_t): fixed-point quotient weights
    uint32_t F16p[kMaxModuli / 2];          // F16[2i] | F16[2i+1] << 16 (0 past s)
    uint32_t W16[kMaxModuli][2 * kMaxLimbs];  // 16-bit limbs of W_t
    uint32_t W16p[kMaxModuli / 2][2 * kMaxLimbs];  // W16[2i][l] | W16[2i+1][l] << 16 (0 past s)
    uint32_t M16[2 * kMaxLimbs];            // 16-bit limbs of M
};
// tables for s = 0..24 (entries 0 and 1 unused); y_t and W_t do not depend on k
void make_tabsets(TabSet* out /* [kMaxModuli + 1] */);

}  // namespace oz2
