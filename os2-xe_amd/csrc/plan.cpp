// plan.cpp - moduli sets, make_plan and the kernel tables (host only).
#include "plan.hpp"

#include <algorithm>
#include <cstring>

namespace oz2 {

namespace {
// Pairwise-coprime moduli for s = 2..24, identical to the sets of the
// reference's src/coprime_tables.hpp:10-111 (COPRIME_DATA at COPRIME_OFFSET[s]).
constexpr uint16_t kFirst6[6] = {256, 255, 253, 251, 247, 241};
constexpr uint16_t kSet7to13[13] = {256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227, 223, 211};
constexpr uint16_t kSet14to24[24] = {256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223,
                                     217, 211, 199, 197, 193, 191, 181, 179, 173, 167, 163, 157};
}  // namespace

const uint16_t* moduli(int s) {
    if (s < kMinModuli || s > kMaxModuli) return nullptr;
    if (s <= 6) return kFirst6;
    if (s <= 13) return kSet7to13;
    return kSet14to24;
}

// ---- Wide -------------------------------------------------------------------
void Wide::mul(uint32_t f) {
    unsigned __int128 c = 0;
    for (int i = 0; i < N; ++i) {
        const unsigned __int128 t = static_cast<unsigned __int128>(v[i]) * f + c;
        v[i] = static_cast<uint64_t>(t);
        c = t >> 64;
    }
    if (c) throw std::overflow_error("oz2: integer overflow in plan arithmetic");
}

uint64_t Wide::divmod(uint64_t d) {
    // schoolbook division by one 64-bit digit: the running remainder stays below d
    unsigned __int128 r = 0;
    for (int i = N - 1; i >= 0; --i) {
        const unsigned __int128 cur = (r << 64) | v[i];
        v[i] = static_cast<uint64_t>(cur / d);
        r = cur % d;
    }
    return static_cast<uint64_t>(r);
}

int Wide::bits() const {
    for (int i = N - 1; i >= 0; --i)
        if (v[i]) return 64 * i + 64 - __builtin_clzll(v[i]);
    return 0;
}

// ---- make_plan (what it computes: hpp:159-200) ------------------------------
namespace {
// inverse of a modulo p (0 < a < p, gcd 1): extended Euclid on small integers
uint32_t inverse_mod_small(uint32_t a, uint32_t p) {
    int64_t g0 = p, g1 = a, x0 = 0, x1 = 1;
    while (g1 != 0) {
        const int64_t q = g0 / g1;
        int64_t t = g0 - q * g1;
        g0 = g1;
        g1 = t;
        t = x0 - q * x1;
        x0 = x1;
        x1 = t;
    }
    if (g0 != 1) throw std::runtime_error("oz2: moduli are not pairwise coprime");
    return static_cast<uint32_t>((x0 % p + p) % p);
}
}  // namespace

Plan make_plan(int s, int64_t k) {
    if (s < kMinModuli || s > kMaxModuli)
        throw std::invalid_argument("oz2: num_moduli must be in [2, " + std::to_string(kMaxModuli) + "]");
    if (k < 1) throw std::invalid_argument("oz2: k must be >= 1");
    const uint16_t* set = moduli(s);
    Plan p;
    p.s = s;
    p.k = k;
    for (int t = 0; t < s; ++t) p.m[t] = set[t];

    // CRT basis: W_t = y_t * prod_{u != t} m_u with y_t the inverse of that product
    // modulo m_t.  The product's residue mod m_t comes from the small moduli directly,
    // so no multi-word division is needed; M is the product of all of them.
    p.M = Wide::of(1);
    for (int t = 0; t < s; ++t) p.M.mul(p.m[t]);
    for (int t = 0; t < s; ++t) {
        Wide rest = Wide::of(1);
        uint32_t rest_mod = 1;
        for (int u = 0; u < s; ++u) {
            if (u == t) continue;
            rest.mul(p.m[u]);
            rest_mod = rest_mod * (p.m[u] % p.m[t]) % p.m[t];
        }
        p.y[t] = inverse_mod_small(rest_mod, p.m[t]);
        rest.mul(p.y[t]);
        p.W[t] = rest;
    }

    // 32-bit words that hold M: floor(log2 M / 32) + 1 with log2 M taken from the bit length
    // (M is no power of two), at least 2; the kernels' limb count per s must agree
    p.NL = std::max(2, (p.M.bits() - 1) / 32 + 1);
    if (p.NL != limbs_for(s)) throw std::logic_error("oz2: limb count mismatch with the moduli table");

    p.tk = scale_bits(s, k);
    if (p.tk < 1)
        throw std::invalid_argument("oz2: num_moduli=" + std::to_string(s) + " is too small for k=" +
                                    std::to_string(k));
    return p;
}

int scale_bits(int s, int64_t k) {
    // M per s, built once
    static const struct Products {
        Wide M[kMaxModuli + 1];
        Products() {
            for (int t = kMinModuli; t <= kMaxModuli; ++t) {
                M[t] = Wide::of(1);
                for (int u = 0; u < t; ++u) M[t].mul(moduli(t)[u]);
            }
        }
    } prod;
    // Since k 2^(2t+2) is an integer, k 2^(2t+2) <= M  <=>  2^(2t+2) <= floor(M / k)  <=>
    // 2t + 2 <= bits(floor(M / k)) - 1, so tk = (bits(floor(M / k)) - 3) / 2 rounded down,
    // exactly, from one division; floor(M / k) < 16 admits no t >= 1
    Wide q = prod.M[s];
    q.divmod(static_cast<uint64_t>(k));
    if (q.bits() < 5) return 0;
    return std::min((q.bits() - 3) / 2, kMaxScaleBits);
}

// ---- tables -------------------------------------------------------------------
namespace {
uint32_t pow2_mod(int e, uint32_t m) {
    uint64_t r = 1 % m;
    for (int i = 0; i < e; ++i) r = (r * 2) % m;
    return static_cast<uint32_t>(r);
}
}  // namespace

void make_tabsets(TabSet* out) {
    for (int s = 0; s <= kMaxModuli; ++s) out[s] = TabSet{};
    for (int s = kMinModuli; s <= kMaxModuli; ++s) {
        const Plan p = make_plan(s, 1);
        TabSet& tb = out[s];
        tb.s = s;
        tb.L16 = 2 * p.NL;
        for (int t = 0; t < s; ++t) {
            const uint32_t m = p.m[t];
            const uint32_t h = (m - 1) / 2;
            tb.m[t] = m;
            tb.mu[t] = static_cast<uint32_t>(((uint64_t{1} << 31) + m - 1) / m);
            tb.h[t] = h;
            tb.c16[t] = pow2_mod(16, m);
            const int offs[3] = {62, 95, 104};
            for (int f = 0; f < 3; ++f) {
                tb.off[f][t] = (m - pow2_mod(offs[f], m) + h) % m;
                tb.offh[f][t] = tb.off[f][t] - h;
            }
            tb.h2[t] = 2 * h;
            for (int j = 0; j < 16; ++j) tb.c8[t][j >> 2] |= static_cast<uint32_t>(pow2_mod(8 * j, m)) << (8 * (j & 3));
            tb.c30[t] = m == 256 ? 0 : m * static_cast<uint32_t>(((uint64_t{1} << 30) + m - 1) / m);
            tb.F16[t] = static_cast<uint32_t>((static_cast<uint64_t>(p.y[t]) * (1u << 17) + m) / (2u * m));
            for (int l = 0; l < tb.L16; ++l) tb.W16[t][l] = (p.W[t].word32(l / 2) >> (16 * (l % 2))) & 0xFFFFu;
        }
        for (int l = 0; l < tb.L16; ++l) tb.M16[l] = (p.M.word32(l / 2) >> (16 * (l % 2))) & 0xFFFFu;
        for (int t = 0; t < s; t += 2) {
            for (int l = 0; l < tb.L16; ++l)
                tb.W16p[t / 2][l] = tb.W16[t][l] | (t + 1 < s ? tb.W16[t + 1][l] << 16 : 0u);
            tb.F16p[t / 2] = tb.F16[t] | (t + 1 < s ? tb.F16[t + 1] << 16 : 0u);
        }
    }
}

int moduli_for_bits(int64_t k, int bits) {
    if (k < 1) return kMaxModuli;  // no plan admits k < 1 (hpp:693-700 then ends at the largest set)
    for (int s = kMinModuli; s <= kMaxModuli; ++s)
        if (scale_bits(s, k) >= bits) return s;
    return kMaxModuli;
}

}  // namespace oz2
