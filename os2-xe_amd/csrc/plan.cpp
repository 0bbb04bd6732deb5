// plan.cpp - moduli sets, make_plan and the kernel tables (host only).
#include "plan.hpp"

#include <algorithm>
#include <cmath>
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

uint32_t Wide::divmod(uint32_t d) {
    unsigned __int128 r = 0;
    for (int i = N - 1; i >= 0; --i) {
        const unsigned __int128 cur = (r << 64) | v[i];
        v[i] = static_cast<uint64_t>(cur / d);
        r = cur % d;
    }
    return static_cast<uint32_t>(r);
}

uint32_t Wide::mod(uint32_t d) const {
    Wide t = *this;
    return t.divmod(d);
}

int Wide::bits() const {
    for (int i = N - 1; i >= 0; --i)
        if (v[i]) return 64 * i + 64 - __builtin_clzll(v[i]);
    return 0;
}

bool Wide::le(const Wide& o) const {
    for (int i = N - 1; i >= 0; --i)
        if (v[i] != o.v[i]) return v[i] < o.v[i];
    return true;
}

Wide Wide::shl(int n) const {
    Wide r;
    const int q = n / 64, b = n % 64;
    for (int i = N - 1; i >= 0; --i) {
        const int src = i - q;
        if (src < 0) continue;
        uint64_t x = v[src] << b;
        if (b && src > 0) x |= v[src - 1] >> (64 - b);
        r.v[i] = x;
    }
    if (bits() + n > 64 * N) throw std::overflow_error("oz2: shift overflow in plan arithmetic");
    return r;
}

double Wide::log2_estimate() const {
    // top three 32-bit words as a double, then scaled: the same estimate as the
    // reference's Big::log2 (hpp:129-135), so tk's starting point matches it
    int top = 2 * N - 1;
    while (top > 0 && word32(top) == 0) --top;
    double x = 0;
    for (int i = top; i >= std::max(0, top - 2); --i) x = x * 4294967296.0 + word32(i);
    return std::log2(x) + 32.0 * std::max(0, top - 2);
}

// ---- make_plan (hpp:159-200) -----------------------------------------------
namespace {
int64_t inverse_mod(int64_t a, int64_t m) {
    int64_t r0 = m, r1 = ((a % m) + m) % m, t0 = 0, t1 = 1;
    while (r1) {
        const int64_t q = r0 / r1;
        const int64_t r2 = r0 - q * r1, t2 = t0 - q * t1;
        r0 = r1; r1 = r2; t0 = t1; t1 = t2;
    }
    if (r0 != 1) throw std::runtime_error("oz2: moduli are not pairwise coprime");
    return ((t0 % m) + m) % m;
}
}  // namespace

Plan make_plan(int s, int64_t k) {
    if (s < kMinModuli || s > kMaxModuli)
        throw std::invalid_argument("oz2: num_moduli must be in [2, " + std::to_string(kMaxModuli) + "]");
    if (k < 1) throw std::invalid_argument("oz2: k must be >= 1");
    Plan p;
    p.s = s;
    p.k = k;
    const uint16_t* mods = moduli(s);
    p.M = Wide::of(1);
    for (int t = 0; t < s; ++t) {
        if (mods[t] < 2 || mods[t] > 256) throw std::runtime_error("oz2: modulus out of int8 range");
        p.M.mul(mods[t]);
    }
    p.log2M = p.M.log2_estimate();
    p.NL = std::max(2, static_cast<int>(std::floor(p.log2M / 32.0)) + 1);
    if (p.NL != limbs_for(s)) throw std::logic_error("oz2: limb count mismatch with the moduli table");
    int tk = static_cast<int>(std::floor((p.log2M - 2.0 - std::log2(static_cast<double>(k))) / 2.0));
    tk = std::min(tk, kMaxScaleBits);
    for (;; --tk) {
        if (tk < 1)
            throw std::invalid_argument("oz2: num_moduli=" + std::to_string(s) + " is too small for k=" +
                                        std::to_string(k));
        if (Wide::of(static_cast<uint64_t>(k)).shl(2 * tk + 2).le(p.M)) break;
    }
    p.tk = tk;
    for (int t = 0; t < s; ++t) {
        const uint32_t m = mods[t];
        Wide Mt = p.M;
        if (Mt.divmod(m) != 0) throw std::runtime_error("oz2: bad modulus table");
        const int64_t y = inverse_mod(Mt.mod(m), m);
        p.W[t] = Mt;
        p.W[t].mul(static_cast<uint32_t>(y));
        p.m[t] = m;
        p.y[t] = static_cast<uint32_t>(y);
    }
    return p;
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
            tb.off[0][t] = (m - pow2_mod(62, m) + h) % m;
            tb.off[1][t] = (m - pow2_mod(104, m) + h) % m;
            tb.offh[0][t] = tb.off[0][t] - h;
            tb.offh[1][t] = tb.off[1][t] - h;
            tb.h2[t] = 2 * h;
            for (int j = 0; j < 16; ++j) tb.c8[t][j >> 2] |= static_cast<uint32_t>(pow2_mod(8 * j, m)) << (8 * (j & 3));
            tb.c30[t] = m == 256 ? 0 : m * static_cast<uint32_t>(((uint64_t{1} << 30) + m - 1) / m);
            tb.frac[t] = static_cast<float>(static_cast<double>(p.y[t]) / m);
            for (int l = 0; l < tb.L16; ++l) tb.W16[t][l] = (p.W[t].word32(l / 2) >> (16 * (l % 2))) & 0xFFFFu;
        }
        for (int l = 0; l < tb.L16; ++l) tb.M16[l] = (p.M.word32(l / 2) >> (16 * (l % 2))) & 0xFFFFu;
        for (int t = 0; t < s; t += 2)
            for (int l = 0; l < tb.L16; ++l)
                tb.W16p[t / 2][l] = tb.W16[t][l] | (t + 1 < s ? tb.W16[t + 1][l] << 16 : 0u);
    }
}

int moduli_for_bits(int64_t k, int bits) {
    for (int s = kMinModuli; s <= kMaxModuli; ++s) {
        try {
            if (make_plan(s, k).tk >= bits) return s;
        } catch (const std::invalid_argument&) {
        }
    }
    return kMaxModuli;
}

}  // namespace oz2
