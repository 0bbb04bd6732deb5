// emu.cpp - host build of the kernels' per-element arithmetic (csrc/math.hpp),
// so the CPU test suite can compare it bit-for-bit with the oracle without a
// GPU.  Test support only (lib/liboz2_emu.so); the product library is liboz2.so.
#include <cstdint>
#include <vector>

#include "../math.hpp"
#include "../planner.hpp"

using namespace oz2;

namespace {
const TabSet* tables() {
    static std::vector<TabSet> t = [] {
        std::vector<TabSet> v(kMaxModuli + 1);
        make_tabsets(v.data());
        return v;
    }();
    return t.data();
}
template <math::Limbs F>
void split_form(const double* v, const int32_t* e, int64_t n, int s, int8_t* planes) {
    const TabSet* tb = tables() + s;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a[4];
        uint64_t bits;
        __builtin_memcpy(&bits, v + i, 8);
        const bool zero = e[i] == INT32_MIN;
        math::to_limbs<F>(bits, zero ? 0 : e[i], zero, a);
        for (int t = 0; t < s; ++t) planes[t * n + i] = static_cast<int8_t>(math::sym_byte<F>(a, tb, t));
    }
}

template <math::Limbs F>
void split_cplx_form(const double* re, const double* im, const int32_t* e, int64_t n, int s, int conj,
                     int8_t* planes) {
    const TabSet* tb = tables() + s;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t ar[4], ai[4];
        uint64_t br, bi;
        __builtin_memcpy(&br, re + i, 8);
        __builtin_memcpy(&bi, im + i, 8);
        const bool zero = e[i] == INT32_MIN;
        const int ee = zero ? 0 : e[i];
        math::to_limbs<F>(br, ee, zero, ar);
        math::to_limbs<F>(bi, ee, zero, ai);
        for (int t = 0; t < s; ++t) {
            const int32_t m = static_cast<int32_t>(tb->m[t]), h = static_cast<int32_t>(tb->h[t]);
            const int32_t s0 = math::sym_value<F>(ar, tb, t);
            int32_t s1 = math::sym_value<F>(ai, tb, t);
            if (conj) s1 = -s1;
            int32_t ss = s0 + s1;
            ss = ss > m - 1 - h ? ss - m : ss;
            ss = ss < -h ? ss + m : ss;
            planes[(0 * s + t) * n + i] = static_cast<int8_t>(s0);
            planes[(1 * s + t) * n + i] = static_cast<int8_t>(s1);
            planes[(2 * s + t) * n + i] = static_cast<int8_t>(ss);
        }
    }
}

}  // namespace

extern "C" {

// symmetric int8 residues of RNE(v_i * 2^e_i): planes[t * n + i]; e = INT32_MIN gives 0.
// the octets of u = 2^OFF + x: 8 bytes for tk <= 62, else 14
void emu_split(const double* v, const int32_t* e, int64_t n, int s, int tk, int8_t* planes) {
    switch (limb_form(tk)) {
        case 0: split_form<math::Limbs::Octets8>(v, e, n, s, planes); break;
        case 1: split_form<math::Limbs::Octets12>(v, e, n, s, planes); break;
        default: split_form<math::Limbs::Octets14>(v, e, n, s, planes); break;
    }
}

// complex: groups Re, Im (negated if conj), Re+Im reduced to [-h, m-1-h]: planes[(g*s + t)*n + i]
void emu_split_cplx(const double* re, const double* im, const int32_t* e, int64_t n, int s, int tk, int conj,
                    int8_t* planes) {
    switch (limb_form(tk)) {
        case 0: split_cplx_form<math::Limbs::Octets8>(re, im, e, n, s, conj, planes); break;
        case 1: split_cplx_form<math::Limbs::Octets12>(re, im, e, n, s, conj, planes); break;
        default: split_cplx_form<math::Limbs::Octets14>(re, im, e, n, s, conj, planes); break;
    }
}

// exact-quotient reduction of modulus t of s over every v < 2^23: entries that differ from v mod m
int64_t emu_mod_m23_mismatches(int s, int t) {
    const TabSet* tb = tables() + s;
    const uint32_t m = tb->m[t];
    int64_t bad = 0;
    for (uint32_t v = 0; v < (1u << 23); ++v) bad += math::mod_m23(v, m, tb->mu[t]) != v % m;
    return bad;
}

// symmetric split reduction of modulus t of s over every dot-chain value v < 2^20 + 2^8
// (w = v - h as the chain leaves it): entries that differ from (v mod m) - h
int64_t emu_sym_m23_mismatches(int s, int t) {
    const TabSet* tb = tables() + s;
    const uint32_t m = tb->m[t], h = tb->h[t];
    int64_t bad = 0;
    for (uint32_t v = 0; v < (1u << 20) + 256; ++v) {
        const int32_t got = static_cast<int32_t>(math::sym_m23(v - h, m, tb->mu[t], tb->h2[t]));
        bad += got != static_cast<int32_t>(v % m) - static_cast<int32_t>(h);
    }
    return bad;
}

// GEMM epilogue: residues of int32 sums, out[t * n + i] in [0, m_t)
void emu_acc_residue(const int32_t* x, int64_t n, int s, uint8_t* out) {
    const TabSet* tb = tables() + s;
    for (int t = 0; t < s; ++t)
        for (int64_t i = 0; i < n; ++i) out[t * n + i] = static_cast<uint8_t>(math::acc_residue(static_cast<uint32_t>(x[i]), tb, t));
}

// CRT of residues res[i * s + t] in [0, m_t): out[i] = RNE(C') * 2^-shift[i]
void emu_crt(const uint8_t* res, const int32_t* shift, int64_t n, int s, double* out) {
    const TabSet* tb = tables() + s;
    const int nl = limbs_for(s);
    for (int64_t i = 0; i < n; ++i) {
        const uint8_t* r = res + i * s;
        switch (nl) {
#define OZ2_EMU_NL(NL)                                     \
    case NL: {                                             \
        math::CrtAcc<NL> a;                                \
        a.init();                                          \
        for (int t = 0; t < s; t += 2)                     \
            a.add2(tb, t / 2, r[t], t + 1 < s ? r[t + 1] : 0u); \
        out[i] = a.finish(tb, shift[i]);                   \
        break;                                             \
    }
            OZ2_EMU_NL(2) OZ2_EMU_NL(3) OZ2_EMU_NL(4) OZ2_EMU_NL(5) OZ2_EMU_NL(6)
#undef OZ2_EMU_NL
        }
    }
}

int emu_exponent_from_maxbits(uint64_t b, int tk) { return math::exponent_from_maxbits(b, tk); }

// make_plan(s, k): tk, NL, y[t], W[t] and M as 5 little-endian 64-bit words each; returns 0, or 1 when
// make_plan throws std::invalid_argument (s out of range, k < 1, s too small for k), 2 for other errors
int emu_plan(int s, int64_t k, int* tk, int* nl, uint32_t* y, uint64_t* W, uint64_t* M) {
    try {
        const Plan p = make_plan(s, k);
        *tk = p.tk;
        *nl = p.NL;
        for (int t = 0; t < s; ++t) {
            y[t] = p.y[t];
            for (int i = 0; i < Wide::N; ++i) W[t * Wide::N + i] = p.W[t].v[i];
        }
        for (int i = 0; i < Wide::N; ++i) M[i] = p.M.v[i];
        return 0;
    } catch (const std::invalid_argument&) {
        return 1;
    } catch (...) {
        return 2;
    }
}

// ---- panel planner (csrc/planner.cpp) ----
// shape: m, n, k, s, cplx, partial_a, partial_b, cus, pad, pow2, per_plane, forced_tile
static CallShape emu_shape(const int64_t* sh) {
    CallShape c;
    c.m = sh[0];
    c.n = sh[1];
    c.k = sh[2];
    c.s = static_cast<int>(sh[3]);
    c.cplx = sh[4] != 0;
    c.partial_a = static_cast<size_t>(sh[5]);
    c.partial_b = static_cast<size_t>(sh[6]);
    c.cus = static_cast<int>(sh[7]);
    c.pad_planes = sh[8] != 0;
    c.pow2 = sh[9] != 0;
    c.per_plane = sh[10] != 0;
    c.forced_tile = static_cast<int>(sh[11]);
    return c;
}
// out: mb, nb, pm, pn, k_chunks, workspace_bytes; returns 0, or 1 when nothing fits
int emu_plan_tiling(const int64_t* sh, uint64_t limit, int64_t* out) {
    try {
        const TileInfo t = plan_tiling(emu_shape(sh), limit);
        out[0] = t.mb;
        out[1] = t.nb;
        out[2] = t.pm;
        out[3] = t.pn;
        out[4] = t.k_chunks;
        out[5] = static_cast<int64_t>(t.workspace_bytes);
        return 0;
    } catch (const std::runtime_error&) {
        return 1;
    }
}
uint64_t emu_layout_total(const int64_t* sh, int64_t mb, int64_t nb) { return layout(emu_shape(sh), mb, nb).total; }
double emu_panel_cost(const int64_t* sh, int64_t mb, int64_t nb) { return panel_cost(emu_shape(sh), mb, nb); }
}
