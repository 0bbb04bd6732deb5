// crt.hip - CRT reconstruction fused with the alpha/beta epilogue, for gfx950.
//
// Reference behaviour: src/ozaki2_gpu.hpp:526-664.  For each entry the s
// residues r_t (bytes from the GEMM epilogue; k-chunks are summed mod m_t)
// determine the exact integer C' with |C'| <= M/4:
//   q  = round(sum_t r_t F_t / 2^16)            (F_t = round(2^16 y_t / m_t); error < 0.05, so q is exact)
//   C' = sum_t r_t W_t - q M                     (16-bit limbs, 32-bit sums)
// C' is rounded once to 53 bits (sticky bit) and scaled by 2^-(D_i + E_j),
// which rounds again only for subnormal results, as in the reference.
// Rows / columns with a non-finite exponent give NaN.  The epilogue is
// C = alpha*v (beta == 0, C not read) or fma(alpha, v, beta*C); complex:
// Re = P0 - P1, Im = P2 - P0 - P1 in residue space (3M form), then
// (ar*vr - ai*vi, ar*vi + ai*vr) + beta*C, each product pair fused as one fma.
#include "grouped.hpp"
#include "kernels.hpp"
#include "math.hpp"

namespace oz2 {
namespace k {
using namespace math;

namespace {

// residue r_t of one entry, summing k-chunks mod m_t
__device__ __forceinline__ uint32_t chunk_sum(uint32_t s, uint32_t m) {
    // s < kch * 256; reduce by repeated folding (kch is small)
    while (s >= m) s -= m;
    return s;
}

// 4 consecutive rows per thread; residues are bytes packed 4 per 32-bit load.
template <int NL, bool CPLX>
__device__ __forceinline__ void crt_body(const CrtArgs& a, uint32_t bx) {
    const auto tb = const_tab(a.tab);
    const int64_t mq = (a.mb + 3) >> 2;
    const int64_t idx = static_cast<int64_t>(bx) * 256 + threadIdx.x;
    const int64_t j = idx / mq;
    const int64_t i0 = (idx - j * mq) * 4;
    if (j >= a.nb) return;
    const int32_t Ej = a.E[j];
    const uint8_t* base = a.R + j * a.ldr + i0;
    const int s = a.s;

    CrtAcc<NL> re[4];
    CrtAcc<NL> im[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        re[e].init();
        if (CPLX) im[e].init();
    }
    if (!CPLX && a.kch == 1) {
        // one k-chunk: the s residue words (4 entries each) are all loaded before any
        // arithmetic, then summed in modulus pairs (2 residues per dot2 step); r = 0 past s
        uint32_t w[kMaxModuli];
#pragma unroll
        for (int t = 0; t < kMaxModuli; ++t)
            w[t] = t < s ? *reinterpret_cast<const uint32_t*>(base + t * a.plane_r) : 0u;
#pragma unroll
        for (int t = 0; t < kMaxModuli; t += 2) {
            if (t < s) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    re[e].add2(tb, t >> 1, (w[t] >> (8 * e)) & 0xFFu, (w[t + 1] >> (8 * e)) & 0xFFu);
            }
        }
    }
    if (CPLX && a.kch == 1) {
        // the same for the three 3M groups, in two halves of the moduli to bound the
        // words held at once; Re = P0 - P1 and Im = P2 - P0 - P1 reduced into [0, m)
        const int64_t gs = static_cast<int64_t>(s) * a.plane_r;
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            constexpr int H = kMaxModuli / 2;
            const int t0 = half * H;
            if (t0 >= s) break;
            uint32_t w0[H], w1[H], w2[H];
#pragma unroll
            for (int u = 0; u < H; ++u) {
                const uint8_t* b = base + (t0 + u) * a.plane_r;
                const bool in = t0 + u < s;
                w0[u] = in ? *reinterpret_cast<const uint32_t*>(b) : 0u;
                w1[u] = in ? *reinterpret_cast<const uint32_t*>(b + gs) : 0u;
                w2[u] = in ? *reinterpret_cast<const uint32_t*>(b + 2 * gs) : 0u;
            }
#pragma unroll
            for (int u = 0; u < H; u += 2) {
                if (t0 + u < s) {
                    const uint32_t m[2] = {tb->m[t0 + u], tb->m[t0 + u + 1]};  // 0 past s (residues 0 there)
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        uint32_t xr[2], xi[2];
#pragma unroll
                        for (int h = 0; h < 2; ++h) {
                            const uint32_t x0 = (w0[u + h] >> (8 * e)) & 0xFFu;
                            const uint32_t x1 = (w1[u + h] >> (8 * e)) & 0xFFu;
                            const uint32_t x2 = (w2[u + h] >> (8 * e)) & 0xFFu;
                            uint32_t r = x0 + m[h] - x1;  // [0, 2m)
                            xr[h] = r >= m[h] ? r - m[h] : r;
                            r = x2 + 2 * m[h] - x0 - x1;  // [0, 3m)
                            r = r >= m[h] ? r - m[h] : r;
                            xi[h] = r >= m[h] ? r - m[h] : r;
                        }
                        re[e].add2(tb, (t0 + u) >> 1, xr[0], xr[1]);
                        im[e].add2(tb, (t0 + u) >> 1, xi[0], xi[1]);
                    }
                }
            }
        }
    }
    // moduli in pairs (2 residues per dot2 step of the limb sums); r = 0 past s
    for (int t = 0; t < s && a.kch > 1; t += 2) {
        const bool two = t + 1 < s;
        const uint32_t m0 = tb->m[t], m1 = two ? tb->m[t + 1] : 1u;
        if (!CPLX) {
            uint32_t r0[4] = {0, 0, 0, 0}, r1[4] = {0, 0, 0, 0};
            for (int c = 0; c < a.kch; ++c) {
                const uint8_t* b = base + c * a.chunk_r + t * a.plane_r;
                const uint32_t w0 = *reinterpret_cast<const uint32_t*>(b);
                const uint32_t w1 = two ? *reinterpret_cast<const uint32_t*>(b + a.plane_r) : 0u;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    r0[e] += (w0 >> (8 * e)) & 0xFFu;
                    r1[e] += (w1 >> (8 * e)) & 0xFFu;
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const uint32_t x0 = a.kch > 1 ? chunk_sum(r0[e], m0) : r0[e];
                const uint32_t x1 = a.kch > 1 ? chunk_sum(r1[e], m1) : r1[e];
                re[e].add2(tb, t >> 1, x0, x1);
            }
        } else {
            uint32_t p0[2][4] = {}, p1[2][4] = {}, p2[2][4] = {};
            const int64_t gs = static_cast<int64_t>(s) * a.plane_r;
            for (int c = 0; c < a.kch; ++c) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h == 1 && !two) break;
                    const uint8_t* b = base + c * a.chunk_r + (t + h) * a.plane_r;
                    const uint32_t w0 = *reinterpret_cast<const uint32_t*>(b);
                    const uint32_t w1 = *reinterpret_cast<const uint32_t*>(b + gs);
                    const uint32_t w2 = *reinterpret_cast<const uint32_t*>(b + 2 * gs);
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        p0[h][e] += (w0 >> (8 * e)) & 0xFFu;
                        p1[h][e] += (w1 >> (8 * e)) & 0xFFu;
                        p2[h][e] += (w2 >> (8 * e)) & 0xFFu;
                    }
                }
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                uint32_t xr[2] = {0, 0}, xi[2] = {0, 0};
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    if (h == 1 && !two) break;
                    const uint32_t m = h ? m1 : m0;
                    const uint32_t x0 = a.kch > 1 ? chunk_sum(p0[h][e], m) : p0[h][e];
                    const uint32_t x1 = a.kch > 1 ? chunk_sum(p1[h][e], m) : p1[h][e];
                    const uint32_t x2 = a.kch > 1 ? chunk_sum(p2[h][e], m) : p2[h][e];
                    xr[h] = chunk_sum(x0 + m - x1, m);           // Re = P0 - P1
                    xi[h] = chunk_sum(x2 + 2 * m - x0 - x1, m);  // Im = P2 - P0 - P1
                }
                re[e].add2(tb, t >> 1, xr[0], xr[1]);
                im[e].add2(tb, t >> 1, xi[0], xi[1]);
            }
        }
    }
    if (!CPLX && i0 + 4 <= a.mb) {
        // four entries of one column: 16-byte accesses to C when it is aligned for them
        double v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int32_t Di = a.D[i0 + e];
            const bool nonfinite = Di == INT32_MIN || Ej == INT32_MIN;
            v[e] = re[e].finish(tb, nonfinite ? 0 : Di + Ej);
            if (nonfinite) v[e] = __builtin_nan("");
        }
        double* cp = a.C + i0 + j * a.ldc;
        if ((reinterpret_cast<uintptr_t>(cp) & 15) == 0) {
            double2* c2 = reinterpret_cast<double2*>(cp);
            if (a.br == 0.0) {
                c2[0] = make_double2(a.ar * v[0], a.ar * v[1]);
                c2[1] = make_double2(a.ar * v[2], a.ar * v[3]);
            } else {
                const double2 x0 = c2[0], x1 = c2[1];
                c2[0] = make_double2(__builtin_fma(a.ar, v[0], a.br * x0.x), __builtin_fma(a.ar, v[1], a.br * x0.y));
                c2[1] = make_double2(__builtin_fma(a.ar, v[2], a.br * x1.x), __builtin_fma(a.ar, v[3], a.br * x1.y));
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                cp[e] = a.br == 0.0 ? a.ar * v[e] : __builtin_fma(a.ar, v[e], a.br * cp[e]);
        }
        return;
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int64_t i = i0 + e;
        if (i >= a.mb) break;
        const int32_t Di = a.D[i];
        const bool nonfinite = Di == INT32_MIN || Ej == INT32_MIN;
        const int shift = nonfinite ? 0 : Di + Ej;
        if (!CPLX) {
            double v = re[e].finish(tb, shift);
            if (nonfinite) v = __builtin_nan("");
            double* cp = a.C + i + j * a.ldc;
            *cp = a.br == 0.0 ? a.ar * v : __builtin_fma(a.ar, v, a.br * *cp);
        } else {
            double vr = re[e].finish(tb, shift), vi = im[e].finish(tb, shift);
            if (nonfinite) vr = vi = __builtin_nan("");
            double* cp = a.C + 2 * (i + j * a.ldc);
            double orr = __builtin_fma(a.ar, vr, -(a.ai * vi));
            double oi = __builtin_fma(a.ar, vi, a.ai * vr);
            if (a.br != 0.0 || a.bi != 0.0) {
                const double cr = cp[0], ci = cp[1];
                orr = orr + __builtin_fma(a.br, cr, -(a.bi * ci));
                oi = oi + __builtin_fma(a.br, ci, a.bi * cr);
            }
            cp[0] = orr;
            cp[1] = oi;
        }
    }
}

template <int NL, bool CPLX>
__global__ void __launch_bounds__(256) crt_kernel(CrtArgs a) {
    crt_body<NL, CPLX>(a, blockIdx.x);
}

template <int NL, bool CPLX>
__global__ void __launch_bounds__(256) crt_grouped_kernel(const CrtArgs* __restrict__ items,
                                                          const uint32_t* __restrict__ start, uint32_t n) {
    const uint32_t p = owner_of(start, n, blockIdx.x);
    crt_body<NL, CPLX>(items[p], blockIdx.x - start[p]);
}

uint32_t crt_blocks(const CrtArgs& a) {
    const int64_t threads = ((a.mb + 3) / 4) * a.nb;
    return static_cast<uint32_t>((threads + 255) / 256);
}

__global__ void scale_c_kernel(double* C, int64_t ldc, int64_t m, int64_t n, double br, double bi, int cplx) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= m * n) return;
    const int64_t j = idx / m, i = idx - j * m;
    if (!cplx) {
        double* cp = C + i + j * ldc;
        *cp = br == 0.0 ? 0.0 : br * *cp;
    } else {
        double* cp = C + 2 * (i + j * ldc);
        if (br == 0.0 && bi == 0.0) {
            cp[0] = 0.0;
            cp[1] = 0.0;
        } else {
            const double r = cp[0], im = cp[1];
            cp[0] = __builtin_fma(br, r, -(bi * im));
            cp[1] = __builtin_fma(br, im, bi * r);
        }
    }
}

template <int NL>
void launch_crt_nl(hipStream_t st, const CrtArgs& a) {
    const unsigned blocks = crt_blocks(a);
    if (a.cplx) hipLaunchKernelGGL((crt_kernel<NL, true>), dim3(blocks), dim3(256), 0, st, a);
    else hipLaunchKernelGGL((crt_kernel<NL, false>), dim3(blocks), dim3(256), 0, st, a);
}

}  // namespace

void crt(hipStream_t st, const CrtArgs& a) {
    if (a.mb <= 0 || a.nb <= 0) return;
    switch (a.L16 / 2) {
        case 2: launch_crt_nl<2>(st, a); break;
        case 3: launch_crt_nl<3>(st, a); break;
        case 4: launch_crt_nl<4>(st, a); break;
        case 5: launch_crt_nl<5>(st, a); break;
        case 6: launch_crt_nl<6>(st, a); break;
        default: break;
    }
}

// grouped CRT: one launch per (limb count, real/complex)
void prep_crt(std::vector<GroupedLaunch>& out, const std::vector<CrtArgs>& jobs) {
    GroupedLaunch by[14];
    for (const CrtArgs& a : jobs) {
        if (a.mb <= 0 || a.nb <= 0) continue;
        const int v = (a.L16 / 2) * 2 + (a.cplx ? 1 : 0);
        by[v].kind = v;
        by[v].add(a, crt_blocks(a));
    }
    for (GroupedLaunch& g : by)
        if (g.n) out.push_back(std::move(g));
}

void launch_crt_grouped(hipStream_t st, const GroupedLaunch& g) {
    if (g.n == 0) return;
    const CrtArgs* items = static_cast<const CrtArgs*>(g.d_items);
    switch (g.kind) {
#define OZ2_NL(NL)                                                                                        \
    case 2 * NL:                                                                                          \
        hipLaunchKernelGGL((crt_grouped_kernel<NL, false>), dim3(g.blocks), dim3(256), 0, st, items, g.d_start, g.n); \
        break;                                                                                            \
    case 2 * NL + 1:                                                                                      \
        hipLaunchKernelGGL((crt_grouped_kernel<NL, true>), dim3(g.blocks), dim3(256), 0, st, items, g.d_start, g.n); \
        break;
        OZ2_NL(2) OZ2_NL(3) OZ2_NL(4) OZ2_NL(5) OZ2_NL(6)
#undef OZ2_NL
        default: break;
    }
}

void scale_c(hipStream_t st, double* C, int64_t ldc, int64_t m, int64_t n, double br, double bi, bool cplx) {
    const int64_t total = m * n;
    if (total <= 0) return;
    hipLaunchKernelGGL(scale_c_kernel, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, st, C, ldc, m,
                       n, br, bi, cplx ? 1 : 0);
}

}  // namespace k
}  // namespace oz2
