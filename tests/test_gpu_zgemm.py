"""GPU parity tests of ZGEMM (3M form) through the C ABI, restating the
reference's complex test matrix (tests/test_ozaki2.cpp:491-506): results are
bit-identical to the oracle and within the complex rigorous bound."""
import numpy as np
import pytest

import oz2
from gpu_cases import bits, complex_case, log, make_cmat, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TC = ["N", "T", "C"]


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


@pytest.mark.parametrize("ta", TC)
@pytest.mark.parametrize("tb", TC)
def test_ops(oz, ta, tb):
    complex_case(oz, ta, tb, 37, 29, 101, 14, "phi", ld_extra=3)


def test_complex_alpha_beta(oz):
    complex_case(oz, "N", "C", 64, 48, 300, 16, "uniform", alpha=0.5 - 1.25j, beta=0.25 + 2.0j, ld_extra=1)


def test_c_only_8_byte_aligned(oz):
    # C one double past a 16-byte boundary, with and without beta: std::complex<double>
    # arrays are 16-byte aligned, a pointer into a double array need not be, and the
    # result must not depend on it
    complex_case(oz, "N", "N", 33, 21, 77, 12, "phi", alpha=1.5 - 0.5j, beta=-0.75 + 0.25j, ld_extra=1, c_shift=1)
    complex_case(oz, "C", "T", 40, 9, 130, 16, "uniform", c_shift=1)


@pytest.mark.parametrize("s", [4, 8, 12, 16, 20, 24])
def test_accuracy_vs_s(oz, s):
    complex_case(oz, "N", "N", 128, 128, 1024, s, "phi")


def test_integers_exact(oz):
    r = complex_case(oz, "N", "T", 60, 70, 900, 14, "ints")
    assert r["bound"] <= 1.0


def test_max_magnitude(oz):
    complex_case(oz, "N", "N", 32, 32, 2048, 12, "maxmag")
    complex_case(oz, "C", "N", 32, 32, 2048, 24, "maxmag")  # wide split (tk > 62)


def test_k_chunks(oz):
    complex_case(oz, "N", "N", 8, 12, 140000, 18, "uniform")
    assert oz.last_tiling().k_chunks == 3  # the plan is at 2k; the chunks follow k


def test_tiling(oz):
    r1 = complex_case(oz, "T", "C", 300, 260, 200, 10, "phi", check_bound=False)
    oz.set_workspace_limit(4 << 20)
    try:
        r2 = complex_case(oz, "T", "C", 300, 260, 200, 10, "phi", check_bound=False)
        ti = oz.last_tiling()
    finally:
        oz.set_workspace_limit(0)
    assert ti.pm * ti.pn > 1
    assert np.array_equal(bits(r1["C"]), bits(r2["C"]))


def test_alpha_zero_scales_c(oz):
    import torch
    rng = np.random.default_rng(0)
    C0 = rng.uniform(-1, 1, 2 * 6)
    dC = torch.from_numpy(C0.copy()).cuda()
    d = torch.zeros(64, dtype=torch.float64, device="cuda")
    oz.zgemm("N", "N", 2, 3, 4, 0j, d, 2, d, 4, 0.5 - 2j, dC, 2, 8)
    got = dC.cpu().numpy().view(np.complex128)
    want = C0.view(np.complex128) * (0.5 - 2j)
    assert np.allclose(got, want, rtol=0, atol=1e-15)
    O.precision_bits(8, 8)


def _sampled_complex(oz, ta, tb, m, n, k, s, dist, ld_extra=0, seed=13, count=256):
    """large ZGEMM (alpha = 1, beta = 0) checked at sampled entries against the closed form"""
    import torch
    rng = np.random.default_rng(seed)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    lda, ldb = ar + ld_extra, br + ld_extra
    A = make_cmat(ar, ac, lda, dist, rng, 1.0).reshape(-1, order="F").view(np.float64).copy()
    B = make_cmat(br, bc, ldb, dist, rng, 1.0).reshape(-1, order="F").view(np.float64).copy()
    log(f"case large zgemm {ta}{tb} m={m} n={n} k={k} s={s} {dist} ld+{ld_extra}")
    dC = to_dev(np.zeros(2 * m * n))
    oz.zgemm(ta, tb, m, n, k, 1.0, to_dev(A), lda, to_dev(B), ldb, 0.0, dC, m, s)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    rows = rng.integers(0, m, count)
    cols = rng.integers(0, n, count)
    rows[:4] = [0, m - 1, 0, m - 1]
    cols[:4] = [0, 0, n - 1, n - 1]
    want = O.gemm_entries(ta, tb, m, n, k, A, lda, B, ldb, s, rows, cols, cplx=True).reshape(-1, 2)
    got = np.stack([C[2 * (rows + cols * m)], C[2 * (rows + cols * m) + 1]], axis=1)
    assert np.array_equal(bits(got), bits(want))


def test_large_short_k_rectangular_tile(oz):
    # k loops up to 1024: 256 x 128 tiles; ragged m, n; padded leading dimensions (odd ld:
    # the contiguous-row split takes its element-wise path)
    _sampled_complex(oz, "C", "N", 1100, 900, 700, 16, "phi", ld_extra=7)
    _sampled_complex(oz, "N", "T", 1000, 1300, 1000, 14, "uniform")


def test_large_wide_split(oz):
    # s = 24: tk > 62, the wide complex split (half-chunk slots), long k: 256 x 256 tiles
    _sampled_complex(oz, "T", "C", 600, 700, 3000, 24, "phi", ld_extra=1)


@pytest.mark.slow
def test_full_size_configuration_sampled(oz):
    # the benchmark configuration at full size: ZGEMM 4096^3, op(A) = C, op(B) = N, s = 16,
    # alpha = (1, 0.5), beta = (0.25, -1), lda = m + 7, ldc = m + 3; sampled entries against
    # the exact closed form with the library's complex epilogue applied exactly on the host
    import torch
    from gpu_cases import fma
    m = n = k = 4096
    lda, ldb, ldc = m + 7, k, m + 3
    ar_, ai_, br_, bi_ = 1.0, 0.5, 0.25, -1.0
    rng = np.random.default_rng(19)
    A = rng.uniform(-1, 1, 2 * lda * m)  # op(A) = C: A is k x m, interleaved (re, im)
    B = rng.uniform(-1, 1, 2 * ldb * n)
    C0 = rng.uniform(-1, 1, 2 * ldc * n)
    log("case full-size zgemm CN 4096^3 s=16 alpha=(1,0.5) beta=(0.25,-1) lda=m+7 ldc=m+3")
    dC = to_dev(C0)
    oz.zgemm("C", "N", m, n, k, complex(ar_, ai_), to_dev(A), lda, to_dev(B), ldb, complex(br_, bi_), dC, ldc, 16)
    torch.cuda.synchronize()
    C = dC.cpu().numpy()
    rows = rng.integers(0, m, 192)
    cols = rng.integers(0, n, 192)
    rows[:4] = [0, m - 1, 0, m - 1]
    cols[:4] = [0, 0, n - 1, n - 1]
    v = O.gemm_entries("C", "N", m, n, k, A, lda, B, ldb, 16, rows, cols, cplx=True).reshape(-1, 2)
    want = []
    for (vr, vi), i, j in zip(v, rows, cols):
        cr, ci = C0[2 * (i + j * ldc)], C0[2 * (i + j * ldc) + 1]
        re = fma(ar_, float(vr), -(ai_ * float(vi))) + fma(br_, cr, -(bi_ * ci))
        im = fma(ar_, float(vi), ai_ * float(vr)) + fma(br_, ci, bi_ * cr)
        want.append((re, im))
    got = np.stack([C[2 * (rows + cols * ldc)], C[2 * (rows + cols * ldc) + 1]], axis=1)
    assert np.array_equal(bits(got), bits(np.array(want)))
