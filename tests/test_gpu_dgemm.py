"""GPU parity tests of DGEMM through the C ABI, restating the reference's test
matrix (tests/test_ozaki2.cpp:448-549): every result must be bit-identical to
the oracle on the same inputs and satisfy the rigorous element-wise bound.
Large shapes are checked on sampled entries against the exact closed form."""
import numpy as np
import pytest
import torch

import oz2
from conftest import make_mat
from gpu_cases import bits, flat, log, real_case, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

OPS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


@pytest.mark.parametrize("ta,tb", OPS)
def test_shapes_and_ops(oz, ta, tb):
    # test:449-455
    real_case(oz, ta, tb, 1, 1, 1, 8, "uniform")
    real_case(oz, ta, tb, 37, 53, 129, 12, "phi", lda_extra=3, ldc_extra=5)
    real_case(oz, ta, tb, 130, 7, 1, 6, "phi", phi=2.0)
    real_case(oz, ta, tb, 77, 65, 300, 14, "uniform", alpha=-1.5, beta=0.25, ldc_extra=2)


@pytest.mark.parametrize("ta,tb", OPS)
def test_odd_leading_dimension_long_rows(oz, ta, tb):
    # lda = ldb = k + 1 (or m/n + 1) odd: every other k-contiguous row is only 8-byte aligned, and
    # rows of 1000 and 3001 doubles run the scale's many-loads-in-flight loop and its remainder
    real_case(oz, ta, tb, 40, 33, 1000, 14, "phi", lda_extra=1, ldc_extra=1)
    real_case(oz, ta, tb, 9, 6, 3001, 16, "uniform", beta=0.5, lda_extra=2)


def test_accuracy_vs_s(oz):
    # test:457-463: every s that the plan accepts
    for s in range(2, 25):
        if O.precision_bits(1024, s) < 0:
            with pytest.raises(oz2.OZ2InvalidArgument):
                real_case(oz, "N", "N", 256, 256, 1024, s, "phi", check_oracle=False, check_bound=False)
            continue
        real_case(oz, "N", "N", 256, 256, 1024, s, "phi")


@pytest.mark.parametrize("phi", [0.5, 2.0, 4.0])
def test_dynamic_range(oz, phi):
    real_case(oz, "N", "N", 200, 300, 2000, 16, "phi", phi=phi)


def test_integers_exact(oz):
    # test:468-469: integer inputs give the exact product
    for ta, tb, m, n, k, s, pad in (("N", "N", 100, 120, 1000, 12, 0), ("T", "T", 63, 31, 777, 14, 1)):
        r = real_case(oz, ta, tb, m, n, k, s, "ints", lda_extra=pad, ldc_extra=pad)
        assert np.array_equal(r["lo"], np.zeros_like(r["lo"]))
        C = r["C"].reshape(n, m + pad)[:, :m].T.reshape(-1, order="F")
        assert np.array_equal(C, r["hi"])


@pytest.mark.parametrize("s", [3, 4, 8, 12, 16, 20, 24])
def test_max_magnitude(oz, s):
    # test:470-471: |C'| ~ M/4, the CRT margin
    real_case(oz, "N", "N", 64, 64, 4096, s, "maxmag")


def test_tiling_bit_identity(oz):
    # test:473-486
    r1 = real_case(oz, "N", "T", 1000, 900, 512, 10, "phi", seed=5, check_bound=False)
    oz.set_workspace_limit(8 << 20)
    try:
        r2 = real_case(oz, "N", "T", 1000, 900, 512, 10, "phi", seed=5, check_bound=False)
        ti = oz.last_tiling()
    finally:
        oz.set_workspace_limit(0)
    log(f"tiling {ti.as_dict()}")
    assert ti.pm * ti.pn > 1
    assert np.array_equal(bits(r1["C"]), bits(r2["C"]))


def test_workspace_too_small(oz):
    oz.set_workspace_limit(1 << 10)
    try:
        with pytest.raises(oz2.OZ2RuntimeError, match="not enough device memory"):
            real_case(oz, "N", "N", 64, 64, 64, 8, "uniform", check_oracle=False, check_bound=False)
    finally:
        oz.set_workspace_limit(0)


def test_k_chunks(oz):
    # test:488-489: long k splits the int32 accumulation (here into equal chunks of at most 65408)
    real_case(oz, "N", "N", 16, 24, 300000, 18, "phi", phi=0.5)
    assert oz.last_tiling().k_chunks == 5
    real_case(oz, "T", "N", 8, 8, 140000, 18, "uniform")
    assert oz.last_tiling().k_chunks == 3
    real_case(oz, "N", "T", 40, 36, 65409, 16, "phi")  # one k past a chunk: two chunks of 32768
    assert oz.last_tiling().k_chunks == 2


def test_non_finite_propagation(oz):
    # test:522-543
    m, n, k = 32, 32, 64
    A = np.full(m * k, 0.5)
    B = np.full(k * n, 0.25)
    A[3 + 10 * m] = np.inf
    B[5 + 7 * k] = np.nan
    dC = to_dev(np.zeros(m * n))
    oz.gemm("N", "N", m, n, k, 1.0, to_dev(A), m, to_dev(B), k, 0.0, dC, m, 8)
    C = dC.cpu().numpy().reshape(n, m).T
    for i in range(m):
        for j in range(n):
            if i == 3 or j == 7:
                assert np.isnan(C[i, j])
            else:
                assert C[i, j] == 0.5 * 0.25 * k


def test_argument_errors(oz):
    d = to_dev(np.zeros(64))
    with pytest.raises(oz2.OZ2InvalidArgument, match="leading dimension"):
        oz.gemm("N", "N", 4, 4, 4, 1.0, d, 3, d, 4, 0.0, d, 4, 8)
    with pytest.raises(oz2.OZ2InvalidArgument, match="negative"):
        oz.gemm("N", "N", -1, 4, 4, 1.0, d, 4, d, 4, 0.0, d, 4, 8)
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz.gemm("N", "N", 4, 4, 4, 1.0, d, 4, d, 4, 0.0, d, 4, 25)
    with pytest.raises(oz2.OZ2InvalidArgument, match="too small"):
        oz.gemm("N", "N", 4, 4, 5000, 1.0, d, 4, d, 5000, 0.0, d, 4, 2)
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz.gemm("X", "N", 4, 4, 4, 1.0, d, 4, d, 4, 0.0, d, 4, 8)


def test_quick_returns_and_scale_c(oz):
    # m or n == 0: no-op; k == 0 or alpha == 0: C = beta*C, exact zero for beta == 0 (hpp:750-751, 1056-1070)
    C0 = np.arange(1.0, 13.0)
    dC = to_dev(C0.copy())
    d = to_dev(np.ones(16))
    oz.gemm("N", "N", 0, 3, 4, 1.0, d, 1, d, 4, 2.0, dC, 4, 8)
    oz.gemm("N", "N", 4, 0, 4, 1.0, d, 4, d, 4, 2.0, dC, 4, 8)
    assert np.array_equal(dC.cpu().numpy(), C0)
    oz.gemm("N", "N", 4, 3, 0, 1.0, d, 4, d, 1, 2.0, dC, 4, 8)
    assert np.array_equal(dC.cpu().numpy(), 2 * C0)
    oz.gemm("N", "N", 4, 3, 4, 0.0, d, 4, d, 4, 0.0, dC, 4, 8)
    assert np.array_equal(dC.cpu().numpy(), np.zeros(12))


def test_trans_c_equals_t_for_real(oz):
    a = real_case(oz, "T", "T", 20, 30, 40, 10, "phi", check_bound=False)
    b = real_case(oz, "C", "C", 20, 30, 40, 10, "phi", check_bound=False, check_oracle=False)
    assert np.array_equal(bits(a["C"]), bits(b["C"]))


def test_stage_times_and_tiling(oz):
    real_case(oz, "N", "N", 512, 512, 512, 16, "uniform", check_bound=False)
    st = oz.stage_times()
    assert st.gemm_ms > 0 and st.split_ms > 0 and st.crt_ms > 0 and st.scale_ms >= 0
    ti = oz.last_tiling()
    assert (ti.mb, ti.nb, ti.pm, ti.pn, ti.k_chunks) == (512, 512, 1, 1, 1)
    assert ti.workspace_bytes > 3 * 16 * 512 * 512


@pytest.mark.parametrize("tile", [256, 128, 64, 256128])
def test_every_gemm_tile_bit_identical(oz, tile):
    oz.set_tuning(gemm_tile=tile)
    try:
        real_case(oz, "T", "N", 300, 260, 700, 13, "phi", check_bound=False)
    finally:
        oz.set_tuning()


def _sampled_large(oz, ta, tb, m, n, k, s, dist, seed=11, count=384):
    rng = np.random.default_rng(seed)
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = flat(make_mat(ar, ac, ar, dist, rng))
    B = flat(make_mat(br, bc, br, dist, rng))
    log(f"case large dgemm {ta}{tb} m={m} n={n} k={k} s={s} {dist}")
    dC = to_dev(np.zeros(m * n))
    oz.gemm(ta, tb, m, n, k, 1.0, to_dev(A), ar, to_dev(B), br, 0.0, dC, m, s)
    C = dC.cpu().numpy()
    rows = rng.integers(0, m, count)
    cols = rng.integers(0, n, count)
    rows[:4] = [0, m - 1, 0, m - 1]
    cols[:4] = [0, 0, n - 1, n - 1]
    want = O.gemm_entries(ta, tb, m, n, k, A, ar, B, br, s, rows, cols)
    got = C[rows + cols * m]
    assert np.array_equal(bits(got), bits(want))


@pytest.mark.slow
def test_large_cases(oz):
    # test:545-549, checked on sampled entries against the exact closed form
    _sampled_large(oz, "N", "N", 2048, 2048, 4096, 16, "sincos")
    _sampled_large(oz, "N", "N", 3000, 2500, 4000, 8, "phi")
    _sampled_large(oz, "T", "N", 2500, 3000, 4000, 20, "phi")


def test_short_k_rectangular_tile(oz):
    # k loops up to 1024 with enough workgroups take the 256 x 128 macro tile (two per CU);
    # ragged m and n against the tile, sampled entries against the exact closed form
    _sampled_large(oz, "N", "T", 1500, 1300, 500, 16, "phi")
    _sampled_large(oz, "T", "N", 1100, 2100, 1000, 14, "sincos")


def test_stream_and_events(oz):
    # deps / done events: the call waits on deps and records done on its stream
    st = torch.cuda.Stream()
    h = oz2.Ozaki2(st)
    try:
        m = n = k = 128
        rng = np.random.default_rng(3)
        A = rng.uniform(-1, 1, m * k)
        B = rng.uniform(-1, 1, k * n)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(np.zeros(m * n))
        ready = torch.cuda.Event()
        ready.record()
        done = torch.cuda.Event()
        done.record()  # creates the event; the library records it again on its stream
        h.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, 14, deps=[ready.cuda_event], done=done.cuda_event)
        done.synchronize()
        ref = np.zeros(m * n)
        O.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, ref, m, 14)
        assert np.array_equal(bits(dC.cpu().numpy()), bits(ref))
    finally:
        h.close()


def test_cpu_baseline_configuration(oz):
    # the reference's CPU-runnable configuration: DGEMM 1024^3 N/N, alpha = 1, beta = 0,
    # U[-1,1), s = 14 (tk = 49), every entry against the oracle and the rigorous bound
    r = real_case(oz, "N", "N", 1024, 1024, 1024, 14, "u11")
    assert r["bound"] <= 1.0


@pytest.fixture(scope="module")
def full8192():
    # the benchmark configuration's inputs: A, B, C0 U[-1,1) of 8192 x 8192 (host and device)
    m = 8192
    rng = np.random.default_rng(17)
    A = rng.uniform(-1, 1, m * m)
    B = rng.uniform(-1, 1, m * m)
    C0 = rng.uniform(-1, 1, m * m)
    return {"A": A, "B": B, "C0": C0, "dA": to_dev(A), "dB": to_dev(B)}


@pytest.mark.slow
@pytest.mark.parametrize("ta,tb,s", [("N", "N", 16), ("N", "T", 16), ("T", "N", 16), ("T", "T", 16),
                                     ("N", "N", 8), ("N", "N", 12)])
def test_full_size_configuration_sampled(oz, full8192, ta, tb, s):
    # the benchmark configuration at full size: DGEMM 8192^3, every op(A)/op(B) at s = 16 and
    # N/N at s = 8 and 12, alpha = 1.5, beta = 0.5, C0 U[-1,1); sampled entries (corners
    # included) against the exact closed form with the library's epilogue fma(alpha, v, beta*c)
    # applied exactly on the host
    from gpu_cases import fma
    m = n = k = 8192
    F = full8192
    rng = np.random.default_rng(s * 4 + "NT".index(ta) * 2 + "NT".index(tb))
    log(f"case full-size dgemm {ta}{tb} 8192^3 s={s} alpha=1.5 beta=0.5")
    dC = to_dev(F["C0"])
    oz.gemm(ta, tb, m, n, k, 1.5, F["dA"], m, F["dB"], k, 0.5, dC, m, s)
    C = dC.cpu().numpy()
    del dC
    rows = rng.integers(0, m, 256)
    cols = rng.integers(0, n, 256)
    rows[:4] = [0, m - 1, 0, m - 1]
    cols[:4] = [0, 0, n - 1, n - 1]
    v = O.gemm_entries(ta, tb, m, n, k, F["A"], m, F["B"], k, s, rows, cols)
    want = np.array([fma(1.5, float(x), 0.5 * F["C0"][i + j * m]) for x, i, j in zip(v, rows, cols)])
    assert np.array_equal(bits(C[rows + cols * m]), bits(want))


@pytest.mark.slow
def test_selector_configurations_sampled(oz):
    # the selector benchmarks at full size, at the s the selector picks there (24): the wide-exponent
    # DGEMM 4000x4000 (k = 8000) and the k-split DGEMM 2048x2048 (k = 200000, four k-chunks of at most
    # 65408), entries (1 + U) * (+-1) * 2^e with e in [-20, 20]; sampled entries against the closed form
    _sampled_large(oz, "N", "N", 4000, 4000, 8000, 24, "wide", seed=23, count=256)
    _sampled_large(oz, "N", "N", 2048, 2048, 200000, 24, "wide", seed=29, count=96)
    assert oz.last_tiling().k_chunks == 4
