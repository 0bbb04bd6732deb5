"""Kernel-level GPU tests through the C ABI, the counterparts of the
reference's unit_split and unit_crt (tests/test_ozaki2.cpp:195-324):
  * split: every int8 residue equals the golden exact value (tests/golden/split.json)
    for both memory orders of op(X), and agrees with the oracle on random data;
  * CRT: every reconstructed value equals the golden RNE(x)*2^-shift (tests/golden/crt.json),
    including +-M/4, 0, ties at 53 bits and subnormal results.
"""
import json
import os
import struct

import numpy as np
import pytest
import torch

import oz2
from gpu_cases import log, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def h2f(h):
    return struct.unpack("<d", struct.pack("<Q", h))[0]


def plane_index(rows_pad, kp):
    """flat offset of (row, k) in one blocked plane: 16-row bands of 1 KB blocks (64 k each),
    64 bytes per row as four 16-byte chunks; chunk c holds the k pairs (2c, 2c+1) of each
    group of 8 k, and chunks are XOR-permuted by (row >> 2) & 3 (oz2_c.h, oz2_unit_split)"""
    r = np.arange(rows_pad)[:, None]
    p = np.arange(kp)[None, :]
    q = p & 63
    c = (q >> 1) & 3
    b = ((q >> 3) << 1) | (q & 1)
    return (((r >> 4) * (kp >> 6) + (p >> 6)) << 10) + ((r & 15) << 6) + ((c ^ ((r >> 2) & 3)) << 4) + b


def run_split(values, ex, s, tk, kcontig, cplx=False, conj=False):
    """values: rows x k (complex if cplx); returns planes [g][t][rows][k]"""
    rows, k = values.shape
    kp = (k + 127) // 128 * 128
    rows_pad = (rows + 31) // 32 * 32
    if kcontig:
        X = np.ascontiguousarray(values)  # row r contiguous: X[r*k + p]
        ldx = k
    else:
        X = np.asfortranarray(values).reshape(-1, order="F")  # X[r + p*rows]
        ldx = rows
    if cplx:
        X = np.ascontiguousarray(X).reshape(-1).view(np.float64)
    g = 3 if cplx else 1
    P = torch.zeros(g * s * rows_pad * kp, dtype=torch.int8, device="cuda")
    oz2.unit_split(to_dev(X.reshape(-1)), ldx, rows, k, kcontig, to_dev(np.asarray(ex, dtype=np.int32)), s, tk, P,
                   rows_pad, kp, cplx=cplx, conj=conj)
    torch.cuda.synchronize()
    flat = P.cpu().numpy().reshape(g, s, rows_pad * kp)
    idx = plane_index(rows_pad, kp)
    assert np.array_equal(np.sort(idx.reshape(-1)), np.arange(rows_pad * kp)), "layout must be a permutation"
    P = flat[:, :, idx]
    assert not P[:, :, rows:, :].any() and not P[:, :, :, k:].any(), "padding must be zero"
    return P[:, :, :rows, :k]


@pytest.mark.parametrize("kcontig", [True, False])
def test_split_golden(kcontig):
    for case in load("split.json"):
        s, tk = case["s"], case["tk"]
        log(f"unit split s={s} tk={tk} kcontig={kcontig}")
        cols = case["cols"]
        vals = np.array([[h2f(b) for b in c["bits"]] for c in cols])  # rows = columns of the golden table
        ex = [c["e"] for c in cols]
        P = run_split(vals, ex, s, tk, kcontig)[0]
        for r, c in enumerate(cols):
            assert P[:, r, :].tolist() == c["res"], (s, tk, r)


@pytest.mark.parametrize("conj", [False, True])
@pytest.mark.parametrize("kcontig", [True, False])
def test_split_complex_matches_oracle(kcontig, conj):
    rng = np.random.default_rng(5)
    for s, tk in ((14, 49), (24, 86)):
        rows, k = 40, 200
        v = (rng.uniform(-1, 1, (rows, k)) + 1j * rng.uniform(-1, 1, (rows, k))) * np.exp2(rng.integers(-8, 8, (rows, 1)))
        v[3, :] = 0
        mx = np.maximum(np.abs(v.real), np.abs(v.imag)).max(axis=1)
        ex = np.array([tk - np.frexp(x)[1] if x else 0 for x in mx], dtype=np.int32)
        P = run_split(v, ex, s, tk, kcontig, cplx=True, conj=conj)
        # oracle: residues of Re, Im (negated for conj) and their sum, reduced to [-h, m-1-h]
        ms = O.moduli(s)
        e_full = np.repeat(ex, k)
        re = O.split_values(v.real.reshape(-1), e_full, s).reshape(s, rows, k).astype(np.int64)
        im = O.split_values(v.imag.reshape(-1), e_full, s).reshape(s, rows, k).astype(np.int64)
        for t, m in enumerate(ms):
            h = (m - 1) // 2
            r0 = np.where(re[t] < -h, re[t] + m, re[t])
            r1 = np.where(im[t] < -h, im[t] + m, im[t])
            if conj:
                r1 = -r1
            ss = (((r0 + r1) % m) + h) % m - h
            for g, want in enumerate((r0, r1, ss)):
                got = P[g, t].astype(np.int64)
                assert np.array_equal(got & 0xFF, want & 0xFF), (s, t, g)


def test_crt_golden():
    for case in load("crt.json"):
        s = case["s"]
        res = np.array(case["res"], dtype=np.uint8)  # [n][s]
        n = res.shape[0]
        shift = np.array(case["shift"], dtype=np.int32)
        log(f"unit crt s={s} n={n}")
        ldr = (n + 3) // 4 * 4
        R = np.zeros((s, 1, ldr), dtype=np.uint8)
        R[:, 0, :n] = res.T
        D = shift - 3
        E = np.array([3], dtype=np.int32)
        dC = torch.zeros(n, dtype=torch.float64, device="cuda")
        oz2.unit_crt(to_dev(R.reshape(-1)), ldr, n, 1, to_dev(D.astype(np.int32)), to_dev(E), s, 1.0, 0.0, dC, n)
        torch.cuda.synchronize()
        got = dC.cpu().numpy().view(np.uint64)
        assert got.tolist() == [int(x) for x in case["expect"]], s


def test_crt_epilogue_and_nan():
    # alpha/beta epilogue fma(alpha, v, beta*C), beta == 0 never reads C, non-finite exponents give NaN
    s = 10
    plan = O.make_plan(s, 16)
    ms = O.moduli(s)
    rng = np.random.default_rng(9)
    n = 64
    M4 = plan.M_int // 4
    xs = [int(rng.integers(0, 2 ** 62)) * (1 if i % 2 else -1) for i in range(n)]
    xs[0] = M4
    res = np.array([[x % m for m in ms] for x in xs], dtype=np.uint8)
    R = np.zeros((s, 1, n), dtype=np.uint8)
    R[:, 0, :] = res.T
    D = rng.integers(-30, 30, n).astype(np.int32)
    D[5] = -2 ** 31
    E = np.array([2], dtype=np.int32)
    C0 = rng.uniform(-1, 1, n)
    for alpha, beta in ((1.0, 0.0), (-1.5, 0.25), (2.0, -3.0)):
        dC = to_dev(C0.copy())
        oz2.unit_crt(to_dev(R.reshape(-1)), n, n, 1, to_dev(D), to_dev(E), s, alpha, beta, dC, n)
        got = dC.cpu().numpy()
        for i in range(n):
            if i == 5:
                assert np.isnan(got[i])
                continue
            v = O.crt_value(plan, res[i], int(D[i]) + 2)
            want = alpha * v if beta == 0 else _fma(alpha, v, beta * C0[i])
            assert got[i].view(np.uint64) == np.float64(want).view(np.uint64), (alpha, beta, i)


def _fma(a, b, c):
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_accuracy_harness_rows():
    # the accuracy benchmark (reference benchmarks/bench_accuracy.cpp): CSV rows per method
    import io
    from oz2 import bench_accuracy
    for cplx in (False, True):
        buf = io.StringIO()
        argv = ["--k", "300", "--s", "6,16", "--dist", "urand", "--mn", "40", "--L", ""] + (["--complex"] if cplx else [])
        bench_accuracy.main(argv, out=buf)
        lines = buf.getvalue().strip().split("\n")
        assert lines[0] == "dist,k,method,s,tk,ulp_max,ulp_mean,cr_frac,ab_max,fro_rel"
        rows = [line.split(",") for line in lines[1:]]
        methods = [r[2] for r in rows]
        assert methods == (["zgemm_gpu", "ozaki2_3m", "ozaki2_3m"] if cplx else ["dgemm_gpu", "ozaki2", "ozaki2"])
        s16 = rows[2]
        assert s16[3] == "16" and float(s16[7]) == 1.0 and float(s16[5]) <= 0.5  # correctly rounded at s = 16
        assert float(rows[1][5]) > 1.0  # s = 6 keeps too few bits


@pytest.mark.parametrize("tile", [256, 128, 64, 256128])
def test_unit_gemm_against_int64_brute_force(tile):
    # the INT8 GEMM with its residue epilogue (the reference's oneMKL gemm_batch, hpp:1029-1033):
    # random int8 planes over the whole int8 range with ragged m, n, k (zero padding up to the macro
    # tile and to kp), every residue byte against (sum of int64 products) mod m_t
    rng = np.random.default_rng(tile)
    tm, tn = (256, 128) if tile == 256128 else (tile, tile)
    cases = ((tm + 37, 2 * tn - 5, 300, 16), (tm - 3, tn + 1, 1, 5), (2 * tm + 1, 3 * tn - 1, 2500, 24),
             (tm + 5, tn - 7, 256, 7))
    for m, n, k, s in cases:
        mpad, npad = -(-m // tm) * tm, -(-n // tn) * tn
        kp = -(-k // 128) * 128
        ldr = mpad + 16
        log(f"unit gemm tile={tile} m={m} n={n} k={k} s={s} (mpad={mpad} npad={npad} kp={kp})")
        a = np.zeros((s, mpad, kp), dtype=np.int8)
        b = np.zeros((s, npad, kp), dtype=np.int8)
        a[:, :m, :k] = rng.integers(-128, 128, (s, m, k))
        b[:, :n, :k] = rng.integers(-128, 128, (s, n, k))
        # the first planes at the extremes: all -128 (the largest sum; for k <= 256 the top of
        # the epilogue's short form), and -128 against 127 (its most negative sum)
        a[0:2, :m, :k] = -128
        b[0:2, :n, :k] = -128
        a[2, :m, :k] = -128
        b[2, :n, :k] = 127
        ia, ib = plane_index(mpad, kp).reshape(-1), plane_index(npad, kp).reshape(-1)
        pa = np.zeros((s, mpad * kp), dtype=np.int8)
        pb = np.zeros((s, npad * kp), dtype=np.int8)
        pa[:, ia] = a.reshape(s, -1)
        pb[:, ib] = b.reshape(s, -1)
        R = torch.full((s * npad * ldr,), 0xEE, dtype=torch.uint8, device="cuda")
        oz2.unit_gemm(to_dev(pa.reshape(-1)), to_dev(pb.reshape(-1)), mpad, npad, kp, s, tile, R, ldr)
        torch.cuda.synchronize()
        got = R.cpu().numpy().reshape(s, npad, ldr)
        mods = O.moduli(s)
        for t in range(s):
            # |sums| < 2^30: the float64 products and sums of these integers are exact
            prod = (a[t].astype(np.float64) @ b[t].astype(np.float64).T).astype(np.int64)
            want = np.mod(prod, int(mods[t])).astype(np.uint8)
            assert np.array_equal(got[t, :, :mpad], want.T), f"tile {tile} plane {t} (m_t = {mods[t]})"
            assert (got[t, :, mpad:] == 0xEE).all(), "bytes past mpad in each column must stay untouched"


def test_unit_gemm_argument_errors():
    P = torch.zeros(1 << 16, dtype=torch.int8, device="cuda")
    R = torch.zeros(1 << 16, dtype=torch.uint8, device="cuda")
    for args in ((100, 128, 128, 2, 128, 128),     # mpad not a tile multiple
                 (128, 128, 100, 2, 128, 128),     # kp not a multiple of 128
                 (128, 128, 65536, 2, 128, 128),   # more than one k-chunk
                 (128, 128, 128, 2, 128, 120),     # ldr below mpad
                 (128, 128, 128, 25, 128, 128),    # num_moduli
                 (128, 128, 128, 2, 96, 128)):     # tile code
        mpad, npad, kp, s, tile, ldr = args
        with pytest.raises(oz2.OZ2InvalidArgument):
            oz2.unit_gemm(P, P, mpad, npad, kp, s, tile, R, ldr)
