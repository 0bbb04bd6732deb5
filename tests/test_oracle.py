"""Pins the CPU oracle (oracle/oz2_oracle.c) before anything is compared with it.

Each oracle function is checked against something other than itself:
  moduli, make_plan, moduli_for_bits  -> tests/golden/plan.json (exact Python integers)
  scaled split                         -> tests/golden/split.json and brute force here
  CRT                                  -> tests/golden/crt.json (exact RNE of Python integers)
  residue-route DGEMM/ZGEMM            -> tests/golden/gemm.json (closed form in Python integers),
                                          the oracle's own 256-bit closed form, integer exactness,
                                          the non-finite known answer (tests/test_ozaki2.cpp:522-543)
  double-double reference              -> exact Fraction sums
  rigorous bound                       -> the oracle result passes; planted mistakes fail
  metrics                              -> hand-computed values
"""
import json
import math
import os
import struct
from fractions import Fraction

import numpy as np
import pytest

from conftest import make_cmat, make_mat
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def h2f(h):
    return struct.unpack("<d", struct.pack("<Q", h))[0]


def f2h(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


def fbits(arr):
    return np.asarray(arr, dtype=np.float64).view(np.uint64)


# ---------------------------------------------------------------- plan
def test_plan_matches_golden():
    g = load("plan.json")
    for p in g["plans"]:
        if p["tk"] is None:
            with pytest.raises(ValueError):
                O.make_plan(p["s"], p["k"])
            continue
        q = O.make_plan(p["s"], p["k"])
        assert O.moduli(p["s"]) == p["moduli"]
        assert q.tk == p["tk"], (p["s"], p["k"])
        assert q.NL == p["NL"]
        assert q.M_int == int(p["M"], 16)
        for t in range(p["s"]):
            assert q.m[t] == p["moduli"][t]
            assert q.y[t] == p["y"][t]
            assert q.W_int(t) == int(p["W"][t], 16)


def test_plan_closed_form_properties():
    # tk is the largest t with k*2^(2t+2) <= M; pairwise coprime; W_t = 1 mod m_t, 0 mod m_u
    for s in range(2, 25):
        ms = O.moduli(s)
        for a in range(s):
            for b in range(a + 1, s):
                assert math.gcd(ms[a], ms[b]) == 1
        for k in (1, 7, 1000, 65536, 300000):
            try:
                p = O.make_plan(s, k)
            except ValueError:
                M = math.prod(ms)
                assert k * 2 ** 4 > M  # no t >= 1 fits
                continue
            M = p.M_int
            assert k * 2 ** (2 * p.tk + 2) <= M
            assert p.tk == 100 or k * 2 ** (2 * p.tk + 4) > M
            for t in range(s):
                W = p.W_int(t)
                for u in range(s):
                    assert W % ms[u] == (1 if u == t else 0)


def test_appendix_values():
    # derived plan table: tk at k=8192 for s = 8/12/16 is 24/39/55; complex plan at 2k
    assert [O.precision_bits(8192, s) for s in (8, 12, 16)] == [24, 39, 55]
    assert O.precision_bits(1024, 14) == 49
    assert O.precision_bits(2 * 4096, 16) == 55
    assert O.precision_bits(200000, 16) == 53
    assert O.moduli_for_bits(1024, 53) == 15
    assert all(O.moduli_for_bits(k, 53) == 16 for k in (2048, 8000, 131071, 200000))
    assert O.precision_bits(1024, 2) == 1 and O.precision_bits(4096, 2) == -1  # s too small -> rejected


def test_moduli_for_bits_golden():
    for r in load("plan.json")["moduli_for_bits"]:
        assert O.moduli_for_bits(r["k"], r["bits"]) == r["s"], r


# ---------------------------------------------------------------- exponents
def test_exponent_rules():
    tk = 40
    assert O.exponent_from_maxbits(0, tk) == 0
    assert O.exponent_from_maxbits(f2h(math.inf), tk) == -2 ** 31
    assert O.exponent_from_maxbits(f2h(math.nan) & 0x7FFFFFFFFFFFFFFF, tk) == -2 ** 31
    for v in (1.0, 0.75, 3.0, 2.0 ** -1074, 1e300):
        assert O.exponent_from_maxbits(f2h(v), tk) == tk - math.frexp(v)[1]


# ---------------------------------------------------------------- split
def test_split_matches_golden():
    for case in load("split.json"):
        s = case["s"]
        for col in case["cols"]:
            v = np.array([h2f(b) for b in col["bits"]])
            e = np.full(v.size, col["e"], dtype=np.int32)
            for i in range(0, v.size, 7):
                assert O.scaled_int(v[i], col["e"]) == int(col["x"][i])
            planes = O.split_values(v, e, s)
            assert planes.tolist() == col["res"], (s, case["tk"])


def test_split_brute_force(rng):
    # RNE(v * 2^e) with Fractions, residue by Python %, for random values and exponents
    for s in (5, 13, 24):
        ms = O.moduli(s)
        v = rng.uniform(-1, 1, 300) * np.exp2(rng.integers(-60, 5, 300).astype(float))
        v[:5] = [0.0, -0.0, 5e-324, -2.5, 0.5]
        e = rng.integers(0, 90, 300).astype(np.int32)
        planes = O.split_values(v, e, s)
        for i in range(v.size):
            x = round(Fraction(float(v[i])) * Fraction(2) ** int(e[i]))
            if abs(x) >= 2 ** 100:
                continue
            for t, m in enumerate(ms):
                h = (m - 1) // 2
                r = ((x % m) + h) % m - h
                assert planes[t, i] == ((r + 128) % 256) - 128


# ---------------------------------------------------------------- CRT
def test_crt_matches_golden():
    for case in load("crt.json"):
        s = case["s"]
        plan = O.make_plan(s, 16)
        for res, sh, ex in zip(case["res"], case["shift"], case["expect"]):
            got = O.crt_value(plan, res, sh)
            assert f2h(got) == ex, (s, sh)


# ---------------------------------------------------------------- end to end
def _run_golden_case(c, closed):
    A = np.array([h2f(b) for b in c["A"]])
    B = np.array([h2f(b) for b in c["B"]])
    m, n, k, s = c["m"], c["n"], c["k"], c["s"]
    if c["kind"] == "d":
        C = np.zeros(m * n)
        if closed:
            O.dgemm_closed(c["ta"], c["tb"], m, n, k, A, c["lda"], B, c["ldb"], C, m, s)
        else:
            O.dgemm(c["ta"], c["tb"], m, n, k, 1.0, A, c["lda"], B, c["ldb"], 0.0, C, m, s)
    else:
        C = np.zeros(2 * m * n)
        if closed:
            O.zgemm_closed(c["ta"], c["tb"], m, n, k, A, c["lda"], B, c["ldb"], C, m, s)
        else:
            O.zgemm(c["ta"], c["tb"], m, n, k, 1 + 0j, A, c["lda"], B, c["ldb"], 0j, C, m, s)
    return C


@pytest.mark.parametrize("closed", [False, True])
def test_gemm_matches_golden(closed):
    for c in load("gemm.json"):
        C = _run_golden_case(c, closed)
        assert fbits(C).tolist() == [int(x) for x in c["C"]], (c["kind"], c["ta"], c["tb"], c["m"], c["n"], c["k"])


OPS = [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")]


def _flat(X):
    return np.asfortranarray(X).reshape(-1, order="F")


@pytest.mark.parametrize("ta,tb", OPS)
def test_residue_route_equals_closed_form(ta, tb, rng):
    for (m, n, k, s, dist, pad) in [(37, 53, 129, 12, "phi", 3), (33, 9, 300, 20, "wide", 1), (16, 16, 64, 24, "u11", 0)]:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        A = _flat(make_mat(ar, ac, ar + pad, dist, rng))
        B = _flat(make_mat(br, bc, br + pad, dist, rng))
        C1 = np.zeros(m * n)
        C2 = np.zeros(m * n)
        O.dgemm(ta, tb, m, n, k, 1.0, A, ar + pad, B, br + pad, 0.0, C1, m, s)
        O.dgemm_closed(ta, tb, m, n, k, A, ar + pad, B, br + pad, C2, m, s)
        assert np.array_equal(fbits(C1), fbits(C2))


def test_k_chunks_equal_closed_form(rng):
    # k > 131071 splits the int32 accumulation into chunks (hpp:944-945, 590-601)
    m, n, k, s = 4, 3, 140000, 18
    A = _flat(make_mat(k, m, k, "u11", rng))
    B = _flat(make_mat(k, n, k, "u11", rng))
    C1 = np.zeros(m * n)
    C2 = np.zeros(m * n)
    O.dgemm("T", "N", m, n, k, 1.0, A, k, B, k, 0.0, C1, m, s)
    O.dgemm_closed("T", "N", m, n, k, A, k, B, k, C2, m, s)
    assert np.array_equal(fbits(C1), fbits(C2))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "N"), ("T", "C"), ("C", "C")])
def test_zgemm_residue_route_equals_closed_form(ta, tb, rng):
    m, n, k, s, pad = 9, 7, 51, 14, 2
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = make_cmat(ar, ac, ar + pad, "phi", rng).reshape(-1, order="F").view(np.float64)
    B = make_cmat(br, bc, br + pad, "phi", rng).reshape(-1, order="F").view(np.float64)
    C1 = np.zeros(2 * m * n)
    C2 = np.zeros(2 * m * n)
    O.zgemm(ta, tb, m, n, k, 1 + 0j, A, ar + pad, B, br + pad, 0j, C1, m, s)
    O.zgemm_closed(ta, tb, m, n, k, A, ar + pad, B, br + pad, C2, m, s)
    assert np.array_equal(fbits(C1), fbits(C2))


def test_integers_exact(rng):
    m, n, k = 100, 120, 1000
    A = rng.integers(-1000, 1001, (m, k)).astype(np.float64)
    B = rng.integers(-1000, 1001, (k, n)).astype(np.float64)
    C = np.zeros(m * n)
    O.dgemm("N", "N", m, n, k, 1.0, _flat(A), m, _flat(B), k, 0.0, C, m, 12)
    exact = (A.astype(np.int64) @ B.astype(np.int64)).astype(np.float64)
    assert np.array_equal(C.reshape(n, m).T, exact)


def test_non_finite_propagation():
    m, n, k = 32, 32, 64
    A = np.full(m * k, 0.5)
    B = np.full(k * n, 0.25)
    A[3 + 10 * m] = np.inf
    B[5 + 7 * k] = np.nan
    C = np.zeros(m * n)
    O.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, 8)
    C = C.reshape(n, m).T
    for i in range(m):
        for j in range(n):
            if i == 3 or j == 7:
                assert np.isnan(C[i, j])
            else:
                assert C[i, j] == 0.5 * 0.25 * k


def test_argument_errors():
    A = np.zeros(16)
    with pytest.raises(ValueError):
        O.dgemm("N", "N", 4, 4, 4, 1.0, A, 3, A, 4, 0.0, A, 4, 8)  # lda < m
    with pytest.raises(ValueError):
        O.dgemm("N", "N", -1, 4, 4, 1.0, A, 4, A, 4, 0.0, A, 4, 8)
    with pytest.raises(ValueError):
        O.dgemm("N", "N", 4, 4, 4, 1.0, A, 4, A, 4, 0.0, A, 4, 25)  # s out of range
    with pytest.raises(ValueError):
        O.dgemm("N", "N", 4, 4, 4, 1.0, A, 4, A, 4, 0.0, A, 4, 1)
    big = np.zeros(5000 * 4)
    with pytest.raises(ValueError):
        O.dgemm("T", "N", 4, 4, 5000, 1.0, big, 5000, big, 5000, 0.0, A, 4, 2)  # s too small for k


def test_k0_alpha0_scale_c():
    C = np.arange(1.0, 7.0)
    A = np.ones(6)
    O.dgemm("N", "N", 2, 3, 0, 1.0, A, 2, A, 1, 2.0, C, 2, 8)
    assert C.tolist() == [2.0, 4.0, 6.0, 8.0, 10.0, 12.0]
    O.dgemm("N", "N", 2, 3, 1, 0.0, A, 2, A, 1, 0.0, C, 2, 8)
    assert C.tolist() == [0.0] * 6


# ---------------------------------------------------------------- dd reference, bound, metrics
def test_dd_reference_is_exact_to_106_bits(rng):
    m, n, k = 5, 4, 300
    A = rng.uniform(-1, 1, (m, k)) * np.exp2(rng.integers(-30, 30, (m, k)))
    B = rng.uniform(-1, 1, (k, n))
    hi, lo = O.dd_gemm("N", "N", m, n, k, _flat(A), m, _flat(B), k)
    for j in range(n):
        for i in range(m):
            exact = sum(Fraction(float(A[i, p])) * Fraction(float(B[p, j])) for p in range(k))
            absab = sum(abs(float(A[i, p]) * float(B[p, j])) for p in range(k))
            err = abs(Fraction(hi[i + j * m]) + Fraction(lo[i + j * m]) - exact)
            assert err <= Fraction(2.0 ** -100) * Fraction(absab)


def _bound_case(ta, tb, m, n, k, s, dist, alpha, beta, rng, phi=1.0, pad=0):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = _flat(make_mat(ar, ac, ar + pad, dist, rng, phi))
    B = _flat(make_mat(br, bc, br + pad, dist, rng, phi))
    ldc = m + pad
    C0 = rng.uniform(-1, 1, ldc * n)
    C = C0.copy()
    O.dgemm(ta, tb, m, n, k, alpha, A, ar + pad, B, br + pad, beta, C, ldc, s)
    hi, lo = O.dd_gemm(ta, tb, m, n, k, A, ar + pad, B, br + pad)
    r = O.bound_ratio(ta, tb, m, n, k, A, ar + pad, B, br + pad, alpha, beta, C0, C, ldc, hi, lo, s)
    pad_ok = all(C[i + j * ldc] == C0[i + j * ldc] for j in range(n) for i in range(m, ldc))
    return r, pad_ok, (A, B, C0, C, hi, lo, ldc)


@pytest.mark.parametrize("ta,tb", OPS)
def test_bound_shapes_and_ops(ta, tb, rng):
    for (m, n, k, s, dist, alpha, beta, pad) in [(1, 1, 1, 8, "uniform", 1.0, 0.0, 0), (37, 53, 129, 12, "phi", 1.0, 0.0, 3),
                                                  (130, 7, 1, 6, "phi", 1.0, 0.0, 0),
                                                  (77, 65, 300, 14, "uniform", -1.5, 0.25, 2)]:
        r, pad_ok, _ = _bound_case(ta, tb, m, n, k, s, dist, alpha, beta, rng, phi=1.0 if k != 1 else 2.0, pad=pad)
        assert r <= 1.0 and pad_ok, (m, n, k, s)


def test_bound_accuracy_vs_s(rng):
    for s in range(2, 25):
        if O.precision_bits(256, s) < 0:
            continue
        r, _, _ = _bound_case("N", "N", 64, 64, 256, s, "phi", 1.0, 0.0, rng)
        assert r <= 1.0, s


def test_bound_max_magnitude(rng):
    # |C'| ~ M/4: the CRT margin case
    for s in (3, 4, 8, 12, 16, 20, 24):
        r, _, _ = _bound_case("N", "N", 16, 16, 1024, s, "maxmag", 1.0, 0.0, rng)
        assert r <= 1.0, s


def test_bound_detects_planted_mistakes(rng):
    r, _, (A, B, C0, C, hi, lo, ldc) = _bound_case("N", "N", 40, 30, 200, 16, "uniform", 1.0, 0.0, rng)
    assert r <= 1.0
    bad = C.copy()
    bad[17] = np.nextafter(np.nextafter(np.nextafter(bad[17], np.inf), np.inf), np.inf) * (1 + 2.0 ** -45)
    assert O.bound_ratio("N", "N", 40, 30, 200, A, 40, B, 200, 1.0, 0.0, C0, bad, ldc, hi, lo, 16) > 1.0
    # a result computed with fewer moduli than claimed breaks the bound
    C4 = C0.copy()
    O.dgemm("N", "N", 40, 30, 200, 1.0, A, 40, B, 200, 0.0, C4, ldc, 6)
    assert O.bound_ratio("N", "N", 40, 30, 200, A, 40, B, 200, 1.0, 0.0, C0, C4, ldc, hi, lo, 16) > 1.0


def test_zgemm_bound(rng):
    for ta in ("N", "T", "C"):
        for tb in ("N", "T", "C"):
            m, n, k, s, pad = 11, 9, 41, 14, 3
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            A = make_cmat(ar, ac, ar + pad, "phi", rng).reshape(-1, order="F").view(np.float64)
            B = make_cmat(br, bc, br + pad, "phi", rng).reshape(-1, order="F").view(np.float64)
            ldc = m + pad
            C0 = rng.uniform(-1, 1, 2 * ldc * n)
            for alpha, beta in ((1 + 0j, 0j), (0.5 - 1.25j, 0.25 + 2j)):
                C = C0.copy()
                O.zgemm(ta, tb, m, n, k, alpha, A, ar + pad, B, br + pad, beta, C, ldc, s)
                hi, lo = O.dd_zgemm(ta, tb, m, n, k, A, ar + pad, B, br + pad)
                r = O.zbound_ratio(ta, tb, m, n, k, A, ar + pad, B, br + pad, alpha, beta, C0, C, ldc, hi, lo, s)
                assert r <= 1.0, (ta, tb, alpha)


def test_metrics_hand_computed():
    hi = np.array([1.0, 2.0, 0.0, -4.0])
    lo = np.array([0.0, 2.0 ** -60, 0.0, 0.0])
    C = np.array([1.0, 2.0, 0.0, -4.0 - 2 * 8.881784197001252e-16])
    r = O.metrics(4, 1, C, 4, hi, lo)
    # entry 1: err 2^-60 / ulp(2)=2^-51 -> 2^-9; entry 3: err 2 ulps of 4 -> 2
    assert r["ulp_max"] == 2.0
    assert abs(r["ulp_mean"] - (2.0 ** -9 + 2.0) / 4) < 1e-15
    assert r["cr_frac"] == 0.75
