"""CPU checks of the kernels' per-element arithmetic (os2-xe_amd/csrc/math.hpp),
compiled for the host as lib/liboz2_emu.so, against the oracle and the golden
vectors: the exact split (both scaling widths, complex groups), the GEMM
residue epilogue, the CRT reconstruction and the exponent rule.  The GPU tests
then only have to establish that the kernels apply this arithmetic to the
right elements."""
import ctypes
import json
import math
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
_emu = ctypes.CDLL(os.path.join(ROOT, "os2-xe_amd", "lib", "liboz2_emu.so"))
vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
_emu.emu_split.argtypes = [vp, vp, i64, i32, i32, vp]
_emu.emu_split_cplx.argtypes = [vp, vp, vp, i64, i32, i32, i32, vp]
_emu.emu_acc_residue.argtypes = [vp, i64, i32, vp]
_emu.emu_crt.argtypes = [vp, vp, i64, i32, vp]
_emu.emu_exponent_from_maxbits.argtypes = [ctypes.c_uint64, i32]
_emu.emu_plan.argtypes = [i32, i64, vp, vp, vp, vp, vp]


def P(a):
    return a.ctypes.data_as(vp)


def load(name):
    with open(os.path.join(GOLD, name)) as f:
        return json.load(f)


def h2f(h):
    return struct.unpack("<d", struct.pack("<Q", h))[0]


def emu_split(v, e, s, tk):
    v = np.ascontiguousarray(v, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.int32)
    out = np.zeros((s, v.size), dtype=np.int8)
    _emu.emu_split(P(v), P(e), v.size, s, tk, P(out))
    return out


def test_split_golden():
    for case in load("split.json"):
        for col in case["cols"]:
            v = np.array([h2f(b) for b in col["bits"]])
            e = np.full(v.size, col["e"], dtype=np.int32)
            got = emu_split(v, e, case["s"], case["tk"])
            assert got.tolist() == col["res"], (case["s"], case["tk"])


@pytest.mark.parametrize("s,tk", [(2, 6), (8, 24), (13, 44), (16, 55), (18, 62), (19, 63), (24, 86), (24, 91),
                                  (24, 94), (24, 95), (24, 100)])
def test_split_random_vs_oracle(s, tk, rng):
    n = 20000
    v = rng.uniform(-1, 1, n) * np.exp2(rng.integers(-70, 1, n).astype(float))
    v[:8] = [0.0, -0.0, 5e-324, -5e-324, 1.0 - 2 ** -53, -(1.0 - 2 ** -53), 0.5, -0.75]
    mx = np.abs(v).max()
    e = np.full(n, tk - math.frexp(mx)[1], dtype=np.int32)
    e[:50] = rng.integers(tk - 60, tk + 1, 50)  # other exponents, |v 2^e| < 2^tk still
    e[50] = -2 ** 31  # non-finite row: all residues zero
    got = emu_split(v, e, s, tk)
    want = O.split_values(v, np.where(e == -2 ** 31, 0, e), s)
    want[:, 50] = 0
    assert np.array_equal(got, want)


@pytest.mark.parametrize("conj", [0, 1])
@pytest.mark.parametrize("s,tk", [(14, 49), (16, 55), (24, 86), (24, 94), (24, 100)])
def test_split_complex_vs_oracle(s, tk, conj, rng):
    n = 5000
    re = rng.uniform(-1, 1, n)
    im = rng.uniform(-1, 1, n)
    e = np.full(n, tk, dtype=np.int32)
    out = np.zeros((3, s, n), dtype=np.int8)
    _emu.emu_split_cplx(P(re), P(im), P(e), n, s, tk, conj, P(out))
    r0 = O.split_values(re, e, s).astype(np.int64)
    r1 = O.split_values(im, e, s).astype(np.int64)
    for t, m in enumerate(O.moduli(s)):
        h = (m - 1) // 2
        a = np.where(r0[t] < -h, r0[t] + m, r0[t])
        b = np.where(r1[t] < -h, r1[t] + m, r1[t])
        if conj:
            b = -b
        ss = (((a + b) % m) + h) % m - h
        for g, want in enumerate((a, b, ss)):
            assert np.array_equal(out[g, t].astype(np.int64) & 0xFF, want & 0xFF), (t, g)


def test_exact_quotient_every_value():
    # mod_m23 (split and GEMM epilogue): q = mulhi24(2v, ceil(2^31 / m)) is floor(v / m)
    # for every v < 2^23 and every modulus of every table (m = 256 never takes this path)
    _emu.emu_mod_m23_mismatches.restype = ctypes.c_int64
    _emu.emu_mod_m23_mismatches.argtypes = [i32, i32]
    seen = set()
    for s in range(2, 25):
        for t, m in enumerate(O.moduli(s)):
            if m == 256 or m in seen:
                continue
            seen.add(m)
            assert _emu.emu_mod_m23_mismatches(s, t) == 0, m
    assert len(seen) == 25


def test_symmetric_reduction_every_value():
    # sym_m23 (real and complex split): from w = v - h (mod 2^32), the chain's start folded
    # with -h, it gives (v mod m) - h for every dot-chain value v < 2^20 + 2^8 of every
    # modulus, including v < h where w wraps below zero
    _emu.emu_sym_m23_mismatches.restype = ctypes.c_int64
    _emu.emu_sym_m23_mismatches.argtypes = [i32, i32]
    seen = set()
    for s in range(2, 25):
        for t, m in enumerate(O.moduli(s)):
            if m == 256 or m in seen:
                continue
            seen.add(m)
            assert _emu.emu_sym_m23_mismatches(s, t) == 0, m
    assert len(seen) == 25


@pytest.mark.parametrize("s", [2, 7, 14, 16, 24])
def test_acc_residue(s, rng):
    # documented domain |x| < 2^30 - 255 (GEMM accumulators of one k-chunk: |x| <= 65408 * 2^14),
    # its edges, and the largest fold value y = (2^15 - 1) c16 + 65535 (x + c30 = 2^31 - 1)
    lim = 2 ** 30 - 256
    x = rng.integers(-lim, lim + 1, 100000, dtype=np.int64)
    x[:10] = [0, -1, 1, -lim, lim, 65408 * 2 ** 14, -65408 * 2 ** 14, 65408 * 127 * 127, -65408 * 128 * 127, 2 ** 14]
    for j, m in enumerate(O.moduli(s)[1:]):
        x[10 + j] = 2 ** 31 - 1 - m * ((2 ** 30 + m - 1) // m)  # u = 2^31 - 1 for modulus m
    out = np.zeros((s, x.size), dtype=np.uint8)
    xi = x.astype(np.int32)
    _emu.emu_acc_residue(P(xi), x.size, s, P(out))
    for t, m in enumerate(O.moduli(s)):
        assert np.array_equal(out[t].astype(np.int64), x % m), m


def test_crt_golden():
    for case in load("crt.json"):
        s = case["s"]
        res = np.array(case["res"], dtype=np.uint8)
        sh = np.array(case["shift"], dtype=np.int32)
        out = np.zeros(len(sh))
        _emu.emu_crt(P(res), P(sh), len(sh), s, P(out))
        assert out.view(np.uint64).tolist() == [int(x) for x in case["expect"]], s


@pytest.mark.parametrize("s", list(range(2, 25)))
def test_crt_random_vs_oracle(s, rng):
    plan = O.make_plan(s, 16)
    ms = O.moduli(s)
    M4 = plan.M_int // 4
    n = 2000
    xs = [int(rng.integers(0, 2 ** 62)) * int(rng.integers(0, 2 ** 62)) * int(rng.integers(0, 2 ** 62)) for _ in range(n)]
    xs = [(x % (M4 + 1)) * (1 if i % 3 else -1) for i, x in enumerate(xs)]
    xs[0], xs[1], xs[2] = M4, -M4, 0
    res = np.array([[x % m for m in ms] for x in xs], dtype=np.uint8)
    sh = rng.integers(-40, 40, n).astype(np.int32)
    out = np.zeros(n)
    _emu.emu_crt(P(res), P(sh), n, s, P(out))
    want = np.array([O.crt_value(plan, res[i], int(sh[i])) for i in range(n)])
    assert np.array_equal(out.view(np.uint64), want.view(np.uint64))


def test_fixed_point_quotient_margin():
    # The CRT takes q = round(sum_t r_t F_t / 2^16), F_t = round(2^16 y_t / m_t).  Its error
    # against the exact sum_t r_t y_t / m_t, at most sum_t (m_t - 1) |F_t / 2^16 - y_t / m_t|,
    # must leave the 1/4 margin of |C'| <= M/4 short of 1/2 for every moduli set; and the
    # residues of the extreme C' = +-M/4 (and of -1, every residue m_t - 1) give the exact q.
    from fractions import Fraction
    for s in range(2, 25):
        plan = O.make_plan(s, 16)
        ms, ys = O.moduli(s), [int(plan.y[t]) for t in range(s)]
        F = [(y * 2 ** 17 + m) // (2 * m) for y, m in zip(ys, ms)]
        err = sum(Fraction(m - 1) * abs(Fraction(f, 2 ** 16) - Fraction(y, m)) for f, y, m in zip(F, ys, ms))
        assert err < Fraction(1, 4) - Fraction(1, 100), (s, float(err))
        M = plan.M_int
        for x in (M // 4, -(M // 4), -1, 0, 1):
            r = [x % m for m in ms]
            exact = sum(Fraction(ri * y, m) for ri, y, m in zip(r, ys, ms))
            q = (sum(ri * f for ri, f in zip(r, F)) + 2 ** 15) >> 16
            assert q == round(exact) and abs(exact - q) <= Fraction(1, 4), (s, x)


def test_exponent_rule():
    for b, tk in ((0, 30), (0x7FF0000000000000, 30), (0x7FF8000000000000, 5), (1, 55), (0x3FF0000000000000, 55),
                  (0x000FFFFFFFFFFFFF, 55), (0x7FEFFFFFFFFFFFFF, 100)):
        assert _emu.emu_exponent_from_maxbits(b, tk) == O.exponent_from_maxbits(b, tk)


def emu_plan(s, k):
    tk, nl = ctypes.c_int(0), ctypes.c_int(0)
    y = np.zeros(24, dtype=np.uint32)
    W = np.zeros(24 * 5, dtype=np.uint64)
    M = np.zeros(5, dtype=np.uint64)
    rc = _emu.emu_plan(s, k, ctypes.byref(tk), ctypes.byref(nl), P(y), P(W), P(M))
    if rc:
        return rc, None

    def big(words):
        return sum(int(w) << (64 * i) for i, w in enumerate(words))
    return 0, {"tk": tk.value, "NL": nl.value, "y": [int(v) for v in y[:s]],
               "W": [big(W[5 * t:5 * t + 5]) for t in range(s)], "M": big(M)}


def test_plan_matches_golden():
    # the library's make_plan (exact integer form) against the fixture of exact Python integers
    for g in load("plan.json")["plans"]:
        rc, p = emu_plan(g["s"], g["k"])
        if g["tk"] is None:
            assert rc == 1, (g["s"], g["k"])
            continue
        assert rc == 0
        assert p["tk"] == g["tk"] and p["NL"] == g["NL"] and p["y"] == g["y"], (g["s"], g["k"])
        assert p["M"] == int(g["M"], 16) and p["W"] == [int(w, 16) for w in g["W"]]


def test_plan_tk_at_every_boundary():
    # tk is the largest t <= 100 with k 2^(2t+2) <= M: at k_t = floor(M / 2^(2t+2)) it is t (or the cap),
    # one past it t - 1.  The oracle restates the reference's rule (a double-precision estimate of
    # (log2 M - 2 - log2 k) / 2, then decrements): it agrees for every k below 2^50; from about 2^50.4 on
    # (k that no GEMM in device memory can have) its estimate can start one below the largest t
    for s in range(2, 25):
        M = math.prod(O.moduli(s))
        for t in range(1, 101):
            kt = M >> (2 * t + 2)
            if kt < 1:
                break
            for k in (kt - 1, kt, kt + 1):
                if not 1 <= k < 2 ** 62:
                    continue
                want = max((u for u in range(1, 101) if (k << (2 * u + 2)) <= M), default=None)
                rc, p = emu_plan(s, k)
                if want is None:
                    assert rc == 1 and O.precision_bits(k, s) < 0
                else:
                    assert rc == 0 and p["tk"] == want, (s, k)
                    ref = O.precision_bits(k, s)
                    if k < 2 ** 50:
                        assert ref == want, (s, k)
                    else:
                        assert ref in (want, want - 1), (s, k)
