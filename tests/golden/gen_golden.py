"""Generate the golden vectors under tests/golden/ with exact Python integers.

Nothing here imports or runs the reference or the oracle: every value follows
from the mathematics of Ozaki Scheme II as the reference defines it
(src/ozaki2_gpu.hpp) --

* plan.json   moduli per s, M, tk, NL, y_t = (M/m_t)^-1 mod m_t, W_t = (M/m_t) y_t
              (make_plan, hpp:159-200; tk = largest t <= 100 with k*2^(2t+2) <= M)
* split.json  symmetric int8 residues of RNE(v*2^e) (unit_split, tests/test_ozaki2.cpp:195-254)
* crt.json    RNE(x)*2^-shift from the residues of integers |x| <= M/4 (unit_crt, test:256-324)
* gemm.json   small DGEMM/ZGEMM cases whose C follows from the closed form
              C_ij = RNE(sum_p RNE(a_ip 2^D_i) RNE(b_pj 2^E_j)) 2^-(D_i+E_j)

Run:  python tests/golden/gen_golden.py   (rewrites the JSON files in place)
"""
import json
import math
import os
import random
import struct
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))

# the pairwise-coprime moduli sets (reference src/coprime_tables.hpp:10-111, s = 2..24)
_BASE16 = [256, 253, 251, 249, 247, 241, 239, 235, 233, 229, 227, 223, 217, 211, 199, 197]
SETS = {
    2: [256, 255], 3: [256, 255, 253], 4: [256, 255, 253, 251], 5: [256, 255, 253, 251, 247],
    6: [256, 255, 253, 251, 247, 241],
    7: [256, 253, 251, 249, 247, 245, 241],
    8: [256, 253, 251, 249, 247, 245, 241, 239],
    9: [256, 253, 251, 249, 247, 245, 241, 239, 233],
    10: [256, 253, 251, 249, 247, 245, 241, 239, 233, 229],
    11: [256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227],
    12: [256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227, 223],
    13: [256, 253, 251, 249, 247, 245, 241, 239, 233, 229, 227, 223, 211],
}
_tail = [193, 191, 181, 179, 173, 167, 163, 157]
for _s in range(14, 25):
    SETS[_s] = (_BASE16 + _tail)[:_s] if _s >= 16 else _BASE16[:_s]


def prod(xs):
    r = 1
    for x in xs:
        r *= x
    return r


def tk_for(M, k):
    """Largest t <= 100 with k * 2^(2t+2) <= M (hpp:177-186), or None."""
    t = 100
    while t >= 1:
        if k << (2 * t + 2) <= M:
            return t
        t -= 1
    return None


def plan(s, k):
    mods = SETS[s]
    M = prod(mods)
    tk = tk_for(M, k)
    nl = max(2, (M.bit_length() - 1) // 32 + 1)
    ys, Ws = [], []
    for m in mods:
        Mt = M // m
        y = pow(Mt % m, -1, m)
        ys.append(y)
        Ws.append(Mt * y)
    return {"s": s, "k": k, "moduli": mods, "M": hex(M), "tk": tk, "NL": nl, "y": ys, "W": [hex(w) for w in Ws]}


def f2hex(v):
    return struct.unpack("<Q", struct.pack("<d", v))[0]


def frexp_exp(v):
    return math.frexp(v)[1]


def rne_int(fr):
    """Round a Fraction to the nearest integer, ties to even."""
    return round(fr)


def scaled(v, e):
    return rne_int(Fraction(v) * (Fraction(2) ** e))


def sym(x, m):
    h = (m - 1) // 2
    r = ((x % m) + h) % m - h
    return ((r + 128) % 256) - 128  # stored as int8 (m = 256 keeps +128 as -128)


def rne_ldexp(x, shift):
    """float(x) is RNE to 53 bits; ldexp then rounds once more only when subnormal."""
    if x == 0:
        return 0.0
    try:
        return math.ldexp(float(x), -shift)
    except OverflowError:
        return math.inf if x > 0 else -math.inf


def gen_plans():
    ks = [1, 16, 1024, 2048, 4096, 8000, 8192, 131071, 200000, 400000]
    out = []
    for s in range(2, 25):
        for k in ks:
            p = plan(s, k)
            if p["tk"] is not None:
                out.append(p)
            else:
                out.append({"s": s, "k": k, "tk": None})
    mfb = []
    for k in [1, 64, 512, 1024, 1025, 2048, 4096, 8000, 8192, 65537, 131071, 200000]:
        for bits in [24, 40, 53, 55, 62, 80, 86, 90]:
            best = 24
            for s in range(2, 25):
                t = tk_for(prod(SETS[s]), k)
                if t is not None and t >= bits:
                    best = s
                    break
            mfb.append({"k": k, "bits": bits, "s": best})
    return {"plans": out, "moduli_for_bits": mfb}


def gen_split():
    rng = random.Random(20260601)
    cases = []
    for S, tk in [(8, 24), (12, 40), (16, 55), (18, 62), (20, 70), (24, 85), (24, 100), (3, 5)]:
        rows = 96
        cols = []
        c0 = [rng.uniform(-1.0, 1.0) for _ in range(rows)]
        c1 = [rng.uniform(-1.0, 1.0) * math.ldexp(1.0, -rng.randrange(120)) for _ in range(rows)]
        c2 = []
        for r in range(rows):
            if r % 7 == 0:
                c2.append(0.0)
            elif r % 7 == 1:
                c2.append(-0.0)
            else:
                c2.append(rng.uniform(-1.0, 1.0) * math.ldexp(1.0, -1030 - rng.randrange(40)))
        # ties: values whose scaled form lands exactly on .5 (RNE to even)
        # column max 0.75 fixes e = tk, so (2j+1) * 2^(-tk-1) scales to j + 1/2 exactly
        c3 = [0.75]
        for r in range(1, rows):
            j = rng.randrange(1 << min(tk - 2, 50)) if tk > 2 else 0
            c3.append(math.ldexp(2 * j + 1, -tk - 1) * (1 if r % 2 else -1))
        for col in (c0, c1, c2, c3):
            mx = max(abs(v) for v in col)
            e = tk - frexp_exp(mx) if mx != 0 else 0
            xs = [scaled(v, e) for v in col]
            res = [[sym(x, m) for x in xs] for m in SETS[S]]
            cols.append({"bits": [f2hex(v) for v in col], "e": e, "x": [str(x) for x in xs], "res": res})
        cases.append({"s": S, "tk": tk, "cols": cols})
    return cases


def gen_crt():
    rng = random.Random(777)
    cases = []
    for S in range(2, 25):
        M = prod(SETS[S])
        Mq = M // 4
        n = 48
        xs, shifts = [], []
        for i in range(n):
            if i < 4:
                x = Mq
            elif i == 4:
                x = 0
            else:
                bits = rng.randrange(1, Mq.bit_length() + 1)
                x = min(rng.getrandbits(bits), Mq)
            if rng.random() < 0.5 and x != 0:
                x = -x
            if i == 3:
                x = -Mq
            xs.append(x)
            shifts.append(rng.randrange(-20, 21) + 3)
        # subnormal results (the reference's ldexp rounds a second time there)
        for j in range(4):
            x = rng.getrandbits(min(100, Mq.bit_length() - 1)) | 1
            xs.append(x)
            shifts.append(x.bit_length() + 1050 + j * 5)
        # an exact tie at 53 bits when the value has more than 53 significant bits
        if Mq.bit_length() > 60:
            x = (((1 << 52) | 5) << 2) | 2  # 55 bits, round bit set, sticky clear, m53 odd -> rounds up
            xs.append(x)
            shifts.append(0)
            x = (((1 << 52) | 4) << 2) | 2  # tie, m53 even -> stays
            xs.append(x)
            shifts.append(0)
        res = [[x % m for m in SETS[S]] for x in xs]
        exp = [f2hex(rne_ldexp(x, sh)) for x, sh in zip(xs, shifts)]
        cases.append({"s": S, "x": [str(x) for x in xs], "shift": shifts, "res": res, "expect": exp})
    return cases


def closed_real(ta, tb, m, n, k, A, lda, B, ldb, s):
    p = plan(s, k)
    tk = p["tk"]

    def opA(i, q):
        return A[i + q * lda] if ta == "N" else A[q + i * lda]

    def opB(q, j):
        return B[q + j * ldb] if tb == "N" else B[j + q * ldb]

    D = []
    for i in range(m):
        mx = max(abs(opA(i, q)) for q in range(k))
        D.append(tk - frexp_exp(mx) if mx != 0 else 0)
    E = []
    for j in range(n):
        mx = max(abs(opB(q, j)) for q in range(k))
        E.append(tk - frexp_exp(mx) if mx != 0 else 0)
    X = [[scaled(opA(i, q), D[i]) for q in range(k)] for i in range(m)]
    Y = [[scaled(opB(q, j), E[j]) for q in range(k)] for j in range(n)]
    C = []
    for j in range(n):
        for i in range(m):
            acc = sum(X[i][q] * Y[j][q] for q in range(k))
            C.append(rne_ldexp(acc, D[i] + E[j]))
    return C


def closed_cplx(ta, tb, m, n, k, A, lda, B, ldb, s):
    """A, B interleaved (re, im) lists."""
    p = plan(s, 2 * k)
    tk = p["tk"]

    def opA(i, q):
        idx = i + q * lda if ta == "N" else q + i * lda
        re, im = A[2 * idx], A[2 * idx + 1]
        return re, (-im if ta == "C" else im)

    def opB(q, j):
        idx = q + j * ldb if tb == "N" else j + q * ldb
        re, im = B[2 * idx], B[2 * idx + 1]
        return re, (-im if tb == "C" else im)

    D, E = [], []
    for i in range(m):
        mx = max(max(abs(opA(i, q)[0]), abs(opA(i, q)[1])) for q in range(k))
        D.append(tk - frexp_exp(mx) if mx != 0 else 0)
    for j in range(n):
        mx = max(max(abs(opB(q, j)[0]), abs(opB(q, j)[1])) for q in range(k))
        E.append(tk - frexp_exp(mx) if mx != 0 else 0)
    C = []
    for j in range(n):
        for i in range(m):
            re = im = 0
            for q in range(k):
                ar, ai = opA(i, q)
                br, bi = opB(q, j)
                xr, xi = scaled(ar, D[i]), scaled(ai, D[i])
                yr, yi = scaled(br, E[j]), scaled(bi, E[j])
                re += xr * yr - xi * yi
                im += xr * yi + xi * yr
            C.extend([rne_ldexp(re, D[i] + E[j]), rne_ldexp(im, D[i] + E[j])])
    return C


def gen_gemm():
    rng = random.Random(4242)
    cases = []

    def mat(rows, cols, ld, kind):
        out = []
        for c in range(cols):
            for r in range(ld):
                if r >= rows:
                    out.append(1e300)  # padding must never be read
                elif kind == "u":
                    out.append(rng.uniform(-1.0, 1.0))
                elif kind == "wide":
                    out.append((1 + rng.random()) * rng.choice([-1, 1]) * math.ldexp(1.0, rng.randrange(-20, 21)))
                elif kind == "int":
                    out.append(float(rng.randrange(-1000, 1001)))
                else:
                    raise ValueError(kind)
        return out

    real_specs = [
        ("N", "N", 5, 7, 9, 8, "u", 0), ("T", "N", 6, 4, 11, 12, "u", 2), ("N", "T", 3, 5, 17, 16, "wide", 1),
        ("T", "T", 4, 4, 33, 20, "wide", 3), ("N", "N", 7, 3, 64, 24, "u", 0), ("N", "N", 6, 6, 40, 6, "int", 0),
        ("T", "N", 2, 9, 129, 14, "u", 0), ("N", "N", 1, 1, 1, 2, "u", 0),
    ]
    for ta, tb, m, n, k, s, kind, pad in real_specs:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        A = mat(ar, ac, ar + pad, kind)
        B = mat(br, bc, br + pad, kind)
        C = closed_real(ta, tb, m, n, k, A, ar + pad, B, br + pad, s)
        cases.append({"kind": "d", "ta": ta, "tb": tb, "m": m, "n": n, "k": k, "s": s, "lda": ar + pad,
                      "ldb": br + pad, "A": [f2hex(v) for v in A], "B": [f2hex(v) for v in B],
                      "C": [f2hex(v) for v in C]})
    cplx_specs = [("N", "N", 3, 4, 7, 10, 0), ("C", "N", 4, 3, 9, 16, 1), ("T", "C", 2, 5, 12, 14, 2),
                  ("N", "T", 3, 3, 20, 22, 0)]
    for ta, tb, m, n, k, s, pad in cplx_specs:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        A = [v for pair in zip(mat(ar, ac, ar + pad, "u"), mat(ar, ac, ar + pad, "u")) for v in pair]
        B = [v for pair in zip(mat(br, bc, br + pad, "u"), mat(br, bc, br + pad, "u")) for v in pair]
        C = closed_cplx(ta, tb, m, n, k, A, ar + pad, B, br + pad, s)
        cases.append({"kind": "z", "ta": ta, "tb": tb, "m": m, "n": n, "k": k, "s": s, "lda": ar + pad,
                      "ldb": br + pad, "A": [f2hex(v) for v in A], "B": [f2hex(v) for v in B],
                      "C": [f2hex(v) for v in C]})
    return cases


def main():
    with open(os.path.join(HERE, "plan.json"), "w") as f:
        json.dump(gen_plans(), f)
    with open(os.path.join(HERE, "split.json"), "w") as f:
        json.dump(gen_split(), f)
    with open(os.path.join(HERE, "crt.json"), "w") as f:
        json.dump(gen_crt(), f)
    with open(os.path.join(HERE, "gemm.json"), "w") as f:
        json.dump(gen_gemm(), f)


if __name__ == "__main__":
    main()
