"""Inputs of the accuracy harness (liboz2_check.so oz2chk_gen, oz2/bench_accuracy.py): the
reference's generator (benchmarks/bench_accuracy.cpp:41-69) must give the same matrices.
The uniform draws are pinned against an independent implementation of the published
MT19937-64 algorithm (itself checked against the 10000th output the C++ standard mandates
for the default seed) and the uniform-distribution rule U = x / 2^64; the structure of the
spike and cancel inputs against their definitions.  The normal draws (phiX) follow the
C++ library's normal_distribution, which no fixture here pins."""
import math

import numpy as np
import pytest

from oz2 import check

M64 = (1 << 64) - 1


def mt19937_64(seed):
    n, m, a = 312, 156, 0xB5026F5AA96619E9
    mt = [seed & M64]
    for i in range(1, n):
        mt.append((6364136223846793005 * (mt[-1] ^ (mt[-1] >> 62)) + i) & M64)
    idx = n
    upper, lower = 0xFFFFFFFF80000000, 0x7FFFFFFF
    while True:
        if idx >= n:
            for i in range(n):
                x = (mt[i] & upper) | (mt[(i + 1) % n] & lower)
                xa = x >> 1
                if x & 1:
                    xa ^= a
                mt[i] = mt[(i + m) % n] ^ xa
            idx = 0
        y = mt[idx]
        idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000
        y ^= (y << 37) & 0xFFF7EEE000000000
        y ^= y >> 43
        yield y & M64


def test_reference_generator_itself():
    g = mt19937_64(5489)
    for _ in range(9999):
        next(g)
    assert next(g) == 9981545732273789042  # [rand.predef]: mt19937_64, default seed


@pytest.mark.parametrize("seed", [11 + 256, 77 + 4096, 5489])
def test_urand_matches_mt19937_64(seed):
    rows, cols = 7, 9
    X = check.gen("urand", rows, cols, seed, True)
    g = mt19937_64(seed)
    want = np.array([2.0 * (float(next(g)) / 2.0 ** 64) - 1.0 for _ in range(rows * cols)])
    assert np.array_equal(X, want)  # column-major: entry (r, c) is draw r + c rows


def test_spike_and_cancel_structure():
    rows, cols = 12, 10
    base = check.gen("urand", rows, cols, 5, True).reshape(cols, rows).T
    spike = check.gen("spike", rows, cols, 5, True).reshape(cols, rows).T
    for r in range(rows):
        c = (r * 7919) % cols
        assert spike[r, c] == base[r, c] * 2.0 ** 40
        others = np.delete(np.arange(cols), c)
        assert np.array_equal(spike[r, others], base[r, others])
    assert np.array_equal(check.gen("spike", rows, cols, 5, False), check.gen("urand", rows, cols, 5, False))
    ca = check.gen("cancel", rows, cols, 5, True).reshape(cols, rows).T
    h = cols // 2
    assert np.array_equal(ca[:, h:2 * h], base[:, :h]) and np.array_equal(ca[:, :h], base[:, :h])
    bb = check.gen("urand", cols, rows, 9, False).reshape(rows, cols).T  # B: cols x rows here
    cb = check.gen("cancel", cols, rows, 9, False).reshape(rows, cols).T
    hb = cols // 2
    assert np.array_equal(cb[hb:2 * hb], -bb[:hb] + np.ldexp(bb[hb:2 * hb], -30))


def test_sincos_and_phi():
    A = check.gen("sincos", 4, 5, 1, True).reshape(5, 4).T
    B = check.gen("sincos", 4, 5, 1, False).reshape(5, 4).T
    for r in range(4):
        for c in range(5):
            assert A[r, c] == pytest.approx(math.sin(r + c), abs=1e-15)
            assert B[r, c] == pytest.approx(math.cos(r * c), abs=1e-15)
    P = check.gen("phi2", 50, 40, 3, True)
    assert np.array_equal(P, check.gen("phi2", 50, 40, 3, True))  # deterministic
    assert np.abs(P).max() < 0.5 * math.exp(2 * 10)  # |U - 0.5| < 1/2, |N| far below 10
    with pytest.raises(ValueError):
        check.gen("phi", 2, 2, 1, True)
