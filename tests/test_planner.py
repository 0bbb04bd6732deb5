"""CPU checks of the panel planner (os2-xe_amd/csrc/planner.cpp, host build in
lib/liboz2_emu.so), the counterpart of the reference's tiling planner
(src/ozaki2_gpu.hpp:1072-1166).

What pins it: every plan fits its limit and its panels cover C; it is the
cheapest of all balanced panel pairs under its own cost model (exhaustive
search); more memory never gives a costlier plan; and the reference's
single-tile cases (whole problem resident when memory allows) come out as one
panel pair.  Results do not depend on the panels (GPU tiling tests)."""
import ctypes
import os
import random

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_emu = ctypes.CDLL(os.path.join(ROOT, "os2-xe_amd", "lib", "liboz2_emu.so"))
vp, i64 = ctypes.c_void_p, ctypes.c_int64
_emu.emu_plan_tiling.argtypes = [vp, ctypes.c_uint64, vp]
_emu.emu_layout_total.argtypes = [vp, i64, i64]
_emu.emu_layout_total.restype = ctypes.c_uint64
_emu.emu_panel_cost.argtypes = [vp, i64, i64]
_emu.emu_panel_cost.restype = ctypes.c_double

GiB = 1 << 30


def shape(m, n, k, s, cplx=False, pa=0, pb=0, cus=256, pad=1, pow2=1, per_plane=0, tile=-1):
    return np.array([m, n, k, s, int(cplx), pa, pb, cus, pad, pow2, per_plane, tile], dtype=np.int64)


def plan(sh, limit):
    out = np.zeros(6, dtype=np.int64)
    rc = _emu.emu_plan_tiling(sh.ctypes.data_as(vp), limit, out.ctypes.data_as(vp))
    return None if rc else dict(zip(("mb", "nb", "pm", "pn", "kch", "ws"), (int(v) for v in out)))


def total(sh, mb, nb):
    return int(_emu.emu_layout_total(sh.ctypes.data_as(vp), mb, nb))


def cost(sh, mb, nb):
    return float(_emu.emu_panel_cost(sh.ctypes.data_as(vp), mb, nb))


def ceil_div(a, b):
    return -(-a // b)


def balanced(x):
    """x itself and round_up(ceil(x / p), 64) for every panel count p"""
    out = {x}
    for p in range(2, ceil_div(x, 64) + 2):
        out.add(ceil_div(ceil_div(x, p), 64) * 64)
    return sorted(out, reverse=True)


def check_plan(sh, limit, p):
    m, n = int(sh[0]), int(sh[1])
    assert p["ws"] == total(sh, p["mb"], p["nb"]) <= limit
    assert p["pm"] == ceil_div(m, p["mb"]) and p["pn"] == ceil_div(n, p["nb"])
    assert (p["pm"] - 1) * p["mb"] < m <= p["pm"] * p["mb"]
    assert (p["pn"] - 1) * p["nb"] < n <= p["pn"] * p["nb"]
    for b, x in ((p["mb"], m), (p["nb"], n)):
        assert b == x or b % 64 == 0


def test_single_panel_when_memory_allows():
    # the configurations of the benchmarks: everything in one panel pair on a 288 GB GPU
    budget = 250 * GiB
    for m, n, k, s, cplx in ((8192, 8192, 8192, 16, False), (4096, 4096, 4096, 16, True),
                             (4000, 8000, 4000, 24, False), (1024, 1024, 1024, 14, False)):
        p = plan(shape(m, n, k, s, cplx), budget)
        assert (p["mb"], p["nb"], p["pm"], p["pn"]) == (m, n, 1, 1), p
    p = plan(shape(2048, 2048, 200000, 16), budget)
    assert (p["pm"], p["pn"], p["kch"]) == (1, 1, 4)


def test_nothing_fits():
    assert plan(shape(4096, 4096, 8192, 16), 1 << 20) is None
    # the smallest panel pair of a long-k product needs its 64 rows of planes
    sh = shape(64, 64, 200000, 24)
    need = total(sh, 64, 64)
    assert plan(sh, need - 1) is None
    assert plan(sh, need)["ws"] == need


@pytest.mark.parametrize("seed", range(6))
def test_cheapest_balanced_panels(seed):
    # the planner's choice against exhaustive search over every balanced (mb, nb)
    rnd = random.Random(seed)
    for _ in range(12):
        m, n = rnd.randint(1, 3000), rnd.randint(1, 3000)
        k = rnd.choice((1, 100, 1000, 5000, 70000))
        s = rnd.choice((2, 8, 16, 24))
        cplx = rnd.random() < 0.3
        sh = shape(m, n, k, s, cplx, pa=rnd.choice((0, 8 * m)), pb=rnd.choice((0, 8 * n)),
                   pad=rnd.randint(0, 1), pow2=0)
        full = total(sh, m, n)
        limit = int(full * rnd.choice((0.05, 0.2, 0.5, 0.9, 1.0, 2.0)))
        p = plan(sh, limit)
        cands = [(cost(sh, mb, nb), mb, nb) for mb in balanced(m) for nb in balanced(n)
                 if total(sh, mb, nb) <= limit]
        if not cands:
            assert p is None, (m, n, k, s, limit)
            continue
        assert p is not None
        check_plan(sh, limit, p)
        best = min(c[0] for c in cands)
        assert cost(sh, p["mb"], p["nb"]) <= best * (1 + 1e-9), (m, n, k, s, cplx, limit, p, min(cands))


def test_more_memory_is_never_costlier():
    rnd = random.Random(99)
    for _ in range(20):
        m, n, k, s = rnd.randint(100, 6000), rnd.randint(100, 6000), rnd.randint(64, 20000), rnd.randint(2, 24)
        sh = shape(m, n, k, s, pow2=0)
        full = total(sh, m, n)
        prev = None
        for f in (0.03, 0.1, 0.3, 1.0, 3.0):
            p = plan(sh, int(full * f))
            if p is None:
                continue
            c = cost(sh, p["mb"], p["nb"])
            if prev is not None:
                assert c <= prev * (1 + 1e-9)
            prev = c
        assert plan(sh, full)["pm"] * plan(sh, full)["pn"] == 1


def test_pow2_candidates_and_tuning_fields():
    sh = shape(5000, 7000, 3000, 16)
    limit = total(sh, 5000, 7000) // 6
    a = plan(sh, limit)
    check_plan(sh, limit, a)
    b = plan(shape(5000, 7000, 3000, 16, pow2=0), limit)
    check_plan(sh, limit, b)
    # power-of-two panels are extra candidates: never costlier than without them
    assert cost(sh, a["mb"], a["nb"]) <= cost(sh, b["mb"], b["nb"]) * (1 + 1e-9)
    # the plane stagger only adds bytes
    assert total(shape(4096, 4096, 4096, 16, pad=1), 4096, 4096) > total(shape(4096, 4096, 4096, 16, pad=0), 4096, 4096)
    # a forced tile sets the padding: 64 x 64 tiles pad 100 rows to 128, 256 x 256 tiles to 256
    assert total(shape(100, 100, 128, 2, tile=64), 100, 100) < total(shape(100, 100, 128, 2, tile=256), 100, 100)


def test_large_shapes_plan_quickly():
    # many distinct panel counts (m = 10^6) must not make planning slow
    import time
    sh = shape(1_000_000, 300_000, 4096, 16)
    t0 = time.perf_counter()
    p = plan(sh, 20 * GiB)
    assert time.perf_counter() - t0 < 0.5
    check_plan(sh, 20 * GiB, p)
