"""Batched DGEMMs (oz2_batch_*): one grouped launch sequence for many problems of
different shapes, optionally replayed from a hipGraph.  Every problem's C must be
bit-identical to the oracle (the reference's algorithm restated, run on the same
inputs with the same num_moduli) and to oz2_dgemm, whatever the macro tiles, graph
use or replay count; problems with k == 0 or alpha == 0 follow the C = beta*C rule
(hpp:750-751) and beta == 0 never reads C (hpp:623)."""
import numpy as np
import pytest
import torch

import oz2
from conftest import make_mat
from gpu_cases import bits, flat, log, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu

# (ta, tb, m, n, k, s, dist, alpha, beta)
PROBLEMS = [
    ("N", "N", 1, 1, 1, 8, "u11", 1.0, 0.0),
    ("N", "N", 37, 53, 129, 12, "uniform", 1.5, 0.5),
    ("T", "N", 300, 260, 200, 16, "phi", -2.0, 0.0),
    ("N", "T", 64, 64, 64, 14, "u11", 1.0, 1.0),
    ("T", "T", 129, 70, 1000, 16, "wide", 0.75, -0.25),
    ("N", "N", 513, 129, 777, 20, "u11", 1.0, 0.0),
    ("N", "T", 96, 200, 5000, 24, "wide", 1.0, 0.0),
    ("T", "N", 256, 256, 256, 10, "uniform", 2.0, 3.0),
    ("N", "N", 40, 30, 0, 16, "u11", 1.0, 0.5),      # k == 0: C = beta C
    ("N", "N", 33, 17, 50, 16, "u11", 0.0, -1.0),    # alpha == 0: C = beta C
    ("N", "N", 0, 17, 50, 16, "u11", 1.0, 0.0),      # empty: no-op
    ("T", "T", 200, 1, 300, 18, "phi", 1.0, 0.0),
]


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


def build(problems, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for (ta, tb, m, n, k, s, dist, alpha, beta) in problems:
        ar, ac = (m, k) if ta == "N" else (k, m)
        br, bc = (k, n) if tb == "N" else (n, k)
        lda, ldb, ldc = max(1, ar + 3), max(1, br + 1), max(1, m + 2)
        A = make_mat(ar, ac, lda, dist, rng)
        B = make_mat(br, bc, ldb, dist, rng)
        C0 = make_mat(m, n, ldc, "u11", rng)
        if beta == 0.0 and m and n:
            C0[:m, :] = np.nan  # must not be read
        out.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, alpha=alpha, beta=beta, s=s, lda=lda, ldb=ldb, ldc=ldc,
                        A=to_dev(flat(A)), B=to_dev(flat(B)), C0=flat(C0), Ah=flat(A), Bh=flat(B)))
    return out


def oracle_c(p):
    """the oracle's C for problem p (host copies of its inputs)"""
    C = p["C0"].copy()
    if p["m"] and p["n"]:
        O.dgemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["Ah"], p["lda"], p["Bh"], p["ldb"], p["beta"], C,
                p["ldc"], p["s"])
    return C


def reference(oz, p):
    C = to_dev(p["C0"].copy())
    oz.gemm(p["ta"], p["tb"], p["m"], p["n"], p["k"], p["alpha"], p["A"], p["lda"], p["B"], p["ldb"], p["beta"], C,
            p["ldc"], p["s"])
    p["want"] = oracle_c(p)
    return C


def desc(p, C, s=None):
    return dict(ta=p["ta"], tb=p["tb"], m=p["m"], n=p["n"], k=p["k"], alpha=p["alpha"], A=p["A"], lda=p["lda"],
                B=p["B"], ldb=p["ldb"], beta=p["beta"], C=C, ldc=p["ldc"], num_moduli=p["s"] if s is None else s)


def check_same(ref, got, what, ps=None):
    """every batch result against oz2_dgemm's (ref) and, given the problems, against the oracle"""
    for i, (r, g) in enumerate(zip(ref, got)):
        a, b = r.cpu().numpy(), g.cpu().numpy()
        nbad = int(np.count_nonzero(bits(a) != bits(b)))
        assert nbad == 0, f"{what}: problem {i} has {nbad} entries different from oz2_dgemm"
        if ps is not None:
            w = ps[i]["want"] if "want" in ps[i] else oracle_c(ps[i])
            nbad = int(np.count_nonzero(bits(b) != bits(w)))
            assert nbad == 0, f"{what}: problem {i} has {nbad} entries different from the oracle"


@pytest.mark.parametrize("graph", [False, True])
def test_batch_matches_single_calls(oz, graph):
    log(f"batch of {len(PROBLEMS)} problems, graph={graph}")
    ps = build(PROBLEMS)
    ref = [reference(oz, p) for p in ps]
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], graph=graph)
    assert b.moduli() == [p["s"] for p in ps]
    info = b.info()
    assert info["graph"] == graph and info["launches"] < 25, info
    b.run()
    torch.cuda.synchronize()
    check_same(ref, Cs, "first run", ps)
    b.close()


def test_forced_tiles_identical(oz):
    ps = build(PROBLEMS[:8], seed=11)
    ref = [reference(oz, p) for p in ps]
    for t in (256, 128, 64, 256128):
        log(f"batch with every macro tile {t}")
        Cs = [to_dev(p["C0"].copy()) for p in ps]
        b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], tiles=[t] * len(ps), graph=False)
        assert all(x in (0, t) for x in b.tiles())
        b.run()
        torch.cuda.synchronize()
        check_same(ref, Cs, f"tile {t}", ps)
        b.close()


def test_graph_replay_reads_current_inputs(oz):
    ps = build(PROBLEMS[1:6], seed=3)
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], graph=True)
    for rep in range(3):
        log(f"graph replay {rep}")
        for p, C in zip(ps, Cs):
            C.copy_(to_dev(p["C0"]))
        if rep == 2:  # new contents of the same buffers
            for p in ps:
                p["A"].mul_(-0.5)
                p["B"].add_(0.25)
                p["Ah"] = p["Ah"] * -0.5  # the same IEEE operations on the host copies
                p["Bh"] = p["Bh"] + 0.25
        b.run()
        torch.cuda.synchronize()
        ref = [reference(oz, p) for p in ps]
        check_same(ref, Cs, f"replay {rep}", ps)
    b.close()


def test_k_chunks_in_batch(oz):
    probs = [("N", "T", 40, 24, 140000, 16, "u11", 1.0, 0.0), ("N", "N", 70, 66, 300, 12, "u11", 1.0, 0.5)]
    ps = build(probs, seed=5)
    ref = [reference(oz, p) for p in ps]
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], graph=True)
    b.run()
    torch.cuda.synchronize()
    check_same(ref, Cs, "k-chunked", ps)
    b.close()


def test_plan_chooses_moduli_for_bits(oz):
    ps = build(PROBLEMS[1:4], seed=9)
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    b = oz2.Batch(oz, [desc(p, C, s=0) for p, C in zip(ps, Cs)], bits=53, graph=False)
    assert b.moduli() == [oz2.Ozaki2.moduli_for_bits(p["k"], 53) for p in ps]
    b.close()


def test_batch_argument_errors(oz):
    ps = build(PROBLEMS[1:3], seed=1)
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    d = [desc(p, C) for p, C in zip(ps, Cs)]
    d[1]["ldc"] = 1
    with pytest.raises(oz2.OZ2InvalidArgument, match="batch entry 1"):
        oz2.Batch(oz, d)
    d = [desc(p, C) for p, C in zip(ps, Cs)]
    d[0]["num_moduli"] = 30
    with pytest.raises(oz2.OZ2InvalidArgument, match="num_moduli"):
        oz2.Batch(oz, d)
    d = [desc(p, C) for p, C in zip(ps, Cs)]
    d[0].update(k=140000, ldb=140000, num_moduli=2)  # s too small for k: make_plan's error (hpp:186), before any launch
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz2.Batch(oz, d)
    d = [desc(p, C) for p, C in zip(ps, Cs)]
    with pytest.raises(oz2.OZ2InvalidArgument, match="tile"):
        oz2.Batch(oz, d, tiles=[256, 96])


@pytest.mark.parametrize("persistent,largest_first,tile", [(True, True, None), (False, False, None),
                                                           (True, False, None), (True, True, 256128),
                                                           (True, False, 128), (True, True, 64)])
def test_schedules_identical(oz, persistent, largest_first, tile):
    log(f"batch schedule persistent={persistent} largest_first={largest_first} tile={tile}")
    ps = build(PROBLEMS, seed=13)
    ref = [reference(oz, p) for p in ps]
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], graph=True, persistent=persistent,
                  largest_first=largest_first, tiles=None if tile is None else [tile] * len(ps))
    b.run()
    torch.cuda.synchronize()
    check_same(ref, Cs, f"persistent={persistent} largest_first={largest_first} tile={tile}", ps)
    b.close()


def test_autotuned_plan(oz):
    ps = build(PROBLEMS[:8], seed=17)
    ref = [reference(oz, p) for p in ps]
    Cs = [to_dev(p["C0"].copy()) for p in ps]
    d = [desc(p, C) for p, C in zip(ps, Cs)]
    tune = oz.batch_autotune(d, bits=53)
    log(f"autotune: {tune}")
    assert all(t in (0, 64, 128, 256, 256128) for t in tune["tiles"]) and tune["ms"] > 0
    for p, C in zip(ps, Cs):
        C.copy_(to_dev(p["C0"]))
    b = oz2.Batch(oz, d, graph=True, tiles=tune["tiles"], persistent=tune["persistent"],
                  largest_first=tune["largest_first"])
    b.run()
    torch.cuda.synchronize()
    check_same(ref, Cs, "autotuned", ps)
    b.close()
