"""Seeded randomized parity sweep through the C ABI: DGEMM and ZGEMM with random
shapes (m, n < 320; k up to 3000), every op(A)/op(B), every admissible s, random
alpha/beta, padded leading dimensions, and random tuning (per-plane launches,
plane padding, power-of-two candidates, forced GEMM tiles) and small workspace
limits that force multi-tile runs, column blocks overlapped over two streams.  Every C
must be bit-identical to the oracle on
the same inputs (the reference's algorithm restated; tests/test_ozaki2.cpp:95-194
runs one fixed case per behaviour, this covers their combinations).  Random batches
(oz2_batch_*) of such problems, with random tiles, schedules, order and graph use,
must reproduce the oracle (and oz2_dgemm) bit for bit on every replay."""
import numpy as np
import pytest
import torch

import oz2
from gpu_cases import complex_case, log, real_case, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [101, 202, 303])
def test_random_cases_match_oracle(seed):
    rng = np.random.default_rng(seed)
    oz = oz2.Ozaki2()
    try:
        for c in range(50):
            cplx = rng.random() < 0.3
            m, n = (int(x) for x in rng.integers(1, 320, 2))
            k = int(rng.integers(1, 600)) if rng.random() < 0.5 else int(rng.integers(600, 3000))
            ops = ["N", "T", "C"] if cplx else ["N", "T"]
            ta, tb = str(rng.choice(ops)), str(rng.choice(ops))
            s = int(rng.integers(3, 25))
            while O.precision_bits(2 * k if cplx else k, s) < 0:  # the 3M ZGEMM plans at 2k
                s = int(rng.integers(3, 25))
            dist = str(rng.choice(["uniform", "phi", "ints"]))
            tune = dict(batched_gemm=bool(rng.random() < 0.8), pad_lda=bool(rng.random() < 0.5),
                        pow2_tiles=bool(rng.random() < 0.5), gemm_tile=int(rng.choice([-1, -1, -1, 256, 128, 64, 256128])),
                        overlap_blocks=int(rng.choice([-1, -1, 2, 3])))
            limit = int(rng.choice([4, 8, 16])) << 20 if rng.random() < 0.15 else 0
            oz.set_tuning(**tune)
            oz.set_workspace_limit(limit)
            log(f"random {seed}/{c}: {'z' if cplx else 'd'}gemm {ta}{tb} m={m} n={n} k={k} s={s} {dist} {tune} limit={limit}")
            case_seed = int(rng.integers(1 << 30))
            try:
                if cplx:
                    alpha = complex(*rng.uniform(-2, 2, 2)) if rng.random() < 0.5 else 1 + 0j
                    beta = complex(*rng.uniform(-2, 2, 2)) if rng.random() < 0.5 else 0j
                    complex_case(oz, ta, tb, m, n, k, s, dist, alpha=alpha, beta=beta, ld_extra=int(rng.integers(0, 4)),
                                 seed=case_seed, check_bound=False)
                else:
                    alpha = float(rng.uniform(-2, 2)) if rng.random() < 0.5 else 1.0
                    beta = float(rng.uniform(-2, 2)) if rng.random() < 0.5 else 0.0
                    real_case(oz, ta, tb, m, n, k, s, dist, alpha=alpha, beta=beta, lda_extra=int(rng.integers(0, 4)),
                              ldc_extra=int(rng.integers(0, 4)), seed=case_seed, check_bound=False)
            except oz2.OZ2RuntimeError as e:
                # a workspace limit below one tile's need is reported, not worked around
                assert limit and "not enough device memory" in str(e), e
    finally:
        oz.close()


@pytest.mark.parametrize("seed", [404, 505])
def test_random_batches_match_single_calls(seed):
    # batches of 4-24 random problems (every op, s, alpha/beta, k including 0 and
    # k-chunked lengths), random per-problem tiles, schedule, order and graph use,
    # replayed twice: every problem's C bit-identical to the oracle and to oz2_dgemm with the same s
    from test_gpu_batch import build, check_same, desc, reference
    rng = np.random.default_rng(seed)
    oz = oz2.Ozaki2()
    try:
        for c in range(6):
            problems = []
            for _ in range(int(rng.integers(4, 25))):
                k = int(rng.choice([0, int(rng.integers(1, 700)), int(rng.integers(700, 4000)), 70000 + int(rng.integers(0, 9))]))
                s = int(rng.integers(3, 25))
                while O.precision_bits(max(k, 1), s) < 0:
                    s = int(rng.integers(3, 25))
                m = int(rng.integers(1, 300)) if k < 70000 else int(rng.integers(1, 40))
                n = int(rng.integers(1, 300)) if k < 70000 else int(rng.integers(1, 40))
                problems.append((str(rng.choice(["N", "T"])), str(rng.choice(["N", "T"])), m, n, k, s,
                                 str(rng.choice(["u11", "uniform", "phi", "wide"])),
                                 float(rng.choice([1.0, -1.5, 0.0])), float(rng.choice([0.0, 0.5, -2.0]))))
            ps = build(problems, seed=int(rng.integers(1 << 30)))
            ref = [reference(oz, p) for p in ps]
            tiles = [int(rng.choice([0, 256, 128, 64, 256128])) for _ in ps] if rng.random() < 0.5 else None
            persistent, largest_first, graph = bool(rng.random() < 0.5), bool(rng.random() < 0.5), bool(rng.random() < 0.7)
            log(f"random batch {seed}/{c}: {len(ps)} problems tiles={tiles} persistent={persistent} "
                f"largest_first={largest_first} graph={graph}")
            Cs = [to_dev(p["C0"].copy()) for p in ps]
            b = oz2.Batch(oz, [desc(p, C) for p, C in zip(ps, Cs)], tiles=tiles, persistent=persistent,
                          largest_first=largest_first, graph=graph)
            for _ in range(2):
                for C, p in zip(Cs, ps):
                    C.copy_(to_dev(p["C0"]))
                b.run()
                torch.cuda.synchronize()
                check_same(ref, Cs, f"random batch {seed}/{c}", ps)
            b.close()
    finally:
        oz.close()


@pytest.mark.parametrize("seed", [606])
def test_random_noalloc_matches_oracle(seed):
    # the allocation-free entry point on random shapes, ops, s and alpha/beta in a row (its scratch
    # keeps the previous call's contents), against the oracle
    rng = np.random.default_rng(seed)
    with oz2.Ozaki2() as oz:
        for c in range(20):
            m, n = (int(x) for x in rng.integers(1, 700, 2))
            k = int(rng.integers(1, 5000))
            ta, tb = str(rng.choice(["N", "T"])), str(rng.choice(["N", "T"]))
            s = int(rng.integers(3, 25))
            while O.precision_bits(k, s) < 0:
                s = int(rng.integers(3, 25))
            alpha = float(rng.uniform(-2, 2)) if rng.random() < 0.5 else 1.0
            beta = float(rng.uniform(-2, 2)) if rng.random() < 0.5 else 0.0
            ar, ac = (m, k) if ta == "N" else (k, m)
            br, bc = (k, n) if tb == "N" else (n, k)
            A, B = rng.uniform(-1, 1, ar * ac), rng.uniform(-1, 1, br * bc)
            C0 = rng.uniform(-1, 1, m * n)
            log(f"random noalloc {seed}/{c}: {ta}{tb} m={m} n={n} k={k} s={s} alpha={alpha} beta={beta}")
            dC = to_dev(C0.copy())
            oz.gemm_noalloc(ta, tb, m, n, k, alpha, to_dev(A), ar, to_dev(B), br, beta, dC, m, s)
            torch.cuda.synchronize()
            want = C0.copy()
            O.dgemm(ta, tb, m, n, k, alpha, A, ar, B, br, beta, want, m, s)
            assert np.array_equal(dC.cpu().numpy().view(np.uint64), want.view(np.uint64))
