"""The grouped-DGEMM benchmark workload at full size, checked against the oracle
in the pattern of the reference's run_case (tests/test_ozaki2.cpp:95-193):
256 DGEMMs with m, n, k log-uniform in [64, 4096], half U[-1,1) and half
(1 + U) * (+-1) * 2^e with e in [-12, 12], a 53-bit target, moduli chosen by the
learned selector, macro tiles and schedule from the autotuner, and the launch
sequence replayed from a hipGraph.  For every problem, sampled entries (the four
corners among them) must be bit-identical to the exact closed form at the s the
selector chose, and within the target (1 ulp of the double-double product) -
the oracle decides, not another library call."""
import numpy as np
import pytest
import torch

import oz2
from gpu_cases import bits, log
from oracle import oracle as O
from oz2.workloads import uneven_batch

pytestmark = pytest.mark.gpu


def ulp_of(hi):
    a = np.abs(hi)
    return np.where(hi == 0, 5e-324, np.nextafter(a, np.inf) - a)


@pytest.mark.slow
def test_uneven_batch_learned_graph_against_oracle():
    rng = np.random.default_rng(55)
    with oz2.Ozaki2() as oz:
        probs = uneven_batch(torch, 256, seed=5)
        descs = [dict(p, num_moduli=0) for p in probs]
        tune = oz.batch_autotune(descs, bits=53, select=True)
        log(f"config-5 batch: autotuned persistent={tune['persistent']} largest_first={tune['largest_first']}")
        b = oz2.Batch(oz, descs, bits=53, graph=True, select=True, tiles=tune["tiles"], persistent=tune["persistent"],
                      largest_first=tune["largest_first"])
        try:
            assert b.info()["graph"]
            for p in probs:
                p["C"].fill_(np.nan)
            b.run()
            for p in probs:  # the replay must recompute every entry
                p["C"].fill_(np.nan)
            b.run()
            torch.cuda.synchronize()
            chosen = b.moduli()
        finally:
            b.close()
    log(f"config-5 batch: mean s {np.mean(chosen):.2f}, s in [{min(chosen)}, {max(chosen)}]")
    worst = 0.0
    for i, (p, s) in enumerate(zip(probs, chosen)):
        m, n, k = p["m"], p["n"], p["k"]
        assert 2 <= s <= 24 and O.precision_bits(k, s) > 0
        rows = np.unique(np.concatenate([[0, m - 1], rng.integers(0, m, 10)]))
        cols = np.unique(np.concatenate([[0, n - 1], rng.integers(0, n, 10)]))
        # the sampled rows of op(A) and columns of op(B) carry their own exponents, so the
        # sub-product of those rows and columns has the same entries
        Asub = p["A"].reshape(k, m)[:, torch.as_tensor(rows, device="cuda")].contiguous().cpu().numpy().reshape(-1)
        Bsub = p["B"].reshape(n, k)[torch.as_tensor(cols, device="cuda")].contiguous().cpu().numpy().reshape(-1)
        R, Cn = rows.size, cols.size
        ri, ci = np.repeat(np.arange(R), Cn), np.tile(np.arange(Cn), R)
        want = O.gemm_entries("N", "N", R, Cn, k, Asub, R, Bsub, k, s, ri, ci)
        got = p["C"].cpu().numpy()[rows[ri] + cols[ci] * m]
        nbad = int(np.count_nonzero(bits(got) != bits(want)))
        assert nbad == 0, f"problem {i} ({m}x{n}x{k}, s={s}): {nbad} of {ri.size} sampled entries differ from the oracle"
        hi, lo = O.dd_entries("N", "N", k, Asub, R, Bsub, k, ri, ci)
        ulp = np.abs((got - hi) - lo) / ulp_of(hi)
        worst = max(worst, float(ulp.max()))
        assert ulp.max() <= 1.0, f"problem {i} ({m}x{n}x{k}, wide={p['wide']}): s={s} misses 53 bits ({ulp.max():.3g} ulp)"
    log(f"config-5 batch: worst sampled error {worst:.3f} ulp")
