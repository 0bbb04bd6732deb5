"""GPU test of oz2_dgemm_noalloc, the DGEMM entry point that allocates nothing:
its C must be bit-identical to oz2_dgemm's on the same inputs, over calls in a
row that leave the shared scratch dirty (the second smaller than the first),
over a product that needs several panels to fit the scratch, and it must
reject a product whose smallest panels do not fit without touching C."""
import numpy as np
import pytest
import torch

import oz2
from gpu_cases import bits, log, to_dev

pytestmark = pytest.mark.gpu


def operands(rng, ta, tb, m, n, k):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    return rng.uniform(-1, 1, ar * ac), ar, rng.uniform(-1, 1, br * bc), br


def both(oz, ta, tb, m, n, k, s, alpha, beta, seed):
    """(C from oz2_dgemm, C from oz2_dgemm_noalloc) on the same inputs"""
    rng = np.random.default_rng(seed)
    A, lda, B, ldb = operands(rng, ta, tb, m, n, k)
    ldc = m + 3
    C0 = rng.uniform(-1, 1, ldc * n)
    dA, dB = to_dev(A), to_dev(B)
    d1, d2 = to_dev(C0.copy()), to_dev(C0.copy())
    torch.cuda.synchronize()
    log(f"case noalloc {ta}{tb} m={m} n={n} k={k} s={s} alpha={alpha} beta={beta} seed={seed}")
    oz.gemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, d1, ldc, s)
    oz.gemm_noalloc(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, d2, ldc, s)
    torch.cuda.synchronize()
    return d1.cpu().numpy(), d2.cpu().numpy()


def test_noalloc_equals_dgemm_on_calls_in_a_row():
    with oz2.Ozaki2() as oz:
        # first call fills the scratch, the second (smaller, other ops and s) reuses it dirty
        for case in (("N", "T", 900, 700, 1500, 16, 1.5, 0.5, 1), ("T", "N", 100, 37, 300, 8, 1.0, 0.0, 2),
                     ("C", "C", 65, 129, 1, 24, -2.0, 0.25, 3)):
            c1, c2 = both(oz, *case)
            assert np.array_equal(bits(c1), bits(c2)), case


def test_noalloc_plans_panels_to_fit_scratch():
    with oz2.Ozaki2() as oz:
        # the planes of all of op(A) alone (16 x 2048 x 4096 bytes = 128 MiB) exceed the scratch
        c1, c2 = both(oz, "N", "N", 2048, 1536, 4000, 16, 1.0, 0.5, 4)
        ti = oz.last_tiling()
        log(f"noalloc tiling {ti.as_dict()}")
        assert ti.pm * ti.pn > 1 and ti.workspace_bytes <= oz2.NOALLOC_SCRATCH_BYTES
        assert np.array_equal(bits(c1), bits(c2))
        # two k-chunks (64-row panels of 6 planes of 70016 bytes: 27 MiB per operand), then a smaller product
        c1, c2 = both(oz, "T", "N", 200, 300, 70000, 6, 1.0, 0.0, 5)
        assert oz.last_tiling().k_chunks == 2
        assert np.array_equal(bits(c1), bits(c2))
        c1, c2 = both(oz, "N", "N", 33, 17, 9, 4, 1.0, 0.0, 6)
        assert np.array_equal(bits(c1), bits(c2))


def test_noalloc_rejects_what_does_not_fit_and_allocates_nothing():
    with oz2.Ozaki2() as oz:
        rng = np.random.default_rng(7)
        m = n = 64
        k = 200000  # 64 rows of 24 planes of 200064 bytes per operand: 586 MiB
        dA, dB = to_dev(rng.uniform(-1, 1, m * k)), to_dev(rng.uniform(-1, 1, k * n))
        C0 = rng.uniform(-1, 1, m * n)
        dC = to_dev(C0.copy())
        torch.cuda.synchronize()
        log("case noalloc too large: 64 x 64 x 200000 s=24")
        with pytest.raises(oz2.OZ2InvalidArgument, match="too large for the fixed workspace"):
            oz.gemm_noalloc("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, 24)
        torch.cuda.synchronize()
        assert np.array_equal(bits(dC.cpu().numpy()), bits(C0))
        # device memory is the same before and after a call on a fresh handle (whose own
        # workspace is empty); the first call of the entry point on this handle comes after
        # the inputs exist
        m, n, k = 1000, 800, 2000
        A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
        dA, dB, dC = to_dev(A), to_dev(B), to_dev(np.zeros(m * n))
        torch.cuda.synchronize()
    with oz2.Ozaki2() as fresh:
        fresh.gemm_noalloc("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, 16)  # kernels loaded, runtime warm
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        for _ in range(3):
            fresh.gemm_noalloc("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, 16)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        log(f"free device memory {free0} -> {free1}")
        assert free1 == free0
