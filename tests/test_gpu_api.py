"""GPU tests of the handle's host-side contract: stream switching, workspace queries and the
tiling report (the reference's members set_workspace_limit / last_tiling, src/ozaki2_gpu.hpp:702-737,
plus this library's oz2_set_stream and oz2_workspace_query)."""
import numpy as np
import pytest
import torch

import oz2
from gpu_cases import bits, log, real_case, to_dev
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


def test_stream_switch_orders_work():
    # a long call on stream 1, then the handle moves to stream 2 and a second call rewrites the same
    # workspace: the second stream must wait for the first call, so both results equal the oracle
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    h = oz2.Ozaki2(s1)
    try:
        rng = np.random.default_rng(41)
        m = n = k = 2048
        A1, B1 = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
        m2, n2, k2 = 500, 400, 300
        A2, B2 = rng.uniform(-1, 1, m2 * k2), rng.uniform(-1, 1, k2 * n2)
        dA1, dB1, dC1 = to_dev(A1), to_dev(B1), to_dev(np.zeros(m * n))
        dA2, dB2, dC2 = to_dev(A2), to_dev(B2), to_dev(np.zeros(m2 * n2))
        torch.cuda.synchronize()
        log("case stream switch: dgemm 2048^3 s=16 on stream 1, then 500x400x300 s=16 on stream 2")
        h.gemm("N", "N", m, n, k, 1.0, dA1, m, dB1, k, 0.0, dC1, m, 16)
        h.set_stream(s2)
        h.gemm("N", "N", m2, n2, k2, 1.0, dA2, m2, dB2, k2, 0.0, dC2, m2, 16)
        torch.cuda.synchronize()
        ref2 = np.zeros(m2 * n2)
        O.dgemm("N", "N", m2, n2, k2, 1.0, A2, m2, B2, k2, 0.0, ref2, m2, 16)
        assert np.array_equal(bits(dC2.cpu().numpy()), bits(ref2))
        rows = rng.integers(0, m, 200)
        cols = rng.integers(0, n, 200)
        want = O.gemm_entries("N", "N", m, n, k, A1, m, B1, k, 16, rows, cols)
        got = dC1.cpu().numpy()[rows + cols * m]
        assert np.array_equal(bits(got), bits(want))
    finally:
        h.close()


def test_workspace_query_leaves_last_tiling(oz):
    real_case(oz, "N", "N", 300, 200, 100, 12, "uniform", check_bound=False)
    before = oz.last_tiling().as_dict()
    assert oz.workspace_query("N", "T", 5000, 4000, 3000, 16) > 0
    assert oz.workspace_query("N", "N", 64, 64, 64, 8, cplx=True) > 0
    assert oz.last_tiling().as_dict() == before
    # calls that only scale C need no workspace, as oz2_dgemm accepts them without a plan
    assert oz.workspace_query("N", "N", 100, 100, 0, 16) == 0
    assert oz.workspace_query("N", "N", 0, 100, 10, 16) == 0
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz.workspace_query("N", "N", -1, 100, 10, 16)
    with pytest.raises(oz2.OZ2InvalidArgument, match="too small"):
        oz.workspace_query("N", "N", 4, 4, 5000, 2)


def test_query_predicts_call(oz):
    # the bytes a query reports are the bytes the call then uses
    for ta, tb, m, n, k, s in (("N", "N", 700, 600, 900, 14), ("T", "N", 64, 2000, 70000, 18)):
        q = oz.workspace_query(ta, tb, m, n, k, s)
        real_case(oz, ta, tb, m, n, k, s, "uniform", check_bound=False, check_oracle=False)
        assert oz.last_tiling().workspace_bytes == q
