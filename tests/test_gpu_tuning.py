"""Tuning knobs (reference ozaki2_gpu.hpp:41-45 `Tuning`, set_tuning / tuning at
hpp:712-737): each knob changes how planes are laid out, how the GEMMs are
launched or which tiles are considered, never the result.  Every case is
bit-identical to the oracle (or to the default tuning on the same inputs):
per-plane GEMM launches (hpp:1028-1041, batched_gemm = false), unpadded plane
strides (hpp:669-672, pad_lda = false) and no power-of-two tile candidates
(hpp:1132-1140, pow2_tiles = false)."""
import contextlib

import numpy as np
import pytest

import oz2
from gpu_cases import bits, complex_case, log, real_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


@contextlib.contextmanager
def tuned(oz, **kw):
    oz.set_tuning(**kw)
    try:
        yield
    finally:
        oz.set_tuning()


def test_roundtrip(oz):
    assert oz.tuning() == {"batched_gemm": True, "pad_lda": True, "pow2_tiles": True, "gemm_tile": -1,
                           "overlap_blocks": -1}
    with tuned(oz, batched_gemm=False, pad_lda=False, pow2_tiles=False, gemm_tile=128, overlap_blocks=3):
        assert oz.tuning() == {"batched_gemm": False, "pad_lda": False, "pow2_tiles": False, "gemm_tile": 128,
                               "overlap_blocks": 3}
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz.set_tuning(overlap_blocks=17)
    assert oz.tuning()["batched_gemm"] is True


@pytest.mark.parametrize("tile", [-1, 256, 128, 64, 256128])
def test_per_plane_gemm_matches_oracle(oz, tile):
    with tuned(oz, batched_gemm=False, gemm_tile=tile):
        real_case(oz, "N", "T", 150, 130, 300, 13, "phi", check_bound=False)


def test_per_plane_k_chunks_and_tiling(oz):
    with tuned(oz, batched_gemm=False):
        real_case(oz, "N", "T", 40, 36, 65409, 16, "phi")  # two k-chunks, each one launch per plane
        assert oz.last_tiling().k_chunks == 2
        r1 = real_case(oz, "N", "T", 1000, 900, 512, 10, "phi", seed=5, check_bound=False)
        oz.set_workspace_limit(8 << 20)
        try:
            r2 = real_case(oz, "N", "T", 1000, 900, 512, 10, "phi", seed=5, check_bound=False)
            ti = oz.last_tiling()
        finally:
            oz.set_workspace_limit(0)
    log(f"per-plane tiling {ti.as_dict()}")
    assert ti.pm * ti.pn > 1
    assert np.array_equal(bits(r1["C"]), bits(r2["C"]))


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("C", "T")])
def test_per_plane_complex(oz, ta, tb):
    # 3s planes (Re, Im, Re + Im groups): plane t of the launch sequence uses modulus t mod s
    with tuned(oz, batched_gemm=False):
        complex_case(oz, ta, tb, 45, 38, 150, 14, "phi", alpha=0.5 - 2j, beta=1 + 1j, check_bound=False)


def test_pad_lda(oz):
    # m = n = k = 256: every plane stride (A, B and residues) is 64 KiB, a multiple of the
    # pad period, so the padded layout is 4 KiB longer per plane of each of the three kinds
    r1 = real_case(oz, "N", "N", 256, 256, 256, 16, "phi", check_bound=False)
    ws_pad = oz.last_tiling().workspace_bytes
    with tuned(oz, pad_lda=False):
        r2 = real_case(oz, "N", "N", 256, 256, 256, 16, "phi", check_bound=False)
        ws_plain = oz.last_tiling().workspace_bytes
        complex_case(oz, "N", "N", 256, 256, 256, 16, "uniform", check_bound=False)
    assert np.array_equal(bits(r1["C"]), bits(r2["C"]))
    assert ws_pad - ws_plain == 3 * 16 * 4096
    # a stride that is not a multiple of 64 KiB is left alone
    real_case(oz, "N", "N", 200, 120, 104, 16, "phi", check_bound=False)
    ws_a = oz.last_tiling().workspace_bytes
    with tuned(oz, pad_lda=False):
        real_case(oz, "N", "N", 200, 120, 104, 16, "phi", check_bound=False)
        assert oz.last_tiling().workspace_bytes == ws_a


def test_no_pow2_tiles(oz):
    with tuned(oz, pow2_tiles=False):
        oz.set_workspace_limit(8 << 20)
        try:
            real_case(oz, "T", "N", 1100, 700, 300, 12, "phi", seed=7, check_bound=False)
            ti = oz.last_tiling()
        finally:
            oz.set_workspace_limit(0)
    log(f"no-pow2 tiling {ti.as_dict()}")
    assert ti.pm * ti.pn > 1


@pytest.mark.parametrize("blocks", [2, 3, 5])
def test_column_block_overlap_matches_oracle(oz, blocks):
    # column blocks of each panel pair over two streams (B split and CRT of one block beside the GEMM
    # of another), on real and complex data, with several panels, k-chunks and per-plane launches
    with tuned(oz, overlap_blocks=blocks):
        real_case(oz, "N", "N", 700, 900, 500, 16, "phi", alpha=1.5, beta=0.5, check_bound=False)
        real_case(oz, "T", "T", 300, 1000, 70000, 12, "u11", check_bound=False)
        assert oz.last_tiling().k_chunks == 2
        complex_case(oz, "C", "N", 260, 700, 300, 16, "phi", alpha=1 + 0.5j, beta=0.25 - 1j, check_bound=False)
        oz.set_workspace_limit(8 << 20)
        try:
            real_case(oz, "N", "T", 1000, 1100, 400, 10, "phi", seed=5, check_bound=False)
            ti = oz.last_tiling()
        finally:
            oz.set_workspace_limit(0)
        assert ti.pm * ti.pn > 1
    with tuned(oz, overlap_blocks=blocks, batched_gemm=False):
        real_case(oz, "N", "T", 500, 777, 300, 9, "phi", check_bound=False)


def test_overlap_stage_times_and_graph_capture(oz):
    # stage times come from both streams; a capture of the caller's stream takes the second stream's
    # work into the graph (event fork and join), replayed several times with eager work in between
    import torch
    from gpu_cases import to_dev
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    m, n, k, s = 1024, 2048, 1024, 16
    A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    dA, dB, dC = to_dev(A), to_dev(B), to_dev(np.zeros(m * n))
    want = np.zeros(m * n)
    O.dgemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, want, m, s)
    st = torch.cuda.Stream()
    h = oz2.Ozaki2(st)
    try:
        h.set_tuning(overlap_blocks=4)
        with torch.cuda.stream(st):
            h.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        st.synchronize()
        t = h.stage_times()
        assert min(t.scale_ms, t.split_ms, t.gemm_ms, t.crt_ms) > 0
        assert np.array_equal(bits(dC.cpu().numpy()), bits(want))
        # during capture: no stage events, no device-memory queries (a fixed limit), no growth
        h.set_workspace_limit(4 << 30)
        h.reserve_workspace(h.workspace_query("N", "N", m, n, k, s))
        h.set_profiling(False)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            h.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        for rep in range(3):
            dC.zero_()
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(bits(dC.cpu().numpy()), bits(want)), f"replay {rep}"
    finally:
        h.close()


def test_capture_with_default_settings():
    # a capture of a plain call with the default settings (stage timing on, no workspace limit): the
    # library records no timing events and plans for the workspace reserved before the capture;
    # every replay, with eager work in between, equals the eager result
    import torch
    from gpu_cases import to_dev
    rng = np.random.default_rng(8)
    m, n, k, s = 700, 900, 1500, 16
    A, B = rng.uniform(-1, 1, m * k), rng.uniform(-1, 1, k * n)
    C0 = rng.uniform(-1, 1, m * n)
    dA, dB = to_dev(A), to_dev(B)
    st = torch.cuda.Stream()
    h = oz2.Ozaki2(st)
    try:
        eager = to_dev(C0.copy())
        with torch.cuda.stream(st):
            h.gemm("N", "N", m, n, k, 1.5, dA, m, dB, k, 0.5, eager, m, s)
        st.synchronize()
        want = eager.cpu().numpy()
        h.reserve_workspace(h.workspace_query("N", "N", m, n, k, s))
        dC = to_dev(C0.copy())
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            h.gemm("N", "N", m, n, k, 1.5, dA, m, dB, k, 0.5, dC, m, s)
        for rep in range(3):
            dC.copy_(to_dev(C0))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(bits(dC.cpu().numpy()), bits(want)), f"replay {rep}"
    finally:
        h.close()
