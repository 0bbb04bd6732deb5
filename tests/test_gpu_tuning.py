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
    assert oz.tuning() == {"batched_gemm": True, "pad_lda": True, "pow2_tiles": True, "gemm_tile": -1}
    with tuned(oz, batched_gemm=False, pad_lda=False, pow2_tiles=False, gemm_tile=128):
        assert oz.tuning() == {"batched_gemm": False, "pad_lda": False, "pow2_tiles": False, "gemm_tile": 128}
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
