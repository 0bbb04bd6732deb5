"""Host pieces of the throughput benchmark (oz2/bench_ozaki2.py, reference
benchmarks/bench_ozaki2.cpp): the input hash pinned to the published splitmix64
outputs, and the option parsers (sizes are MxKxN as in the reference)."""
import pytest

from bench_inputs import GAMMA, M64, splitmix64
from oz2 import bench_ozaki2


def test_splitmix64_known_answers():
    # splitmix64 seeded with 1234567: the first five outputs of the published generator
    want = [6457827717110365317, 3203168211198807973, 9817491932198370423, 4593380528125082431,
            16408922859458223821]
    assert [splitmix64((1234567 + i * GAMMA) & M64) for i in range(5)] == want


def test_parse_sizes_and_tuning():
    assert bench_ozaki2.parse_sizes("2048,4000x8000x3000") == [(2048, 2048, 2048), (4000, 8000, 3000)]
    assert len(bench_ozaki2.parse_sizes(bench_ozaki2.DEFAULT_SIZES)) == 7
    assert bench_ozaki2.parse_tuning("batched_gemm=0,gemm_tile=256128") == {"batched_gemm": 0, "gemm_tile": 256128}
    assert bench_ozaki2.parse_tuning("") == {}
    with pytest.raises(ValueError):
        bench_ozaki2.parse_tuning("batched=0")
    with pytest.raises(ValueError):
        bench_ozaki2.parse_tuning("pad_lda")
