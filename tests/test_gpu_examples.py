"""The C++ drop-in (include/oz2/ozaki2.hpp over liboz2.so) run on the GPU through
the reference's usage examples, rebuilt MI355X-side (os2-xe_amd/examples, built by
`make -C os2-xe_amd`): zgemm.cpp (examples/zgemm.cpp:49, zgemm(...).wait() with
op(A) = C and complex alpha, beta) and profiling.cpp (examples/profiling.cpp:30-42:
set_resident_hint, stage_times, last_tiling over s = 8..20), plus dgemm.cpp.  Each
runs as its own process; the C it writes is checked bit for bit against the
oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from gpu_cases import bits, log
from oracle import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "os2-xe_amd", "examples", "bin")


def run(name, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C os2-xe_amd examples)")
    log(f"example {name} {' '.join(args)}")
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=120)
    log(r.stdout)
    assert r.returncode == 0, r.stderr[-2000:]
    return r.stdout


def test_zgemm_example(tmp_path):
    m, n, k, s = 300, 200, 500, 16
    out = run("zgemm", str(m), str(n), str(k), str(s), "--dump", str(tmp_path))
    assert f"tk={O.precision_bits(2 * k, s)}" in out
    rd = lambda f: np.fromfile(tmp_path / f, dtype=np.float64)
    A, B, C0, C = rd("A.bin"), rd("B.bin"), rd("C0.bin"), rd("C.bin")
    want = C0.copy()
    O.zgemm("C", "N", m, n, k, complex(1.0, 0.5), A, k, B, k, complex(0.25, -1.0), want, m, s)
    assert np.array_equal(bits(C), bits(want))
    rel = float(out.split("max |C_oz2 - C_fp64| / max |C_fp64| = ")[1].split()[0])
    assert rel < 1e-14


def test_profiling_example(tmp_path):
    N = 700
    out = run("profiling", str(N), "--dump", str(tmp_path))
    rows = [ln.split() for ln in out.splitlines() if ln.split() and ln.split()[0] in ("8", "12", "14", "16", "20")]
    assert [int(r[0]) for r in rows] == [8, 12, 14, 16, 20]
    for r in rows:
        s, tk = int(r[0]), int(r[1])
        assert tk == O.precision_bits(N, s)
        scale, split, gemm, crt, wall, tflops = (float(x) for x in r[2:8])
        assert min(scale, split, gemm, crt) > 0 and wall > 0 and tflops > 0
        assert r[8] == "1x1"
    rd = lambda f: np.fromfile(tmp_path / f, dtype=np.float64)
    A, B, C = rd("A.bin"), rd("B.bin"), rd("C.bin")
    i = np.arange(N * N, dtype=np.float64)
    # the device's sin/cos of the index: within an ulp of the host's
    assert np.max(np.abs(A - np.sin(i))) < 1e-15 and np.max(np.abs(B - np.cos(i))) < 1e-15
    want = np.zeros(N * N)
    O.dgemm("N", "N", N, N, N, 1.0, A, N, B, N, 0.0, want, N, 20)
    assert np.array_equal(bits(C), bits(want))


def test_dgemm_example():
    out = run("dgemm", "300", "200", "100")
    assert f"moduli s={O.moduli_for_bits(100, 53)}" in out and "batch: learned s =" in out
    assert " schedule, " in out  # the autotuned plan ran through the C++ wrapper
    assert "ozaki1: L=9, 45 INT8 GEMMs, 61 bits per element" in out
