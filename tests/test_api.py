"""Host-side checks of liboz2.so that need no GPU: the library loads, exports
every entry point declared in include/oz2/oz2_c.h, and its host-only plan
queries (precision_bits, moduli_for_bits, moduli) agree with the oracle and
the golden plan table (reference src/ozaki2_gpu.hpp:691-700, 762)."""
import json
import os
import re

import pytest

import oz2
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "oz2", "oz2_c.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(oz2_\w+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(oz2.lib(), name), name
    assert sorted(oz2.EXPORTS) == syms


def test_moduli_match_oracle_and_golden():
    with open(os.path.join(ROOT, "tests", "golden", "plan.json")) as f:
        g = json.load(f)
    sets = {p["s"]: p["moduli"] for p in g["plans"] if "moduli" in p}
    for s in range(2, 25):
        assert oz2.moduli(s) == O.moduli(s) == sets[s]
    for bad in (1, 25, 0, -3):
        with pytest.raises(oz2.OZ2InvalidArgument):
            oz2.moduli(bad)


def test_precision_bits_match_oracle():
    for s in range(2, 25):
        for k in (1, 16, 1000, 1024, 4096, 8192, 131071, 200000, 10 ** 6):
            want = O.precision_bits(k, s)
            if want < 0:
                with pytest.raises(oz2.OZ2InvalidArgument, match="too small"):
                    oz2.Ozaki2.precision_bits(k, s)
            else:
                assert oz2.Ozaki2.precision_bits(k, s) == want
                if O.precision_bits(2 * k, s) > 0:
                    assert oz2.Ozaki2.precision_bits_complex(k, s) == O.precision_bits(2 * k, s)


def test_moduli_for_bits_match_oracle():
    for k in (1, 100, 1024, 2048, 8000, 131071, 200000):
        for bits in (10, 24, 40, 53, 62, 86, 200):
            assert oz2.Ozaki2.moduli_for_bits(k, bits) == O.moduli_for_bits(k, bits)


def test_precision_bits_errors():
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz2.Ozaki2.precision_bits(100, 1)
    with pytest.raises(oz2.OZ2InvalidArgument):
        oz2.Ozaki2.precision_bits(100, 25)
    with pytest.raises(oz2.OZ2InvalidArgument, match="k must be"):
        oz2.Ozaki2.precision_bits(0, 8)


def test_error_message_prefix():
    try:
        oz2.Ozaki2.precision_bits(100, 30)
    except oz2.OZ2InvalidArgument as e:
        assert str(e).startswith("oz2: ")


def test_cpp_dropin_header_compiles_and_links(tmp_path):
    """The C++ drop-in header (include/oz2/ozaki2.hpp) builds the reference-style example and links liboz2.so."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = os.path.join(ROOT, "os2-xe_amd", "examples", "dgemm.cpp")
    lib = os.path.join(ROOT, "os2-xe_amd", "lib")
    out = tmp_path / "dgemm_example"
    r = subprocess.run([hipcc, "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-L", lib, "-loz2",
                        "-o", str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
