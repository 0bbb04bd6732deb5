"""Throughput and per-stage timing of the emulated DGEMM: the counterpart of the
reference's benchmarks/bench_ozaki2.cpp.

Same options (--sizes, --s, --reps, --dgemm, --cpu, --check, --dist), the same inputs
(liboz2_check.so oz2chk_bench_fill: sin / cos by default, the hashed phiX
distribution with --dist phiX; A seeded 1, B 12345), the same measurements and the
same output lines:

* a 256^3 s=8 warm-up call, then the workspace is released;
* per size and s: a cold call after release_workspace (workspace allocation
  included), then --reps timed calls, best and median;
* stage times and tiling of the last call;
* with --dgemm (or --check), the GPU's native FP64 GEMM (rocBLAS through
  torch.matmul, the reference's oneMKL call), best of 3 (1 above 2 TFLOP);
* with --cpu, numpy's BLAS (one warm-up, one timed call);
* with --check, max |C - C_dgemm| / (|A||B|).

Sizes are MxKxN (or one number for a cube), as in the reference.  --tuning applies
Ozaki2.set_tuning before the sweep, e.g. --tuning batched_gemm=0,pad_lda=0.

    python -m oz2.bench_ozaki2 --sizes 4096,8192 --s 8,12,16 --dgemm --check
"""
import argparse
import ctypes
import sys
import time

import torch

import oz2
from oz2 import check

DEFAULT_SIZES = "2048,4096,4000x8000x4000,8192,16384,20480,24576"
TUNING_KEYS = ("batched_gemm", "pad_lda", "pow2_tiles", "gemm_tile")


def parse_sizes(text):
    """'4096,4000x8000x4000' -> [(4096, 4096, 4096), (4000, 8000, 4000)] as (m, k, n)"""
    out = []
    for part in text.split(","):
        if not part:
            continue
        d = [int(v) for v in part.split("x")]
        out.append((d[0], d[0], d[0]) if len(d) == 1 else (d[0], d[1], d[2]))
    return out


def parse_tuning(text):
    """'batched_gemm=0,gemm_tile=128' -> {'batched_gemm': 0, 'gemm_tile': 128}"""
    out = {}
    for part in text.split(","):
        if not part:
            continue
        key, _, val = part.partition("=")
        if key not in TUNING_KEYS or not val:
            raise ValueError(f"unknown tuning entry {part!r} (keys: {', '.join(TUNING_KEYS)})")
        out[key] = int(val)
    return out


def fill(X, rows, cols, is_a, phi, seed):
    L = check.lib()
    L.oz2chk_bench_fill.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                    ctypes.c_double, ctypes.c_uint64]
    rc = L.oz2chk_bench_fill(None, X.data_ptr(), rows, cols, 1 if is_a else 0, phi, seed)
    if rc != 0:
        raise RuntimeError(f"oz2: benchmark fill failed ({rc})")


def _sync_time(fn):
    t0 = time.perf_counter()
    fn()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def run_size(oz, m, k, n, svals, reps, phi, do_dgemm, do_cpu, do_check, out):
    flops = 2.0 * m * n * k
    dA = torch.empty(m * k, dtype=torch.float64, device="cuda")
    dB = torch.empty(k * n, dtype=torch.float64, device="cuda")
    dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
    fill(dA, m, k, True, phi, 1)
    fill(dB, k, n, False, phi, 12345)
    tA = dA.view(k, m).t()  # column-major m x k
    tB = dB.view(n, k).t()
    t_dgemm, ref, absab = 0.0, None, None
    if do_dgemm or do_check:
        res = []

        def native():
            res[:] = [tA @ tB]

        native()
        torch.cuda.synchronize()
        t_dgemm = min(_sync_time(native) for _ in range(1 if flops > 2e12 else 3))
        ref = res[0].t().contiguous().view(-1)  # column-major m x n
        del res
    t_cpu = 0.0
    if do_cpu:
        Ah = dA.cpu().numpy().reshape(k, m).T
        Bh = dB.cpu().numpy().reshape(n, k).T
        Ah @ Bh
        t0 = time.perf_counter()
        Ah @ Bh
        t_cpu = time.perf_counter() - t0
    if do_check and ref is not None:
        absab = (tA.abs() @ tB.abs()).t().contiguous().view(-1)
    rows = []
    for s in svals:
        try:
            tk = oz2.Ozaki2.precision_bits(k, s)
            oz.release_workspace()

            def call():
                oz.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)

            cold = _sync_time(call)
            ts = sorted(_sync_time(call) for _ in range(reps))
            st = oz.stage_times()
            ti = oz.last_tiling()
        except (oz2.OZ2Error, RuntimeError) as e:
            print(f"  {m}x{k}x{n} s={s} FAILED: {e}", file=out, flush=True)
            continue
        best, med = ts[0], ts[len(ts) // 2]
        line = (f"{m:6d}x{k:6d}x{n:6d} s={s:2d} tk={tk:2d} | {best * 1e3:7.3f} ms  {flops / best / 1e12:7.3f} TFLOPS "
                f"(median {flops / med / 1e12:.3f}) | cold {cold:.3f} s | scale {st.scale_ms:.2f} split {st.split_ms:.2f} "
                f"gemm {st.gemm_ms:.2f} ({flops * s / (st.gemm_ms * 1e-3) / 1e12 if st.gemm_ms > 0 else 0.0:.0f} TOPS) "
                f"crt {st.crt_ms:.2f} ms | tiles {ti.pm}x{ti.pn} ws {ti.workspace_bytes / 1e9:.2f} GB")
        if t_dgemm > 0:
            line += f" | dgemm {flops / t_dgemm / 1e12:.3f} TFLOPS (x{t_dgemm / best:.1f})"
        if t_cpu > 0:
            line += f" | cpu {flops / t_cpu / 1e12:.3f} TFLOPS (x{t_cpu / best:.1f})"
        if absab is not None:
            pos = absab > 0
            err = ((dC - ref).abs()[pos] / absab[pos]).max().item() if bool(pos.any()) else 0.0
            line += f" | max|C-C_dgemm|/(|A||B|) {err:.2e}"
        print(line, file=out, flush=True)
        rows.append({"m": m, "k": k, "n": n, "s": s, "tk": tk, "best": best, "median": med, "cold": cold,
                     "dgemm": t_dgemm, "cpu": t_cpu, "flops": flops})
    oz.release_workspace()
    return rows


def main(argv=None, out=sys.stdout):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--sizes", default=DEFAULT_SIZES, help="MxKxN or N, comma separated")
    ap.add_argument("--s", default="8,12,16")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--dgemm", action="store_true")
    ap.add_argument("--cpu", action="store_true")
    ap.add_argument("--check", action="store_true")
    ap.add_argument("--dist", default="sincos", help="sincos or phiX (X >= 0)")
    ap.add_argument("--tuning", default="", help="Ozaki2.set_tuning entries, e.g. batched_gemm=0,pad_lda=0")
    args = ap.parse_args(argv)
    phi = float(args.dist[3:]) if args.dist.startswith("phi") else -1.0
    svals = [int(v) for v in args.s.split(",") if v]
    sizes = parse_sizes(args.sizes)
    tuning = parse_tuning(args.tuning)
    if args.reps < 1:
        raise ValueError("--reps must be at least 1")
    oz = oz2.Ozaki2()
    try:
        oz.set_profiling(True)
        if tuning:
            oz.set_tuning(**tuning)
        print(f"device: {torch.cuda.get_device_name(0)}  tuning: {oz.tuning()}", file=out, flush=True)
        w = torch.zeros(3 * 256 * 256, dtype=torch.float64, device="cuda")
        oz.gemm("N", "N", 256, 256, 256, 1.0, w, 256, w[65536:], 256, 0.0, w[131072:], 256, 8)
        torch.cuda.synchronize()
        oz.release_workspace()
        del w
        rows = []
        for m, k, n in sizes:
            rows += run_size(oz, m, k, n, svals, args.reps, phi, args.dgemm, args.cpu, args.check, out)
    finally:
        oz.close()
    print(f"\n{'size (MxKxN)':<22s} {'s':>3s} {'tk':>3s} {'TFLOPS':>10s} {'ms':>10s} {'vs GPU DGEMM':>12s} {'vs CPU':>10s}",
          file=out)
    for r in rows:
        line = f"{str(r['m']) + 'x' + str(r['k']) + 'x' + str(r['n']):<22s} {r['s']:3d} {r['tk']:3d} " \
               f"{r['flops'] / r['best'] / 1e12:10.3f} {r['best'] * 1e3:10.2f}"
        line += f" {r['dgemm'] / r['best']:11.1f}x" if r["dgemm"] > 0 else f" {'-':>12s}"
        line += f" {r['cpu'] / r['best']:9.1f}x" if r["cpu"] > 0 else f" {'-':>10s}"
        print(line, file=out)
    out.flush()
    return rows


if __name__ == "__main__":
    main()
