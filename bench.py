"""Throughput of the MI355X Ozaki Scheme II DGEMM (one JSON line on stdout).

Workload (one step): C = 1.5*A*B + 0.5*C, DGEMM m = n = k = 8192, op N/N,
s = 16 moduli, A, B, C0 i.i.d. U[-1, 1) (synthetic, seeded), inputs resident
in HBM.  Metric: FP64-equivalent TFLOP/s = 2 m n k / t, reported with the
max / mean ulp error and correctly-rounded fraction of the same inputs'
product against a double-double reference.

Also reported:
  roofline      the dominant kernel (the INT8 MFMA GEMM, one launch of s planes
                per call) against the INT8 dense peak, its time measured with
                HIP events on the library's stream over the timed region;
  cpu_baseline  the oracle's CPU restatement of plain Ozaki-II (OpenMP, all
                host cores) on a bounded sample of the same workload;
  sweep         the other op(A)/op(B) and s values of the same configuration.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "os2-xe_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

INT8_PEAK_TOPS = 5000.0  # MI355X dense INT8 MFMA, 5.0 POP/s
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--m", type=int, default=8192)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--k", type=int, default=8192)
    ap.add_argument("--s", type=int, default=16)
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    return ap.parse_args()


def load_check_lib():
    import ctypes
    path = os.path.join(ROOT, "os2-xe_amd", "lib", "liboz2_check.so")
    L = ctypes.CDLL(path)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    L.oz2chk_dd_gemm.argtypes = [vp, i32, i32, i64, i64, i64, vp, i64, vp, i64, vp, vp, i32]
    return L


def accuracy(oz, torch, dA, dB, m, n, k, s):
    """ulp_max, ulp_mean, cr_frac of alpha=1, beta=0 against double-double (bench_accuracy.cpp:71-94)."""
    import numpy as np
    chk = load_check_lib()
    dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
    oz.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
    hi = torch.empty_like(dC)
    lo = torch.empty_like(dC)
    rc = chk.oz2chk_dd_gemm(None, 0, 0, m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, hi.data_ptr(), lo.data_ptr(), 0)
    assert rc == 0
    torch.cuda.synchronize()
    C, H, L = dC.cpu().numpy(), hi.cpu().numpy(), lo.cpu().numpy()
    err = np.abs((C - H) - L)
    aH = np.abs(H)
    ulp = np.where(H == 0, np.float64(5e-324), np.nextafter(aH, np.inf) - aH)
    eu = err / ulp
    return {"ulp_max": float(eu.max()), "ulp_mean": float(eu.mean()), "cr_frac": float(np.mean(C == H + L)),
            "reference": "double-double GEMM on the GPU, same inputs, alpha=1 beta=0"}


def cpu_baseline(m, n, k, s):
    """Oracle (CPU restatement, OpenMP) on the first `rows` rows of op(A): same k, n, s, distribution."""
    import numpy as np
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    rng = np.random.default_rng(1)
    rows = 3072  # about 10-15 s of OpenMP work on 16 host cores
    A = rng.uniform(-1, 1, rows * k)
    B = rng.uniform(-1, 1, k * n)
    C = rng.uniform(-1, 1, rows * n)
    O.dgemm("N", "N", 8, 64, 256, 1.5, A, rows, B, k, 0.5, C, rows, s, threads)  # load / warm
    t0 = time.perf_counter()
    O.dgemm("N", "N", rows, n, k, 1.5, A, rows, B, k, 0.5, C, rows, s, threads)
    dt = time.perf_counter() - t0
    return {"value": 2.0 * rows * n * k / dt / 1e12, "unit": "FP64-equiv TFLOP/s", "cores": threads, "kind": "port",
            "seconds": dt,
            "sample": f"rows 0..{rows - 1} of the {m}x{n}x{k} s={s} DGEMM (alpha=1.5, beta=0.5): {rows}x{n}x{k}, "
                      f"oracle residue route (split, exact int8 GEMM, CRT) with OpenMP"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import oz2

    m, n, k, s = args.m, args.n, args.k, args.s
    alpha, beta = 1.5, 0.5
    g = torch.Generator(device="cuda")
    g.manual_seed(1)
    dA = torch.rand(m * k, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dB = torch.rand(k * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dC0 = torch.rand(m * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dC = dC0.clone()
    oz = oz2.Ozaki2()
    oz.set_profiling(True)

    for _ in range(args.warmup):
        oz.gemm("N", "N", m, n, k, alpha, dA, m, dB, k, beta, dC, m, s)
    torch.cuda.synchronize()
    oz.profile_begin()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        oz.gemm("N", "N", m, n, k, alpha, dA, m, dB, k, beta, dC, m, s)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st, calls = oz.profile_end()
    ms = (t1 - t0) * 1e3 / args.steps
    tiling = oz.last_tiling().as_dict()
    flops = 2.0 * m * n * k
    value = flops / (ms * 1e-3) / 1e12

    gemm_ms = st.gemm_ms / max(1, calls)  # one GEMM launch per call at this size (one tile, one k-chunk)
    int8_ops = 2.0 * m * n * k * s
    achieved = int8_ops / (gemm_ms * 1e-3) / 1e12
    stage = {kk: v / max(1, calls) for kk, v in st.as_dict().items()}

    out = {
        "metric": "FP64-equiv TFLOP/s at stated max-ulp: DGEMM 8192^3 s=16",
        "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 (INT8 MFMA, int32 accumulate; f64 in/out)", "data": "synthetic U[-1,1) seeded, HBM-resident",
        "config": {"workload": f"DGEMM {m}x{n}x{k} op N/N s={s} alpha=1.5 beta=0.5, A B C0 U[-1,1)", "m": m, "n": n,
                   "k": k, "s": s, "tk": oz2.Ozaki2.precision_bits(k, s), "tiling": tiling},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / INT8_PEAK_TOPS, 4), "traffic": None,
                     "kernel": "igemm_residue_kernel (2*m*n*k*s int8 ops per launch)",
                     "gemm_ms": round(gemm_ms, 4)},
        "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
        "int8_rate_end_to_end": round(int8_ops / (ms * 1e-3) / 1e12, 1),
        "reference_b60_tflops": 9.3,
    }
    if not args.no_accuracy:
        out["accuracy"] = accuracy(oz, torch, dA, dB, m, n, k, s)
    if not args.no_sweep:
        sweep = {}
        for sv in (8, 12, 16):
            for ta, tb in (("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")):
                if sv != s and (ta, tb) != ("N", "N"):
                    continue
                oz.gemm(ta, tb, m, n, k, alpha, dA, m if ta == "N" else k, dB, k if tb == "N" else n, beta, dC, m, sv)
                torch.cuda.synchronize()
                ts = []
                for _ in range(3):
                    a = time.perf_counter()
                    oz.gemm(ta, tb, m, n, k, alpha, dA, m if ta == "N" else k, dB, k if tb == "N" else n, beta, dC,
                            m, sv)
                    torch.cuda.synchronize()
                    ts.append(time.perf_counter() - a)
                sweep[f"s{sv}_{ta}{tb}"] = round(flops / min(ts) / 1e12, 2)
        out["sweep_tflops"] = sweep
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, n, k, s)
    oz.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
