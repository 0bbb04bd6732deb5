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
HBM_PEAK_GBS = 8000.0    # MI355X HBM3E, 8 TB/s


# HBM bytes per INT8 GEMM launch measured with rocprofv3 (FETCH_SIZE and WRITE_SIZE passes of this
# build's bench.py workloads on one MI355X, each scaled by its reading for a 1 GiB copy kernel).
# A bench run cannot count its own traffic (the counters need rocprofv3 around the process), so the
# line quotes these with their source; the GEMM's fetch (its L2 misses) differs by a few % between boxes.
MEASURED_GEMM_TRAFFIC = {
    ("dgemm", 8192, 8192, 8192, 16): (1.4405e10, "profiles/dgemm8192_s16_hbm_traffic.txt (one MI355X, 2026-12-03)"),
    ("zgemm", 4096, 4096, 4096, 16): (5.6388e9, "profiles/zgemm4096_s16_hbm_traffic.txt (one MI355X, 2026-12-03)"),
}


def gemm_traffic(kind, m, n, k, s):
    t = MEASURED_GEMM_TRAFFIC.get((kind, m, n, k, s))
    return (t[0], t[1]) if t else (None, None)


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
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-native", action="store_true",
                    help="skip the native FP64 GEMM (rocBLAS) timing and accuracy on the same inputs")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run three 1 GiB device-to-device copies (the runtime's copy kernel): the "
                         "known-size reference for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    ap.add_argument("--vendor-int8", action="store_true",
                    help="also time the vendor INT8 GEMM (hipBLASLt through torch._int_mm) on s GEMMs of the same "
                         "shape, int32 output, as the rate this GPU reaches with a library kernel")
    ap.add_argument("--vendor-fp8", action="store_true",
                    help="also time the vendor FP8 GEMM (hipBLASLt through torch._scaled_mm) on s GEMMs of the same "
                         "shape: the same nominal peak as INT8, for comparing sustained rates and clocks")
    ap.add_argument("--mfma-rate", action="store_true",
                    help="also time a register-only INT8 MFMA loop on every CU: the rate the matrix cores "
                         "sustain with no memory traffic, the practical ceiling for the GEMM")
    ap.add_argument("--workload", default="dgemm8192",
                    choices=["dgemm8192", "zgemm4096", "wide4000", "ksplit", "batch"],
                    help="dgemm8192 (default line), or one of the other measured configurations")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--overlap", type=int, default=-1,
                    help="column blocks overlapped over two streams (Tuning::overlap_blocks): -1 automatic, 1 none")
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
    from oz2 import check
    tk = oz2_precision_bits(k, s)
    return {"ulp_max": float(eu.max()), "ulp_mean": float(eu.mean()), "cr_frac": float(np.mean(C == H + L)),
            "bound_ratio": check.bound_ratio(m, n, k, dA, m, dB, k, dC, m, hi, lo, tk),
            "bound": f"largest error / the reference's rigorous element-wise bound at s={s} (tk={tk}); <= 1 "
                     "means every entry is within it",
            "reference": "double-double GEMM on the GPU, same inputs, alpha=1 beta=0"}


def oz2_precision_bits(k, s):
    import oz2
    return oz2.Ozaki2.precision_bits(k, s)


def make_batch(torch, count, seed):
    from oz2.workloads import uneven_batch
    return uneven_batch(torch, count, seed)


def batch_bench(oz, torch, count, reps=5, bits=53):
    """learned-s grouped batch (hipGraph) vs the plain sequential fixed-s loop on the uneven mix"""
    import oz2
    from oz2 import check
    probs = make_batch(torch, count, seed=5)
    flops = sum(2.0 * p["m"] * p["n"] * p["k"] for p in probs)
    refs = [check.dd_gemm("N", "N", p["m"], p["n"], p["k"], p["A"], p["lda"], p["B"], p["ldb"]) for p in probs]
    torch.cuda.synchronize()

    def accuracy(svals):
        """ulp_max, ulp_mean and the correctly rounded fraction over every entry of every problem, and
        the largest ratio of an error to the reference's rigorous bound at that problem's s"""
        emax, esum, ncr, cnt, br = 0.0, 0.0, 0.0, 0, 0.0
        for p, (hi, lo), sv in zip(probs, refs, svals):
            mt = check.metrics(p["C"], hi, lo, p["m"], p["ldc"])
            ne = hi.numel()
            emax = max(emax, mt["ulp_max"])
            esum += mt["ulp_mean"] * ne
            ncr += mt["cr_frac"] * ne
            cnt += ne
            tk = oz2.Ozaki2.precision_bits(p["k"], sv)
            br = max(br, check.bound_ratio(p["m"], p["n"], p["k"], p["A"], p["lda"], p["B"], p["ldb"], p["C"],
                                           p["ldc"], hi, lo, tk))
        return {"ulp_max": emax, "ulp_mean": esum / max(cnt, 1), "cr_frac": ncr / max(cnt, 1), "bound_ratio": br}

    def max_ulp():
        return max(float(check.ulp_errors(p["C"], hi, lo, p["m"], p["ldc"]).max()) for p, (hi, lo) in zip(probs, refs))

    def plain(svals):
        for p, sv in zip(probs, svals):
            oz.gemm("N", "N", p["m"], p["n"], p["k"], 1.0, p["A"], p["lda"], p["B"], p["ldb"], 0.0, p["C"], p["ldc"], sv)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    oz.set_profiling(False)
    s_mfb = [oz2.Ozaki2.moduli_for_bits(p["k"], bits) for p in probs]
    t_plain = timed(lambda: plain(s_mfb))
    acc_plain = accuracy(s_mfb)
    ulp_plain = acc_plain["ulp_max"]
    descs = [dict(p, num_moduli=0) for p in probs]
    # the selector's own cost (exponent features of every problem + the network, two stream
    # synchronizations): paid when a plan is made or re-selected, not by its replays
    oz.selector_choose(descs, bits=bits)
    torch.cuda.synchronize()
    t_sel = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        oz.selector_choose(descs, bits=bits)
        t_sel = min(t_sel, time.perf_counter() - t0)
    tune = oz.batch_autotune(descs, bits=bits, select=True)
    b = oz2.Batch(oz, descs, bits=bits, graph=True, select=True, tiles=tune["tiles"],
                  persistent=tune["persistent"], largest_first=tune["largest_first"])
    t_batch = timed(b.run)
    acc_batch = accuracy(b.moduli())
    ulp_batch = acc_batch["ulp_max"]
    # re-selection of the plan's moduli from the buffers' current contents (oz2_batch_reselect:
    # features + network + choice; the contents are unchanged, so no plan is rebuilt)
    b.reselect()
    t_resel = float("inf")
    for _ in range(3):
        t0 = time.perf_counter()
        changed = b.reselect()
        t_resel = min(t_resel, time.perf_counter() - t0)
    assert changed == 0
    s_sel = b.moduli()
    info = b.info()
    # plain loop at the smallest uniform s whose max-ulp is no worse than the batch's
    s_eq, t_eq, ulp_eq = None, None, None
    for sv in range(min(s_mfb), 25):
        plain([sv] * len(probs))
        torch.cuda.synchronize()
        u = max_ulp()
        if u <= ulp_batch:
            s_eq, ulp_eq = sv, u
            acc_eq = accuracy([sv] * len(probs))
            t_eq = timed(lambda: plain([sv] * len(probs)))
            break
    b.close()
    # the same batch with moduli_for_bits per problem: grouping and graph alone
    b2 = oz2.Batch(oz, [dict(p, num_moduli=sv) for p, sv in zip(probs, s_mfb)], graph=True, tiles=tune["tiles"],
                   persistent=tune["persistent"], largest_first=tune["largest_first"])
    t_group = timed(b2.run)
    b2.close()
    # the chosen plan with the other GEMM schedule (persistent workgroups pulling tiles from a
    # counter for the 128- and 64-row groups; 256-wide groups keep one workgroup per tile)
    b3 = oz2.Batch(oz, descs, bits=bits, graph=True, select=True, tiles=tune["tiles"],
                   persistent=not tune["persistent"], largest_first=tune["largest_first"])
    t_other = timed(b3.run)
    b3.close()
    oz.set_profiling(True)
    out = {
        "workload": f"{count} DGEMMs N/N, m n k log-uniform [64,4096], half U[-1,1) half (1+U)*2^e e in [-12,12], "
                    f"alpha=1 beta=0, {bits}-bit target",
        "fp64_flops": flops,
        "plain_fixed_s": {"tflops": round(flops / t_plain / 1e12, 3), "ms": round(t_plain * 1e3, 3),
                          "s": "moduli_for_bits(k, 53) per problem", "max_ulp": ulp_plain,
                          "ulp_mean": acc_plain["ulp_mean"], "cr_frac": acc_plain["cr_frac"],
                          "bound_ratio": acc_plain["bound_ratio"]},
        "grouped_graph_same_s": {"tflops": round(flops / t_group / 1e12, 3), "ms": round(t_group * 1e3, 3)},
        "learned_grouped_graph": {"tflops": round(flops / t_batch / 1e12, 3), "ms": round(t_batch * 1e3, 3),
                                  "max_ulp": ulp_batch, "ulp_mean": acc_batch["ulp_mean"],
                                  "cr_frac": acc_batch["cr_frac"], "bound_ratio": acc_batch["bound_ratio"],
                                  "mean_s": sum(s_sel) / len(s_sel),
                                  "s_hist": {str(v): s_sel.count(v) for v in sorted(set(s_sel))},
                                  "launches_per_run": info["launches"], "workspace_gib": round(info["workspace_bytes"] / 2 ** 30, 2),
                                  "autotuned": {"persistent": tune["persistent"], "largest_first": tune["largest_first"],
                                                "tiles": {("256x128" if t == 256128 else str(t)): tune["tiles"].count(t)
                                                          for t in (64, 128, 256, 256128)},
                                                "tune_ms_no_graph": round(tune["ms"], 3)},
                                  "other_schedule": {"persistent": not tune["persistent"],
                                                     "ms": round(t_other * 1e3, 3)}},
        "speedup_vs_plain": round(t_plain / t_batch, 3),
        "selector": {"ms": round(t_sel * 1e3, 3),
                     "what": "features of all problems + MLP + choice through the Python API (oz2_selector_choose, "
                             "problem list built from tensors), host wall time",
                     "reselect_ms": round(t_resel * 1e3, 3),
                     "reselect_what": "oz2_batch_reselect on the plan (features + MLP + choice from the buffers' "
                                      "current contents, nothing rebuilt), host wall time; paid once per plan or "
                                      "re-selection, not per replay",
                     "learned_plus_reselect_ms": round((t_batch + t_resel) * 1e3, 3)},
    }
    if s_eq is not None:
        out["plain_equal_max_ulp"] = {"s": s_eq, "max_ulp": ulp_eq, "ulp_mean": acc_eq["ulp_mean"],
                                      "cr_frac": acc_eq["cr_frac"], "bound_ratio": acc_eq["bound_ratio"],
                                      "tflops": round(flops / t_eq / 1e12, 3),
                                      "ms": round(t_eq * 1e3, 3)}
        out["speedup_at_equal_max_ulp"] = round(t_eq / t_batch, 3)
        out["speedup_at_equal_max_ulp_selection_every_run"] = round(t_eq / (t_batch + t_resel), 3)
    return out


def hbm_stages(m, n, k, s, cplx, beta_nonzero, stage):
    """algorithmic HBM bytes of the bandwidth-bound stages and their achieved GB/s from the stage times:
    scale reads the operands; the split reads them and writes s (3s complex) int8 planes padded to the
    256-row GEMM tile and k to 128; the CRT reads s (3s) residue bytes per entry and writes C (reading it
    too when beta != 0)"""
    def ru(x, a):
        return (x + a - 1) // a * a
    w, g = (2, 3) if cplx else (1, 1)
    kp = ru(k, 128)
    elems = m * k + k * n
    nbytes = {"scale": 8 * w * elems,
              "split": 8 * w * elems + s * g * (ru(m, 256) + ru(n, 256)) * kp,
              "crt": s * g * m * n + 8 * w * m * n * (2 if beta_nonzero else 1)}
    out = {}
    for name, b in nbytes.items():
        ms = stage.get(f"{name}_ms", 0.0)
        gbs = b / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[name] = {"bytes": b, "ms": round(ms, 4), "GB/s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["peak_GB/s"] = HBM_PEAK_GBS
    return out


def zgemm_line(args, torch):
    """ZGEMM 4096^3, op(A)=C, op(B)=N, s=16, alpha=(1,0.5), beta=(0.25,-1), lda=m+7, ldc=m+3 (3M form)"""
    import oz2
    from oz2 import check
    m = n = k = 4096
    s = 16
    lda, ldb, ldc = k + 7, k, m + 3  # op(A) = A^H: A is k x m
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    dA = torch.rand(2 * lda * m, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dB = torch.rand(2 * ldb * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dC0 = torch.rand(2 * ldc * n, dtype=torch.float64, device="cuda", generator=g) * 2 - 1
    dC = dC0.clone()
    oz = oz2.Ozaki2()
    al, be = complex(1, 0.5), complex(0.25, -1)
    for _ in range(args.warmup):
        oz.zgemm("C", "N", m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, s)
    torch.cuda.synchronize()
    oz.profile_begin()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        oz.zgemm("C", "N", m, n, k, al, dA, lda, dB, ldb, be, dC, ldc, s)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    st, calls = oz.profile_end()
    gemm_ms = st.gemm_ms / max(1, calls)
    int8_ops = 2.0 * m * n * k * 3 * s
    achieved = int8_ops / (gemm_ms * 1e-3) / 1e12
    out = {"metric": "FP64-equiv TFLOP/s at stated max-ulp: ZGEMM 4096^3 s=16 (8mnk per call)",
           "value": round(8.0 * m * n * k / (ms * 1e-3) / 1e12, 3), "unit": "TFLOP/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int8 (3M: 3s INT8 GEMMs; f64 in/out)",
           "data": "synthetic U[-1,1) re and im, HBM-resident",
           "config": {"workload": "ZGEMM 4096x4096x4096 op(A)=C op(B)=N s=16 alpha=(1,0.5) beta=(0.25,-1) "
                                  "lda=k+7 ldc=m+3", "tk": oz2.Ozaki2.precision_bits_complex(k, s),
                      "tiling": oz.last_tiling().as_dict()},
           "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                        "frac": round(achieved / INT8_PEAK_TOPS, 4), "traffic": gemm_traffic("zgemm", m, n, k, s)[0],
                        "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                        "traffic_source": gemm_traffic("zgemm", m, n, k, s)[1],
                        "algorithmic_bytes": 3 * s * (k * (m + n) + m * n),
                        "kernel": "igemm_residue_kernel (2*m*n*k*3s int8 ops per launch)", "gemm_ms": round(gemm_ms, 4)},
           "stage_ms": {kk: round(v / max(1, calls), 4) for kk, v in st.as_dict().items()}}
    out["hbm_stages"] = hbm_stages(m, n, k, s, True, True, out["stage_ms"])
    if args.no_accuracy:
        oz.close()
        return out
    # accuracy with alpha=1, beta=0 against double-double (complex)
    dC1 = torch.empty(2 * ldc * n, dtype=torch.float64, device="cuda")
    oz.zgemm("C", "N", m, n, k, 1.0, dA, lda, dB, ldb, 0.0, dC1, ldc, s)
    hi = torch.empty(2 * m * n, dtype=torch.float64, device="cuda")
    lo = torch.empty_like(hi)
    rc = check.lib().oz2chk_dd_gemm(None, 2, 0, m, n, k, dA.data_ptr(), lda, dB.data_ptr(), ldb, hi.data_ptr(),
                                    lo.data_ptr(), 1)
    assert rc == 0
    torch.cuda.synchronize()
    Cv = dC1.view(n, ldc, 2)[:, :m, :].reshape(-1)
    e = check.ulp_errors(Cv, hi, lo, 2 * m * n, 2 * m * n)
    tk = oz2.Ozaki2.precision_bits(2 * k, s)
    out["accuracy"] = {"ulp_max": float(e.max()), "ulp_mean": float(e.mean()),
                       "cr_frac": float((Cv == hi + lo).double().mean()),
                       "bound_ratio": check.zbound_ratio("C", "N", m, n, k, dA, lda, dB, ldb, dC1, ldc, hi, lo, tk),
                       "bound": f"largest error / the reference's rigorous element-wise bound at s={s} (tk={tk}); "
                                "<= 1 means every entry is within it",
                       "reference": "double-double ZGEMM on the GPU, per real component, alpha=1 beta=0"}
    if not args.no_native:
        out["native_fp64"] = native_fp64(torch, m, n, k, dA, dB, cplx=True, op_a="C", lda=lda)
        out["x_native_fp64"] = round(out["value"] / out["native_fp64"]["tflops"], 3)
    oz.close()
    return out


def selector_line(args, torch, which):
    """wide-exponent DGEMM (4000x8000x4000) or k-split 2048x2048x200000, entries (1+U)*(+-1)*2^e, e in [-20,20]:
    learned s vs moduli_for_bits(k, 53), each with its max-ulp against double-double and its largest ratio to
    the reference's rigorous element-wise bound at its s"""
    import oz2
    from oz2 import check
    m, k, n = (4000, 8000, 4000) if which == "wide4000" else (2048, 200000, 2048)
    g = torch.Generator(device="cuda")
    g.manual_seed(11)

    def gen(cnt):
        e = torch.randint(-20, 21, (cnt,), device="cuda", generator=g).double()
        sg = torch.randint(0, 2, (cnt,), device="cuda", generator=g).double() * 2 - 1
        return (1 + torch.rand(cnt, dtype=torch.float64, device="cuda", generator=g)) * sg * torch.exp2(e)
    dA, dB = gen(m * k), gen(k * n)
    dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
    oz = oz2.Ozaki2()
    hi, lo = check.dd_gemm("N", "N", m, n, k, dA, m, dB, k)
    desc = dict(ta="N", tb="N", m=m, n=n, k=k, A=dA, lda=m, B=dB, ldb=k)
    s_sel, probs, learned = oz.selector_choose([desc], bits=53)
    res = {}
    for name, sv in (("moduli_for_bits", oz2.Ozaki2.moduli_for_bits(k, 53)), ("learned", s_sel[0])):
        oz.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, sv)
        torch.cuda.synchronize()
        ts = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            oz.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, sv)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        met = check.metrics(dC, hi, lo, m, m)
        met["bound_ratio"] = check.bound_ratio(m, n, k, dA, m, dB, k, dC, m, hi, lo,
                                               oz2.Ozaki2.precision_bits(k, sv))
        res[name] = {"s": sv, "ms": round(min(ts) * 1e3, 3), "tflops": round(2.0 * m * n * k / min(ts) / 1e12, 3),
                     **met, "tiling": oz.last_tiling().as_dict()}
    res["learned"]["selector_confident"] = bool(learned[0])
    oz.close()
    return {"metric": f"FP64-equiv TFLOP/s at stated max-ulp: DGEMM {m}x{n}x{k} wide exponents",
            "value": res["learned"]["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8 (INT8 MFMA; f64 in/out)",
            "data": "synthetic (1+U[0,1))*(+-1)*2^e, e uniform integer in [-20,20], HBM-resident",
            "config": {"workload": f"DGEMM m={m} n={n} k={k} N/N alpha=1 beta=0, 53-bit target"}, "runs": res}


def _host_cpu():
    """CPU model name (lscpu) and the logical CPUs of the host"""
    import subprocess
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        model = next((ln.split(":", 1)[1].strip() for ln in out.splitlines() if ln.startswith("Model name")), "")
    except (OSError, subprocess.SubprocessError):
        pass
    return model, os.cpu_count()


def cpu_baseline(m, n, k, s):
    """The CPU-runnable configuration (DGEMM 1024^3 N/N, alpha=1, beta=0, U[-1,1), s=14) through the
    oracle, a C/OpenMP restatement of the reference's algorithm (its SYCL CPU device cannot run here):
    1 warm-up, then best and median wall time of 5 calls, checked against a double-double product with
    the reference's rigorous bound.  Also one call on rows 0..3071 of the GPU line's DGEMM."""
    import numpy as np
    from oracle import oracle as O
    threads = len(os.sched_getaffinity(0))
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    model, ncpu = _host_cpu()
    N, s1 = 1024, 14
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, N * N)
    B = rng.uniform(-1, 1, N * N)
    C = np.zeros(N * N)
    O.dgemm("N", "N", N, N, N, 1.0, A, N, B, N, 0.0, C, N, s1, threads)  # warm-up
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        O.dgemm("N", "N", N, N, N, 1.0, A, N, B, N, 0.0, C, N, s1, threads)
        ts.append(time.perf_counter() - t0)
    hi, lo = O.dd_gemm("N", "N", N, N, N, A, N, B, N, threads)
    bound = O.bound_ratio("N", "N", N, N, N, A, N, B, N, 1.0, 0.0, np.zeros(N * N), C, N, hi, lo, s1)
    acc = O.metrics(N, N, C, N, hi, lo)
    ts.sort()
    flops = 2.0 * N ** 3
    out = {"value": round(flops / ts[0] / 1e12, 5), "unit": "FP64-equiv TFLOP/s", "cores": threads, "kind": "port",
           "label": "CPU restatement (reference SYCL CPU device unavailable)",
           "sample": f"DGEMM {N}x{N}x{N} N/N alpha=1 beta=0 U[-1,1) seed 1, s={s1}: oracle residue route "
                     f"(split, exact int8 GEMM, CRT) with OpenMP on {threads} threads; 1 warm-up, best of 5",
           "best_s": round(ts[0], 4), "median_s": round(ts[2], 4),
           "median_tflops": round(flops / ts[2] / 1e12, 5),
           "nproc": ncpu, "cpu_model": model,
           "accuracy_vs_double_double": {"bound_ratio": bound, **acc}}
    # the GPU line's own workload, on its first 3072 rows (about 10-15 s on 16 host cores)
    rows = 3072
    rng = np.random.default_rng(1)
    A = rng.uniform(-1, 1, rows * k)
    B = rng.uniform(-1, 1, k * n)
    C = rng.uniform(-1, 1, rows * n)
    t0 = time.perf_counter()
    O.dgemm("N", "N", rows, n, k, 1.5, A, rows, B, k, 0.5, C, rows, s, threads)
    dt = time.perf_counter() - t0
    out["gpu_workload_sample"] = {"tflops": round(2.0 * rows * n * k / dt / 1e12, 5), "seconds": round(dt, 3),
                                  "sample": f"rows 0..{rows - 1} of the {m}x{n}x{k} s={s} DGEMM (alpha=1.5, "
                                            f"beta=0.5): {rows}x{n}x{k}, one call"}
    return out


def native_fp64(torch, m, n, k, dA, dB, cplx=False, op_a="N", lda=None, reps=5):
    """The GPU's native FP64 GEMM (rocBLAS through torch.matmul) on the same column-major inputs:
    best time of reps, TFLOP/s (2mnk, complex 8mnk) and max-ulp against double-double (alpha=1, beta=0)."""
    from oz2 import check
    if not cplx:
        A = dA.view(k, m).t()  # column-major m x k
        B = dB.view(n, k).t()
    else:
        ld = lda or (k if op_a != "N" else m)
        Ac = torch.view_as_complex(dA.view(-1, 2))
        A = Ac.view(m, ld)[:, :k].conj() if op_a == "C" else Ac.view(k, ld)[:, :m].t()  # op(A), m x k
        B = torch.view_as_complex(dB.view(-1, 2)).view(n, k).t()
    C = A @ B
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        C = A @ B
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    flops = (8.0 if cplx else 2.0) * m * n * k
    out = {"library": "rocBLAS through torch.matmul (" + ("complex128" if cplx else "float64") + ")",
           "ms": round(best, 4), "tflops": round(flops / best / 1e9, 2)}
    hi = torch.empty((2 if cplx else 1) * m * n, dtype=torch.float64, device="cuda")
    lo = torch.empty_like(hi)
    if not cplx:
        rc = check.lib().oz2chk_dd_gemm(None, 0, 0, m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, hi.data_ptr(),
                                        lo.data_ptr(), 0)
        Cv = C.t().contiguous().view(-1)
    else:
        rc = check.lib().oz2chk_dd_gemm(None, 2 if op_a == "C" else 0, 0, m, n, k, dA.data_ptr(), lda or m,
                                        dB.data_ptr(), k, hi.data_ptr(), lo.data_ptr(), 1)
        Cv = torch.view_as_real(C.t().contiguous()).reshape(-1)
    assert rc == 0
    torch.cuda.synchronize()
    e = check.ulp_errors(Cv, hi, lo, hi.numel(), hi.numel())
    out["ulp_max"] = float(e.max())
    out["cr_frac"] = float((Cv == hi + lo).double().mean())
    return out


def vendor_int8(torch, m, n, k, s, reps=5):
    """hipBLASLt int8 x int8 -> int32 (torch._int_mm) on s GEMMs of m x n x k, best of reps; op(B) = T and N."""
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    A = [torch.randint(-128, 128, (m, k), dtype=torch.int8, device="cuda", generator=g) for _ in range(2)]
    out = {"library": "hipBLASLt (torch._int_mm), int32 output", "gemms": s}
    for layout in ("NT", "NN"):
        B = [torch.randint(-128, 128, (n, k) if layout == "NT" else (k, n), dtype=torch.int8, device="cuda",
                           generator=g) for _ in range(2)]
        B = [b.t() for b in B] if layout == "NT" else B
        torch._int_mm(A[0], B[0])
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            for i in range(s):
                torch._int_mm(A[i & 1], B[i & 1])
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1))
        out[f"{layout}_ms"] = round(best, 4)
        out[f"{layout}_tops"] = round(2.0 * m * n * k * s / best / 1e9, 1)
        del B
    return out


def vendor_fp8(torch, m, n, k, s, reps=5):
    """hipBLASLt float8_e4m3fn x float8_e4m3fn -> bf16 (torch._scaled_mm, unit scales) on s GEMMs of m x n x k,
    op(B) = T, best of reps: the same nominal MFMA peak as INT8 (5 POP/s dense), for comparing the rate this GPU
    sustains with each element type."""
    out = {"library": "hipBLASLt (torch._scaled_mm) float8_e4m3fn, bf16 output", "gemms": s}
    try:
        g = torch.Generator(device="cuda")
        g.manual_seed(4)
        A = [(torch.randn(m, k, device="cuda", generator=g) * 0.1).to(torch.float8_e4m3fn) for _ in range(2)]
        B = [(torch.randn(n, k, device="cuda", generator=g) * 0.1).to(torch.float8_e4m3fn).t() for _ in range(2)]
        one = torch.tensor(1.0, device="cuda")

        def run(i):
            torch._scaled_mm(A[i & 1], B[i & 1], scale_a=one, scale_b=one, out_dtype=torch.bfloat16)

        run(0)
        torch.cuda.synchronize()
    except (RuntimeError, AttributeError) as e:
        out["error"] = str(e)[:200]
        return out
    best = float("inf")
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(s):
            run(i)
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    out["NT_ms"] = round(best, 4)
    out["NT_tflops"] = round(2.0 * m * n * k * s / best / 1e9, 1)
    return out


def mfma_rate(torch, reps=3):
    """sustained INT8 MFMA rate with no memory traffic (liboz2_check.so oz2chk_mfma_rate: 16 independent
    32x32x32 chains per wave, one wave per SIMD, 4 workgroups per CU over the launch), best of reps"""
    import ctypes
    from oz2 import check
    L = check.lib()
    L.oz2chk_mfma_rate.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    blocks, iters = 4 * torch.cuda.get_device_properties(0).multi_processor_count, 4096
    out = torch.empty(blocks * 256, dtype=torch.int32, device="cuda")
    best = float("inf")
    for _ in range(reps + 1):  # the first launch warms the clock up
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        if L.oz2chk_mfma_rate(None, blocks, iters, out.data_ptr()) != 0:
            raise RuntimeError("oz2: MFMA rate kernel failed")
        e1.record()
        torch.cuda.synchronize()
        best = min(best, e0.elapsed_time(e1))
    ops = 2.0 * 32768 * 16 * iters * 4 * blocks
    return {"kernel": "register-only v_mfma_i32_32x32x32_i8 loop", "ms": round(best, 4),
            "tops": round(ops / best / 1e9, 1), "frac_of_peak": round(ops / best / 1e9 / INT8_PEAK_TOPS, 4)}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import oz2

    if args.calibrate_copy:
        nbytes = 1 << 30
        src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record()
        for i in range(3):
            dst.copy_(src)
            ev[i + 1].record()
        torch.cuda.synchronize()
        best = min(ev[i].elapsed_time(ev[i + 1]) for i in range(3))
        print(json.dumps({"calibration": "device-to-device copy", "bytes_read": nbytes, "bytes_written": nbytes,
                          "launches": 3, "best_ms": round(best, 4),
                          "GB/s": round(2 * nbytes / best / 1e6, 1)}))
        return
    if args.workload == "zgemm4096":
        print(json.dumps(zgemm_line(args, torch)))
        return
    if args.workload in ("wide4000", "ksplit"):
        print(json.dumps(selector_line(args, torch, args.workload)))
        return
    if args.workload == "batch":  # the learned-s grouped batch alone (for profiling)
        oz = oz2.Ozaki2()
        probs = make_batch(torch, args.batch_count, seed=5)
        descs = [dict(p, num_moduli=0) for p in probs]
        tune = oz.batch_autotune(descs, bits=53, select=True)
        b = oz2.Batch(oz, descs, bits=53, graph=True, select=True, tiles=tune["tiles"], persistent=tune["persistent"],
                      largest_first=tune["largest_first"])
        for _ in range(args.warmup):
            b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.run()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        flops = sum(2.0 * p["m"] * p["n"] * p["k"] for p in probs)
        print(json.dumps({"metric": "FP64-equiv TFLOP/s: uneven batch, learned s, grouped + hipGraph",
                          "value": round(flops / ms / 1e9, 3), "unit": "TFLOP/s", "ms_per_step": round(ms, 3),
                          "steps": args.steps, "moduli": b.moduli(), "tiles": b.tiles(), "tune": tune, **b.info()}))
        b.close()
        oz.close()
        return
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
    oz.set_tuning(overlap_blocks=args.overlap)

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
        "metric": "FP64-equiv TFLOP/s at stated max-ulp: DGEMM 8192^3 s=16, uneven batch mix",
        "value": round(value, 3), "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int8 (INT8 MFMA, int32 accumulate; f64 in/out)", "data": "synthetic U[-1,1) seeded, HBM-resident",
        "config": {"workload": f"DGEMM {m}x{n}x{k} op N/N s={s} alpha=1.5 beta=0.5, A B C0 U[-1,1)", "m": m, "n": n,
                   "k": k, "s": s, "tk": oz2.Ozaki2.precision_bits(k, s), "tiling": tiling},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / INT8_PEAK_TOPS, 4), "traffic": gemm_traffic("dgemm", m, n, k, s)[0],
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "traffic_source": gemm_traffic("dgemm", m, n, k, s)[1],
                     "algorithmic_bytes": s * k * (m + n) + s * m * n,
                     "kernel": "igemm_residue_kernel (2*m*n*k*s int8 ops per launch)",
                     "gemm_ms": round(gemm_ms, 4)},
        "stage_ms": {kk: round(v, 4) for kk, v in stage.items()},
        "hbm_stages": hbm_stages(m, n, k, s, False, True, stage),
        "int8_rate_end_to_end": round(int8_ops / (ms * 1e-3) / 1e12, 1),
        "reference_b60_tflops": 9.3,
    }
    if not args.no_accuracy:
        out["accuracy"] = accuracy(oz, torch, dA, dB, m, n, k, s)
    if not args.no_native:
        out["native_fp64"] = native_fp64(torch, m, n, k, dA, dB)
        out["x_native_fp64"] = round(value / out["native_fp64"]["tflops"], 3)
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
    if args.vendor_int8:
        out["roofline"]["vendor_int8"] = vendor_int8(torch, m, n, k, s)
    if args.vendor_fp8:
        out["roofline"]["vendor_fp8"] = vendor_fp8(torch, m, n, k, s)
    if args.mfma_rate:
        out["roofline"]["mfma_register_loop"] = mfma_rate(torch)
    if not args.no_batch:
        del dA, dB, dC, dC0
        torch.cuda.empty_cache()
        oz.release_workspace()
        out["batch"] = batch_bench(oz, torch, args.batch_count)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, n, k, s)
    oz.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
