"""Accuracy of the emulated DGEMM / ZGEMM against a double-double reference: the
counterpart of the reference's benchmarks/bench_accuracy.cpp.

Same options (--k, --s, --dist, --mn, --cpu, --complex), the same input generator
(liboz2_check.so oz2chk_gen: std::mt19937_64 seeded 11 + k for A and 77 + k for B;
complex 11 + k / 13 + k for Re / Im of A and 77 + k / 79 + k for B), the same methods
and the same CSV columns:

    dist,k,method,s,tk,ulp_max,ulp_mean,cr_frac,ab_max,fro_rel

with method dgemm_gpu (the GPU's native FP64 GEMM: rocBLAS through torch.matmul),
dgemm_cpu (numpy's BLAS, with --cpu), ozaki2 (this library, one row per s whose plan
admits k), ozaki1 (the Ozaki Scheme I baseline of oz2/ozaki1.py, one row per --L, with
s = L(L+1)/2 GEMMs and tk = 7L - 2); complex: zgemm_gpu, zgemm_cpu, ozaki2_3m.
m = n = mn, alpha = 1, beta = 0, op = N.

    python -m oz2.bench_accuracy --k 256,4096 --s 8-24 --dist urand,phi2
"""
import argparse
import sys

import numpy as np
import torch

import oz2
from oz2 import check, ozaki1


def parse_ints(text):
    """'4-8,12' -> [4, 5, 6, 7, 8, 12]"""
    out = []
    for part in text.split(","):
        if not part:
            continue
        if "-" in part[1:]:
            lo, hi = part.split("-", 1)
            out.extend(range(int(lo), int(hi) + 1))
        else:
            out.append(int(part))
    return out


def _row(out, dist, k, method, s, tk, met):
    print(f"{dist},{k},{method},{s},{tk},{met['ulp_max']:.4g},{met['ulp_mean']:.4g},{met['cr_frac']:.6f},"
          f"{met['ab_max']:.4g},{met['fro_rel']:.4g}", file=out, flush=True)


def run_real(oz, dist, k, m, n, svals, cpu, out, Ls=()):
    A = check.gen(dist, m, k, 11 + k, True)
    B = check.gen(dist, k, n, 77 + k, False)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    hi, lo = check.dd_gemm("N", "N", m, n, k, dA, m, dB, k)
    tA = dA.view(k, m).t()  # column-major m x k
    tB = dB.view(n, k).t()
    absab = (tA.abs() @ tB.abs()).t().contiguous().view(-1)

    def metrics(C):
        return check.full_metrics(C, hi, lo, absab)

    _row(out, dist, k, "dgemm_gpu", 0, 53, metrics((tA @ tB).t().contiguous().view(-1)))
    if cpu:
        Ch = A.reshape(k, m).T @ B.reshape(n, k).T
        _row(out, dist, k, "dgemm_cpu", 0, 53, metrics(torch.from_numpy(np.ascontiguousarray(Ch.T).reshape(-1)).cuda()))
    dC = torch.empty(m * n, dtype=torch.float64, device="cuda")
    for s in svals:
        try:
            tk = oz2.Ozaki2.precision_bits(k, s)
        except oz2.OZ2InvalidArgument:
            continue
        oz.gemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        _row(out, dist, k, "ozaki2", s, tk, metrics(dC))
    for L in Ls:
        try:
            ozaki1.gemm(m, n, k, dA, m, dB, k, dC, m, L)
        except ValueError:
            continue
        _row(out, dist, k, "ozaki1", ozaki1.num_gemms(L), ozaki1.precision_bits(L), metrics(dC))


def run_complex(oz, dist, k, m, n, svals, cpu, out):
    Ar, Ai = check.gen(dist, m, k, 11 + k, True), check.gen(dist, m, k, 13 + k, True)
    Br, Bi = check.gen(dist, k, n, 77 + k, False), check.gen(dist, k, n, 79 + k, False)
    A = np.stack([Ar, Ai], axis=1).reshape(-1)  # interleaved (re, im), column-major
    B = np.stack([Br, Bi], axis=1).reshape(-1)
    dA = torch.from_numpy(A).cuda()
    dB = torch.from_numpy(B).cuda()
    hi = torch.empty(2 * m * n, dtype=torch.float64, device="cuda")
    lo = torch.empty_like(hi)
    rc = check.lib().oz2chk_dd_gemm(None, 0, 0, m, n, k, dA.data_ptr(), m, dB.data_ptr(), k, hi.data_ptr(),
                                    lo.data_ptr(), 1)
    if rc != 0:
        raise RuntimeError("oz2: dd reference failed")
    cA = torch.view_as_complex(dA.view(-1, 2)).view(k, m).t()
    cB = torch.view_as_complex(dB.view(-1, 2)).view(n, k).t()
    ab1 = (torch.from_numpy(np.abs(Ar) + np.abs(Ai)).cuda().view(k, m).t()
           @ torch.from_numpy(np.abs(Br) + np.abs(Bi)).cuda().view(n, k).t())
    absab = ab1.t().contiguous().view(-1).repeat_interleave(2)  # the same bound for Re and Im

    def metrics(C):
        return check.full_metrics(C, hi, lo, absab)

    def flat(Cc):
        return torch.view_as_real(Cc.t().contiguous()).reshape(-1)

    _row(out, "z" + dist, k, "zgemm_gpu", 0, 53, metrics(flat(cA @ cB)))
    if cpu:
        Ah = (Ar + 1j * Ai).reshape(k, m).T
        Bh = (Br + 1j * Bi).reshape(n, k).T
        _row(out, "z" + dist, k, "zgemm_cpu", 0, 53, metrics(flat(torch.from_numpy(Ah @ Bh).cuda())))
    dC = torch.empty(2 * m * n, dtype=torch.float64, device="cuda")
    for s in svals:
        try:
            tk = oz2.Ozaki2.precision_bits_complex(k, s)
        except oz2.OZ2InvalidArgument:
            continue
        oz.zgemm("N", "N", m, n, k, 1.0, dA, m, dB, k, 0.0, dC, m, s)
        _row(out, "z" + dist, k, "ozaki2_3m", s, tk, metrics(dC))


def main(argv=None, out=sys.stdout):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--k", default="256,1024,4096,16384")
    ap.add_argument("--s", default="4-24")
    ap.add_argument("--dist", default="urand,phi0.5,phi2,phi4")
    ap.add_argument("--mn", type=int, default=512)
    ap.add_argument("--cpu", action="store_true")
    ap.add_argument("--complex", action="store_true")
    ap.add_argument("--L", default="4-9", help="Ozaki Scheme I slice counts (real runs; empty: none)")
    args = ap.parse_args(argv)
    ks, svals, dists = parse_ints(args.k), parse_ints(args.s), [d for d in args.dist.split(",") if d]
    Ls = parse_ints(args.L)
    oz = oz2.Ozaki2()
    print("dist,k,method,s,tk,ulp_max,ulp_mean,cr_frac,ab_max,fro_rel", file=out, flush=True)
    try:
        for dist in dists:
            for k in ks:
                if args.complex:
                    run_complex(oz, dist, k, args.mn, args.mn, svals, args.cpu, out)
                else:
                    run_real(oz, dist, k, args.mn, args.mn, svals, args.cpu, out, Ls)
    finally:
        oz.close()


if __name__ == "__main__":
    main()
