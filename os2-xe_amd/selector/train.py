"""Train the learned moduli selector on data the method itself generates.

  python os2-xe_amd/selector/train.py --out DIR [--problems N] [--seed S]

1. Data: random DGEMMs (shapes log-uniform, operands from a mix of value
   distributions, from narrow U[-1,1) to wide exponent ranges).  For each
   problem the library computes its 48 selector features (selector_kernels.hip),
   runs Ozaki II at every admissible s = 2..24 and measures the largest error in
   ulp against a double-double product on the GPU (oz2.check).  A bit target b
   is met when that error is <= 2^(53-b) ulp; labels y[s][b] are those
   outcomes for b in TARGETS.
2. Training: the 48-64-64-23 MLP is trained with the library's own HIP
   kernels (oz2_mlp_forward / oz2_mlp_backward / oz2_mlp_adam) on masked binary
   cross-entropy; torch tensors only hold device memory.
3. Threshold: the smallest probability threshold whose choices (the library's
   rule: smallest admissible s with p_s >= threshold, else the largest
   admissible s) meet the target on at least 99.5 % of held-out examples,
   searched in [0.5, 0.999].
4. Output in DIR: dataset.npz, selector_weights.inc (the form compiled into
   liboz2.so: copy it to os2-xe_amd/csrc/) and report.json.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import torch  # noqa: E402  (before oz2: liboz2 binds to torch's HIP runtime)

import oz2  # noqa: E402
from oz2 import check  # noqa: E402

TARGETS = (24, 32, 40, 48, 53)
SMIN, SMAX = 2, 24


def log(msg):
    print(msg, flush=True)


def sample_values(rng, dist, rows, cols):
    """host float64 (rows, cols) from one of the training distributions"""
    if dist == "u11":
        return rng.uniform(-1, 1, (rows, cols))
    if dist.startswith("wide"):
        emax = int(dist[4:])
        e = rng.integers(-emax, emax + 1, (rows, cols))
        return (1 + rng.random((rows, cols))) * rng.choice([-1.0, 1.0], (rows, cols)) * np.exp2(e)
    if dist.startswith("phi"):
        phi = float(dist[3:])
        return rng.uniform(-0.5, 0.5, (rows, cols)) * np.exp(phi * rng.standard_normal((rows, cols)))
    if dist == "rowscale":  # each row / column its own scale, narrow inside
        return rng.uniform(-1, 1, (rows, cols)) * np.exp2(rng.integers(-30, 31, (rows, 1)))
    if dist == "sparse":
        v = rng.uniform(-1, 1, (rows, cols))
        v[rng.random((rows, cols)) < 0.5] = 0
        return v
    if dist == "ints":
        return rng.integers(-1000, 1001, (rows, cols)).astype(np.float64)
    if dist == "spike":  # narrow values plus a few large entries
        v = rng.uniform(-1, 1, (rows, cols))
        mask = rng.random((rows, cols)) < 0.01
        v[mask] *= np.exp2(rng.integers(10, 25, mask.sum()))
        return v
    raise ValueError(dist)


DISTS = ["u11", "u11", "wide4", "wide8", "wide12", "wide20", "phi0.5", "phi1", "phi2", "phi4", "rowscale", "sparse",
         "ints", "spike"]


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


def admissible(k):
    out = []
    for s in range(SMIN, SMAX + 1):
        try:
            oz2.Ozaki2.precision_bits(k, s)
        except oz2.OZ2InvalidArgument:
            continue
        out.append(s)
    return out


def make_problem(rng):
    m, n = log_uniform_int(rng, 16, 2048), log_uniform_int(rng, 16, 2048)
    k = log_uniform_int(rng, 16, 16384)
    ta, tb = rng.choice(["N", "T"]), rng.choice(["N", "T"])
    da, db = rng.choice(DISTS), rng.choice(DISTS)
    if rng.random() < 0.6:
        db = da
    return dict(m=m, n=n, k=k, ta=str(ta), tb=str(tb), da=str(da), db=str(db))


def run_problem(oz, pr, rng):
    m, n, k, ta, tb = pr["m"], pr["n"], pr["k"], pr["ta"], pr["tb"]
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    A = torch.from_numpy(np.asfortranarray(sample_values(rng, pr["da"], ar, ac)).reshape(-1, order="F")).cuda()
    B = torch.from_numpy(np.asfortranarray(sample_values(rng, pr["db"], br, bc)).reshape(-1, order="F")).cuda()
    desc = dict(ta=ta, tb=tb, m=m, n=n, k=k, A=A, lda=ar, B=B, ldb=br)
    feat = oz.selector_features([desc], bits=53)[0]
    hi, lo = check.dd_gemm(ta, tb, m, n, k, A, ar, B, br)
    C = torch.empty(m * n, dtype=torch.float64, device="cuda")
    umax = np.full(SMAX - SMIN + 1, np.inf)
    valid = np.zeros(SMAX - SMIN + 1, dtype=bool)
    for s in admissible(k):
        oz.gemm(ta, tb, m, n, k, 1.0, A, ar, B, br, 0.0, C, m, s)
        e = check.ulp_errors(C, hi, lo, m, m)
        umax[s - SMIN] = float(e.max())
        valid[s - SMIN] = True
    return feat, umax, valid


def with_bits(feat, bits):
    f = feat.copy()
    lk = np.float32(feat[0]) * np.float32(16)
    f[1] = np.float32(bits) / np.float32(64)
    f[46] = lk * np.float32(bits) / np.float32(1024)
    return f


def expand(feats, umax, valid):
    """rows per (problem, target): X, Y (met), M (valid s)"""
    X, Y, M, P, T = [], [], [], [], []
    for i in range(len(feats)):
        for b in TARGETS:
            X.append(with_bits(feats[i], b))
            Y.append((umax[i] <= 2.0 ** (53 - b)).astype(np.float32))
            M.append(valid[i].astype(np.float32))
            P.append(i)
            T.append(b)
    return (np.array(X, np.float32), np.array(Y, np.float32), np.array(M, np.float32), np.array(P), np.array(T))


def init_weights(rng):
    from_shapes = [(64, 48), (64,), (64, 64), (64,), (23, 64), (23,)]
    out = []
    for sh in from_shapes:
        if len(sh) == 2:
            out.append(rng.normal(0, math.sqrt(2.0 / sh[1]), sh).astype(np.float32).reshape(-1))
        else:
            out.append(np.zeros(sh, np.float32))
    return np.concatenate(out)


def train(oz, X, Y, M, rng, steps=20000, batch=1024, lr=2e-3, wd=1e-5):
    dev = "cuda"
    W = torch.from_numpy(init_weights(rng)).to(dev)
    Mm = torch.zeros_like(W)
    Vv = torch.zeros_like(W)
    G = torch.zeros_like(W)
    Xd, Yd, Md = (torch.from_numpy(a).to(dev) for a in (X, Y, M))
    nb = min(batch, len(X))
    H1 = torch.empty(nb * 64, device=dev)
    H2 = torch.empty_like(H1)
    D1 = torch.empty_like(H1)
    D2 = torch.empty_like(H1)
    Z = torch.empty(nb * 23, device=dev)
    D3 = torch.empty_like(Z)
    L = torch.empty_like(Z)
    g = torch.Generator(device="cpu").manual_seed(int(rng.integers(1 << 30)))
    t0 = time.time()
    for step in range(1, steps + 1):
        idx = torch.randint(0, len(X), (nb,), generator=g).to(dev)
        xb, yb, mb = Xd[idx].contiguous(), Yd[idx].contiguous(), Md[idx].contiguous()
        norm = float(mb.sum().clamp(min=1))
        oz.mlp_forward(xb, nb, W, H1, H2, Z, None)
        oz.mlp_backward(xb, nb, W, H1, H2, Z, yb, mb, 1.0 / norm, D1, D2, D3, G, L)
        cur_lr = lr * (0.5 * (1 + math.cos(math.pi * step / steps)) * 0.95 + 0.05)
        oz.mlp_adam(W, G, Mm, Vv, step, cur_lr, 0.9, 0.999, 1e-8, wd)
        if step % 2000 == 0 or step == 1:
            torch.cuda.synchronize()
            log(f"  step {step}: batch loss {float(L.sum()):.4f} ({time.time() - t0:.1f} s)")
    torch.cuda.synchronize()
    return W.cpu().numpy()


def predict(oz, W, X):
    dev = "cuda"
    Xd = torch.from_numpy(X).to(dev)
    P = torch.empty(len(X) * 23, device=dev)
    Wd = torch.from_numpy(W).to(dev)
    oz.mlp_forward(Xd, len(X), Wd, None, None, None, P)
    torch.cuda.synchronize()
    return P.cpu().numpy().reshape(len(X), 23)


def choose(P, M, tau):
    """the library's rule: smallest admissible s with p >= tau, else the largest admissible s"""
    out = np.zeros(len(P), dtype=int)
    for r in range(len(P)):
        adm = [t for t in range(23) if M[r, t] > 0]
        ok = [t for t in adm if P[r, t] >= tau]
        out[r] = (ok[0] if ok else adm[-1]) + SMIN
    return out


def evaluate(P, Y, M, fallback, tau):
    s = choose(P, M, tau)
    met = np.array([Y[r, s[r] - SMIN] > 0 for r in range(len(s))])
    fb_met = np.array([Y[r, fallback[r] - SMIN] > 0 for r in range(len(s))])
    opt = np.array([next((t + SMIN for t in range(23) if Y[r, t] > 0 and M[r, t] > 0), SMAX) for r in range(len(s))])
    return s, met, fb_met, opt


def write_inc(path, W, tau, info):
    with open(path, "w") as f:
        f.write("// Weights of the moduli selector as trained for this library; written by\n")
        f.write("// os2-xe_amd/selector/train.py (" + info + ").\n")
        f.write("// Layout: W1[64][48] b1[64] W2[64][64] b2[64] W3[23][64] b3[23], row-major.\n")
        f.write(f"static const float kTrainedWeights[{len(W)}] = {{\n")
        for i in range(0, len(W), 8):
            f.write("    " + ", ".join(f"{float(x):.9g}f" for x in W[i:i + 8]) + ",\n")
        f.write("};\n")
        f.write(f"constexpr float kTrainedThreshold = {tau:.6g}f;\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--problems", type=int, default=6000)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--data", default=None, help="reuse a dataset.npz instead of generating")
    args = ap.parse_args()
    os.makedirs(args.out, exist_ok=True)
    rng = np.random.default_rng(args.seed)
    oz = oz2.Ozaki2()
    if args.data:
        d = np.load(args.data)
        feats, umax, valid, ks = d["feats"], d["umax"], d["valid"], d["k"]
    else:
        feats, umax, valid, ks, meta = [], [], [], [], []
        t0 = time.time()
        for i in range(args.problems):
            pr = make_problem(rng)
            f, u, v = run_problem(oz, pr, rng)
            feats.append(f)
            umax.append(u)
            valid.append(v)
            ks.append(pr["k"])
            meta.append(pr)
            if (i + 1) % 250 == 0:
                log(f"data: {i + 1}/{args.problems} problems ({time.time() - t0:.0f} s)")
        feats, umax, valid, ks = np.array(feats), np.array(umax), np.array(valid), np.array(ks)
        np.savez_compressed(os.path.join(args.out, "dataset.npz"), feats=feats, umax=umax, valid=valid, k=ks)
        with open(os.path.join(args.out, "problems.json"), "w") as f:
            json.dump(meta, f)
    n = len(feats)
    perm = rng.permutation(n)
    nval = max(1, n // 6)
    val, tr = perm[:nval], perm[nval:]
    Xt, Yt, Mt, _, _ = expand(feats[tr], umax[tr], valid[tr])
    Xv, Yv, Mv, Pv, Tv = expand(feats[val], umax[val], valid[val])
    log(f"train rows {len(Xt)}, held-out rows {len(Xv)}")
    W = train(oz, Xt, Yt, Mt, rng, steps=args.steps)
    Pr = predict(oz, W, Xv)
    fb = np.array([oz2.Ozaki2.moduli_for_bits(int(ks[val][p_i]), int(b)) for p_i, b in zip(Pv, Tv)])
    # threshold: smallest tau (on a grid) whose choices meet the target on >= 99.5 % of held-out rows
    tau_best = 0.999
    for tau in np.round(np.concatenate([np.arange(0.5, 0.99, 0.01), np.arange(0.99, 0.9991, 0.001)]), 4):
        _, met, _, _ = evaluate(Pr, Yv, Mv, fb, tau)
        if met.mean() >= 0.995:
            tau_best = float(tau)
            break
    s, met, fb_met, opt = evaluate(Pr, Yv, Mv, fb, tau_best)
    report = {"problems": int(n), "held_out_problems": int(nval), "targets": list(TARGETS),
              "threshold": tau_best, "held_out_rows": int(len(Xv)),
              "learned_met_frac": float(met.mean()), "fallback_met_frac": float(fb_met.mean()),
              "learned_mean_s": float(s.mean()), "fallback_mean_s": float(fb.mean()), "oracle_min_mean_s": float(opt.mean())}
    for b in TARGETS:
        sel = Tv == b
        report[f"bits{b}"] = {"learned_met": float(met[sel].mean()), "fallback_met": float(fb_met[sel].mean()),
                              "learned_mean_s": float(s[sel].mean()), "fallback_mean_s": float(fb[sel].mean()),
                              "min_mean_s": float(opt[sel].mean())}
    log(json.dumps(report, indent=1))
    with open(os.path.join(args.out, "report.json"), "w") as f:
        json.dump(report, f, indent=1)
    np.save(os.path.join(args.out, "weights.npy"), W)
    write_inc(os.path.join(args.out, "selector_weights.inc"), W, tau_best,
              f"{n} problems, seed {args.seed}, {args.steps} Adam steps, held-out met fraction "
              f"{report['learned_met_frac']:.4f}")
    log("done")


if __name__ == "__main__":
    main()
