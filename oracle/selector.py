"""CPU oracle for the learned moduli selector (test infrastructure only).

The selector has no counterpart in the reference, so this file restates the
library's own definitions in float64 numpy and the tests pin it by
mathematics:
  * features(): the exponent-gap statistics of os2-xe_amd/csrc/selector_kernels.hip
    (same sampling of rows and k positions, same bins, same normalisation);
    the integer counts behind every histogram feature are exact, so the GPU
    features must match to float32 rounding;
  * forward() / backward() / adam(): the 48-64-64-23 MLP with ReLU hidden
    layers, sigmoid outputs and masked binary cross-entropy; backward() is
    checked against central finite differences of forward(), and the GPU
    kernels against backward().
"""
import numpy as np

F, H, O = 48, 64, 23
ROWS, COLS = 128, 2048
BIN_LO = [0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64]
SIZES = [H * F, H, H * H, H, O * H, O]
NPARAMS = sum(SIZES)


def unpack(W):
    W = np.asarray(W, dtype=np.float64)
    out, o = [], 0
    for n in SIZES:
        out.append(W[o:o + n])
        o += n
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(H, F), b1, W2.reshape(H, H), b2, W3.reshape(O, H), b3


def forward(X, W):
    """X (nb, 48) -> (H1, H2, Z, P)"""
    W1, b1, W2, b2, W3, b3 = unpack(W)
    X = np.asarray(X, dtype=np.float64)
    H1 = np.maximum(X @ W1.T + b1, 0.0)
    H2 = np.maximum(H1 @ W2.T + b2, 0.0)
    Z = H2 @ W3.T + b3
    return H1, H2, Z, 1.0 / (1.0 + np.exp(-Z))


def loss(X, W, Y, M, inv_norm):
    _, _, Z, _ = forward(X, W)
    M = np.ones_like(Z) if M is None else M
    return float(np.sum(M * (np.maximum(Z, 0) - Z * Y + np.log1p(np.exp(-np.abs(Z))))) * inv_norm)


def backward(X, W, Y, M, inv_norm):
    """gradient of loss() with respect to the flat parameter vector"""
    W1, b1, W2, b2, W3, b3 = unpack(W)
    X = np.asarray(X, dtype=np.float64)
    H1, H2, Z, P = forward(X, W)
    M = np.ones_like(Z) if M is None else M
    D3 = M * (P - Y) * inv_norm
    D2 = (D3 @ W3) * (H2 > 0)
    D1 = (D2 @ W2) * (H1 > 0)
    g = [D1.T @ X, D1.sum(0), D2.T @ H1, D2.sum(0), D3.T @ H2, D3.sum(0)]
    return np.concatenate([x.reshape(-1) for x in g])


def adam(W, G, M, V, step, lr, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    """one Adam step (Kingma and Ba 2015) with bias correction; returns new (W, M, V)"""
    g = G + wd * W
    M = b1 * M + (1 - b1) * g
    V = b2 * V + (1 - b2) * g * g
    mh = M / (1 - b1 ** step)
    vh = V / (1 - b2 ** step)
    return W - lr * mh / (np.sqrt(vh) + eps), M, V


def _side_stats(X, ld, contig, rows, k):
    """counts for the sampled rows of one side; X flat column-major float64"""
    nr = min(rows, ROWS)
    npos = min(k, COLS)
    cnt = np.zeros(len(BIN_LO), dtype=np.int64)
    zeros = gap_sum = gap_max = wide = 0
    rowexp = []
    pos = (np.arange(npos, dtype=np.int64) * k) // npos
    for r in range(nr):
        i = r * rows // nr
        idx = pos + i * ld if contig else i + pos * ld
        bits = X[idx].view(np.uint64) & np.uint64(0x7FFFFFFFFFFFFFFF)
        e = (bits >> np.uint64(52)).astype(np.int64)
        e0 = int(e.max())
        nz = bits != 0
        zeros += int((~nz).sum())
        g = np.where((e >= 0x7FF) | (e0 >= 0x7FF), 0, e0 - e)[nz]
        if g.size:
            gap_sum += int(g.sum())
            gm = int(g.max())
            gap_max = max(gap_max, gm)
            wide += 1 if gm >= 24 else 0
            b = np.searchsorted(BIN_LO, g, side="right") - 1
            cnt += np.bincount(b, minlength=len(BIN_LO))
        rowexp.append(e0)
    return cnt, zeros, gap_sum, gap_max, wide, np.array(rowexp, dtype=np.float64)


def _quantile(cnt, total, q):
    if total == 0:
        return 0.0
    want = q * total
    run = 0.0
    for b, c in enumerate(cnt):
        run += c
        if run >= want:
            return float(BIN_LO[b])
    return float(BIN_LO[-1])


def features(ta, tb, m, n, k, A, lda, B, ldb, bits):
    """48 features of one problem (A, B flat column-major float64 host arrays)"""
    f = np.zeros(F, dtype=np.float32)
    f32 = np.float32
    lk = np.log2(f32(max(k, 1)), dtype=np.float32)
    f[0] = lk / f32(16)
    f[1] = f32(bits) / f32(64)
    f[2] = np.log2(f32(max(m, 1)), dtype=np.float32) / f32(16)
    f[3] = np.log2(f32(max(n, 1)), dtype=np.float32) / f32(16)
    sides = [(A, lda, ta != "N", m), (B, ldb, tb == "N", n)]
    for s, (X, ld, contig, rows) in enumerate(sides):
        cnt, zeros, gsum, gmax, wide, rowexp = _side_stats(np.asarray(X, dtype=np.float64), ld, contig, rows, k)
        nz = int(cnt.sum())
        allc = nz + zeros
        h = 4 + 14 * s
        for b in range(len(BIN_LO)):
            f[h + b] = f32(cnt[b] / nz) if nz else 0
        f[h + 13] = f32(zeros / allc) if allc else 0
        g = 32 + 4 * s
        f[g] = f32(gsum / nz) / f32(32) if nz else 0
        f[g + 1] = f32(min(gmax, 128)) / f32(64)
        f[g + 2] = f32(_quantile(cnt, nz, f32(0.90))) / f32(64)
        f[g + 3] = f32(_quantile(cnt, nz, f32(0.99))) / f32(64)
        nr = len(rowexp)
        if nr:
            f[40 + 2 * s] = f32(rowexp.max() - rowexp.min()) / f32(64)
            f[41 + 2 * s] = f32(np.abs(rowexp - rowexp.mean()).sum() / nr) / f32(16)
            f[44 + s] = f32(wide) / f32(nr)
    f[46] = lk * f32(bits) / f32(1024)
    f[47] = 1
    return f
