"""Learned moduli selector on the GPU: the feature kernel, the MLP forward,
backward and Adam kernels against the CPU oracle (oracle/selector.py, itself
pinned by finite differences in tests/test_selector_oracle.py), training with
those kernels, and the fallback rule."""
import numpy as np
import pytest
import torch

import oz2
from conftest import make_mat
from gpu_cases import flat, log, to_dev
from oracle import oracle as O
from oracle import selector as S

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def oz():
    h = oz2.Ozaki2()
    yield h
    h.close()


FEATURE_CASES = [  # ta, tb, m, n, k, dist
    ("N", "N", 1, 1, 1, "u11"),
    ("N", "T", 37, 53, 129, "wide"),
    ("T", "N", 300, 130, 200, "phi"),
    ("T", "T", 129, 70, 5000, "wide"),
    ("N", "N", 200, 257, 3000, "ints"),
    ("N", "T", 64, 64, 64, "sparse"),
    ("N", "N", 2500, 40, 300, "wide"),    # strided sampled rows 19-20 apart: no two share a line
    ("N", "T", 1000, 20, 2500, "phi"),    # 7-8 apart, k subsampled; op(B) strided, 20 rows: a partial group
]


def _problem(ta, tb, m, n, k, dist, rng):
    ar, ac = (m, k) if ta == "N" else (k, m)
    br, bc = (k, n) if tb == "N" else (n, k)
    if dist == "sparse":
        A = make_mat(ar, ac, ar, "u11", rng)
        A[rng.random(A.shape) < 0.4] = 0
        B = make_mat(br, bc, br, "u11", rng)
        B[:, 0] = 0
    else:
        A = make_mat(ar, ac, ar, dist, rng, phi=2.0)
        B = make_mat(br, bc, br, dist, rng, phi=2.0)
    return flat(A), ar, flat(B), br


def test_features_match_oracle(oz):
    rng = np.random.default_rng(21)
    for (ta, tb, m, n, k, dist) in FEATURE_CASES:
        log(f"selector features {ta}{tb} m={m} n={n} k={k} {dist}")
        A, lda, B, ldb = _problem(ta, tb, m, n, k, dist, rng)
        d = dict(ta=ta, tb=tb, m=m, n=n, k=k, A=to_dev(A), lda=lda, B=to_dev(B), ldb=ldb)
        for bits in (40, 53):
            got = oz.selector_features([d], bits=bits)[0]
            want = S.features(ta, tb, m, n, k, A, lda, B, ldb, bits)
            assert np.allclose(got, want, rtol=2e-6, atol=1e-7), (np.nonzero(~np.isclose(got, want))[0], got, want)


def test_features_of_a_batch_match_one_by_one(oz):
    rng = np.random.default_rng(22)
    ds = []
    for (ta, tb, m, n, k, dist) in FEATURE_CASES:
        A, lda, B, ldb = _problem(ta, tb, m, n, k, dist, rng)
        ds.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, A=to_dev(A), lda=lda, B=to_dev(B), ldb=ldb))
    all_f = oz.selector_features(ds, bits=53)
    for i, d in enumerate(ds):
        assert np.array_equal(all_f[i], oz.selector_features([d], bits=53)[0])


def _rand_mlp(seed, nb):
    rng = np.random.default_rng(seed)
    W = rng.normal(0, 0.2, S.NPARAMS).astype(np.float32)
    X = rng.normal(0, 1, (nb, S.F)).astype(np.float32)
    Y = (rng.random((nb, S.O)) < 0.5).astype(np.float32)
    M = (rng.random((nb, S.O)) < 0.85).astype(np.float32)
    return W, X, Y, M


def test_mlp_forward_matches_oracle(oz):
    W, X, _, _ = _rand_mlp(1, 37)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    H1, H2 = torch.empty(37 * 64, device="cuda"), torch.empty(37 * 64, device="cuda")
    Z, P = torch.empty(37 * 23, device="cuda"), torch.empty(37 * 23, device="cuda")
    oz.mlp_forward(dev(X), 37, dev(W), H1, H2, Z, P)
    torch.cuda.synchronize()
    h1, h2, z, p = S.forward(X, W)
    assert np.allclose(H1.cpu().numpy().reshape(37, 64), h1, rtol=1e-5, atol=1e-5)
    assert np.allclose(H2.cpu().numpy().reshape(37, 64), h2, rtol=1e-5, atol=1e-5)
    assert np.allclose(Z.cpu().numpy().reshape(37, 23), z, rtol=1e-5, atol=1e-5)
    assert np.allclose(P.cpu().numpy().reshape(37, 23), p, rtol=1e-5, atol=1e-6)


def test_mlp_gradient_matches_oracle(oz):
    nb = 53
    W, X, Y, M = _rand_mlp(2, nb)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    Xd, Wd = dev(X), dev(W)
    H1, H2, D1, D2 = (torch.empty(nb * 64, device="cuda") for _ in range(4))
    Z, D3, L = (torch.empty(nb * 23, device="cuda") for _ in range(3))
    G = torch.empty(S.NPARAMS, device="cuda")
    inv = 1.0 / float(M.sum())
    oz.mlp_forward(Xd, nb, Wd, H1, H2, Z, None)
    oz.mlp_backward(Xd, nb, Wd, H1, H2, Z, dev(Y), dev(M), inv, D1, D2, D3, G, L)
    torch.cuda.synchronize()
    g = S.backward(X, W, Y, M, inv)
    got = G.cpu().numpy()
    scale = np.abs(g).max()
    assert np.max(np.abs(got - g)) <= 2e-5 * scale, np.max(np.abs(got - g)) / scale
    assert np.isclose(float(L.sum()), S.loss(X, W, Y, M, inv), rtol=1e-5)


def test_adam_matches_oracle(oz):
    rng = np.random.default_rng(7)
    W = rng.normal(0, 1, S.NPARAMS).astype(np.float32)
    G = rng.normal(0, 1, S.NPARAMS).astype(np.float32)
    M = rng.normal(0, 0.1, S.NPARAMS).astype(np.float32)
    V = rng.random(S.NPARAMS).astype(np.float32) * 0.01
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    Wd, Gd, Md, Vd = dev(W), dev(G), dev(M), dev(V)
    oz.mlp_adam(Wd, Gd, Md, Vd, 5, 1e-3, 0.9, 0.999, 1e-8, 1e-4)
    torch.cuda.synchronize()
    w, m, v = S.adam(W.astype(np.float64), G.astype(np.float64), M.astype(np.float64), V.astype(np.float64), 5, 1e-3,
                     wd=1e-4)
    # the kernel holds 1 - beta2 in float32 (0.00099998713 for 0.999): up to 1.3e-5 relative on the g^2 term
    assert np.allclose(Wd.cpu().numpy(), w, rtol=1e-6, atol=1e-6)
    assert np.allclose(Md.cpu().numpy(), m, rtol=2e-5, atol=1e-7)
    assert np.allclose(Vd.cpu().numpy(), v, rtol=4e-5, atol=1e-9)


def test_training_with_the_kernels_fits_a_rule(oz):
    """labels y_t = [x . u_t > 0]: 400 Adam steps with the library's kernels must cut the loss by 3x"""
    rng = np.random.default_rng(8)
    nb = 256
    U = rng.normal(0, 1, (S.O, S.F))
    X = rng.normal(0, 1, (2048, S.F)).astype(np.float32)
    Y = (X @ U.T > 0).astype(np.float32)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
    W = dev((rng.normal(0, 0.1, S.NPARAMS)).astype(np.float32))
    Mm, Vv, G = torch.zeros_like(W), torch.zeros_like(W), torch.zeros_like(W)
    H1, H2, D1, D2 = (torch.empty(nb * 64, device="cuda") for _ in range(4))
    Z, D3, L = (torch.empty(nb * 23, device="cuda") for _ in range(3))
    Xd, Yd = dev(X), dev(Y)
    losses = []
    for step in range(1, 401):
        i0 = ((step - 1) * nb) % 2048
        xb, yb = Xd[i0:i0 + nb].contiguous(), Yd[i0:i0 + nb].contiguous()
        oz.mlp_forward(xb, nb, W, H1, H2, Z, None)
        oz.mlp_backward(xb, nb, W, H1, H2, Z, yb, None, 1.0 / (nb * S.O), D1, D2, D3, G, L)
        oz.mlp_adam(W, G, Mm, Vv, step, 3e-3)
        if step in (1, 400):
            torch.cuda.synchronize()
            losses.append(float(L.sum()))
    log(f"training loss {losses[0]:.4f} -> {losses[1]:.4f}")
    assert losses[1] < losses[0] / 3


def test_unloaded_selector_falls_back(oz):
    rng = np.random.default_rng(9)
    saved, tau = oz.selector_weights()
    try:
        oz.selector_set_weights(np.zeros(0, np.float32), 0.5)
        ds = []
        for (ta, tb, m, n, k, dist) in FEATURE_CASES[1:4]:
            A, lda, B, ldb = _problem(ta, tb, m, n, k, dist, rng)
            ds.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, A=to_dev(A), lda=lda, B=to_dev(B), ldb=ldb))
        s, probs, learned = oz.selector_choose(ds, bits=53)
        assert s == [oz2.Ozaki2.moduli_for_bits(d["k"], 53) for d in ds]
        assert not any(learned) and not probs.any()
        with pytest.raises(oz2.OZ2InvalidArgument):
            oz.selector_set_weights(np.zeros(5, np.float32), 0.5)
    finally:
        if saved is not None:
            oz.selector_set_weights(saved, tau)


GAIN_CASES = [  # dist, m, n, k
    ("u11", 256, 192, 3000), ("u11", 400, 300, 700), ("wide12", 300, 256, 2000), ("wide12", 128, 512, 5000),
    ("wide12", 512, 128, 900), ("wide20", 256, 256, 4000), ("ints", 300, 300, 2000), ("ints", 128, 256, 8000),
    ("phi1", 256, 300, 1500), ("u11", 100, 700, 12000),
]


def _gain_matrix(dist, rows, cols, rng):
    if dist == "u11":
        v = rng.uniform(-1, 1, (rows, cols))
    elif dist.startswith("wide"):
        e = int(dist[4:])
        v = (1 + rng.random((rows, cols))) * rng.choice([-1.0, 1.0], (rows, cols)) * np.exp2(
            rng.integers(-e, e + 1, (rows, cols)).astype(float))
    elif dist == "ints":
        v = rng.integers(-1000, 1001, (rows, cols)).astype(float)
    else:
        v = rng.uniform(-0.5, 0.5, (rows, cols)) * np.exp(rng.standard_normal((rows, cols)))
    return np.asfortranarray(v).reshape(-1, order="F")


def test_trained_selector_gain(oz):
    """53-bit target (every entry within 1 ulp of the exact product): the trained network against
    the method without it (moduli_for_bits(k, 53)), on problems it was not trained on"""
    from oz2 import check
    saved, _ = oz.selector_weights()
    if saved is None:
        pytest.skip("no trained selector weights are compiled into this build")
    rng = np.random.default_rng(2025)
    rows = []
    for (dist, m, n, k) in GAIN_CASES:
        A = to_dev(_gain_matrix(dist, m, k, rng))
        B = to_dev(_gain_matrix(dist, k, n, rng))
        d = dict(ta="N", tb="N", m=m, n=n, k=k, A=A, lda=m, B=B, ldb=k)
        s_l = oz.selector_choose([d], bits=53)[0][0]
        s_f = oz2.Ozaki2.moduli_for_bits(k, 53)
        hi, lo = check.dd_gemm("N", "N", m, n, k, A, m, B, k)
        C = torch.empty(m * n, dtype=torch.float64, device="cuda")
        res = {}
        for name, sv in (("learned", s_l), ("fallback", s_f)):
            oz.gemm("N", "N", m, n, k, 1.0, A, m, B, k, 0.0, C, m, sv)
            res[name] = float(check.ulp_errors(C, hi, lo, m, m).max())
        rows.append((dist, m, n, k, s_l, res["learned"], s_f, res["fallback"]))
        log(f"{dist} {m}x{n}x{k}: learned s={s_l} max {res['learned']:.3g} ulp | moduli_for_bits s={s_f} "
            f"max {res['fallback']:.3g} ulp")
    met_l = sum(r[5] <= 1.0 for r in rows)
    met_f = sum(r[7] <= 1.0 for r in rows)
    log(f"target met: learned {met_l}/{len(rows)}, moduli_for_bits {met_f}/{len(rows)}")
    assert met_l >= len(rows) - 1 and met_l > met_f
    for r in rows:
        if r[0] == "ints":  # exact integer data needs a handful of moduli, not 15-16
            assert r[4] <= 8 < r[6] and r[5] <= 1.0, r


def test_choice_rule_threshold_and_largest_admissible(oz):
    # the choice rule with weights whose outputs are known: zero weights and output biases b3 give
    # probabilities sigmoid(b3) for every problem.  Outputs from s = 7 up above the threshold: the
    # smallest such admissible s (7, or the smallest admissible above it).  Every output below the
    # threshold: the largest admissible s (24), not moduli_for_bits.
    rng = np.random.default_rng(4)
    saved, tau = oz.selector_weights()
    b3 = oz2.SEL_PARAMS - oz2.SEL_OUTPUTS
    try:
        ds = []
        for (ta, tb, m, n, k, dist) in FEATURE_CASES[1:4]:
            A, lda, B, ldb = _problem(ta, tb, m, n, k, dist, rng)
            ds.append(dict(ta=ta, tb=tb, m=m, n=n, k=k, A=to_dev(A), lda=lda, B=to_dev(B), ldb=ldb))
        W = np.zeros(oz2.SEL_PARAMS, np.float32)
        W[b3:] = -20.0
        W[b3 + 5:] = 20.0
        oz.selector_set_weights(W, 0.5)
        s, probs, learned = oz.selector_choose(ds, bits=53)
        for d, sv in zip(ds, s):
            want = min(v for v in range(7, 25) if O.precision_bits(d["k"], v) > 0)
            assert sv == want, (d["k"], sv)
        assert all(learned)
        W[b3:] = -20.0
        oz.selector_set_weights(W, 0.5)
        s, probs, learned = oz.selector_choose(ds, bits=53)
        assert s == [24] * len(ds) and all(learned) and (probs < 0.5).all()
    finally:
        if saved is not None:
            oz.selector_set_weights(saved, tau)


def test_batch_reselect_after_new_contents(oz):
    # the learned s is fixed when a plan is made: integer data need few moduli; the same buffers then
    # hold wide-exponent data, which the old s cannot carry to 53 bits.  reselect() chooses again from
    # the current contents and rebuilds the plan; the rebuilt plan meets the target.
    from oz2 import check
    saved, _ = oz.selector_weights()
    if saved is None:
        pytest.skip("no trained selector weights are compiled into this build")
    rng = np.random.default_rng(31)
    m, n, k = 256, 192, 2000
    A = to_dev(_gain_matrix("ints", m, k, rng))
    B = to_dev(_gain_matrix("ints", k, n, rng))
    C = torch.empty(m * n, dtype=torch.float64, device="cuda")
    b = oz2.Batch(oz, [dict(ta="N", tb="N", m=m, n=n, k=k, alpha=1.0, A=A, lda=m, B=B, ldb=k, beta=0.0, C=C, ldc=m)],
                  bits=53, graph=True, select=True)
    try:
        s_int = b.moduli()[0]
        A.copy_(to_dev(_gain_matrix("wide20", m, k, rng)))
        B.copy_(to_dev(_gain_matrix("wide20", k, n, rng)))
        hi, lo = check.dd_gemm("N", "N", m, n, k, A, m, B, k)
        b.run()
        torch.cuda.synchronize()
        stale = float(check.ulp_errors(C, hi, lo, m, m).max())
        assert b.moduli()[0] == s_int
        assert b.reselect() == 1
        s_wide = b.moduli()[0]
        b.run()
        torch.cuda.synchronize()
        fresh = float(check.ulp_errors(C, hi, lo, m, m).max())
        log(f"reselect: ints s={s_int} on wide data {stale:.3g} ulp; reselected s={s_wide} {fresh:.3g} ulp")
        assert s_wide > s_int and stale > 1.0 and fresh <= 1.0
        assert b.reselect() == 0  # same contents, same choice
    finally:
        b.close()
