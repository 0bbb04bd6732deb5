"""CPU checks of the selector's oracle (oracle/selector.py), pinned by mathematics:
central finite differences for the gradient, the closed form of Adam's first step,
hand-computed exponent-gap features, and the layout of the compiled-in weights."""
import os
import re

import numpy as np

from oracle import selector as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_parameter_count():
    assert S.NPARAMS == 8791


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    W = rng.normal(0, 0.3, S.NPARAMS)
    X = rng.normal(0, 1, (6, S.F))
    Y = (rng.random((6, S.O)) < 0.5).astype(np.float64)
    M = (rng.random((6, S.O)) < 0.8).astype(np.float64)
    g = S.backward(X, W, Y, M, 0.1)
    idx = rng.choice(S.NPARAMS, 300, replace=False)
    h = 1e-6
    for i in idx:
        Wp, Wm = W.copy(), W.copy()
        Wp[i] += h
        Wm[i] -= h
        fd = (S.loss(X, Wp, Y, M, 0.1) - S.loss(X, Wm, Y, M, 0.1)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_masked_outputs_have_no_gradient():
    rng = np.random.default_rng(4)
    W = rng.normal(0, 0.3, S.NPARAMS)
    X = rng.normal(0, 1, (3, S.F))
    Y = np.ones((3, S.O))
    M = np.zeros((3, S.O))
    assert not np.any(S.backward(X, W, Y, M, 1.0))


def test_adam_first_step_moves_by_lr_times_sign():
    rng = np.random.default_rng(5)
    W = rng.normal(0, 1, 50)
    G = rng.normal(0, 1, 50)
    W1, M1, V1 = S.adam(W, G, np.zeros(50), np.zeros(50), 1, 0.01, eps=0.0)
    assert np.allclose(W1, W - 0.01 * np.sign(G), rtol=0, atol=1e-12)
    assert np.allclose(M1, 0.1 * G) and np.allclose(V1, 0.001 * G * G)


def test_features_hand_computed():
    # op(A) = A (N): 2 rows x 4 k; row 0 = [1, 0.5, 0.25, 0], row 1 = [8, -8, 2**-30, 1]
    A = np.array([[1.0, 0.5, 0.25, 0.0], [8.0, -8.0, 2.0 ** -30, 1.0]]).reshape(-1, order="F")
    # op(B) = B (N): 4 k x 1 column, all equal
    B = np.array([3.0, 3.0, 3.0, 3.0])
    f = S.features("N", "N", 2, 1, 4, A, 2, B, 4, 53)
    assert f[0] == np.float32(2 / 16) and f[1] == np.float32(53 / 64) and f[47] == 1
    # A gaps of nonzero elements: row 0 -> 0,1,2 ; row 1 -> 0,0,33,3 ; zeros: 1 of 8
    hist = f[4:17]
    want = np.zeros(13)
    for g in (0, 1, 2, 0, 0, 33, 3):
        want[np.searchsorted(S.BIN_LO, g, side="right") - 1] += 1
    assert np.allclose(hist, want / 7)
    assert np.isclose(f[17], 1 / 8)
    assert np.isclose(f[32], (0 + 1 + 2 + 0 + 0 + 33 + 3) / 7 / 32) and np.isclose(f[33], 33 / 64)
    assert np.isclose(f[44], 0.5)  # one of two rows has a gap >= 24
    # B: single column, no gaps, no zeros
    assert f[18] == 1 and not f[19:32].any() and f[36] == 0 and f[45] == 0


def test_compiled_weights_file_layout():
    text = open(os.path.join(ROOT, "os2-xe_amd", "csrc", "selector_weights.inc")).read()
    n = int(re.search(r"kTrainedWeights\[(\d+)\]", text).group(1))
    assert n in (1, S.NPARAMS)
    tau = float(re.search(r"kTrainedThreshold = ([0-9.eE+-]+)f", text).group(1))
    assert 0 < tau < 1
