"""Pins the float64 scheduler-network oracle (oracle/mlp_oracle.py).

Backpropagation is checked against central finite differences of the loss;
the forward pass against a hand-computed network; Adam's first step against
its closed form (every parameter moves by lr * sign(g)); the mask against
dropping the masked targets.  A swapped layer index, a missing ReLU gate or
a wrong loss normalisation fails one of these.
"""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def rand_problem(seed, B=13):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.5, M.NPARAM)
    X = rng.uniform(-1, 1, (B, M.IN))
    T = rng.normal(0, 1, (B, M.OUT))
    return w, X, T


@pytest.mark.parametrize("seed", range(3))
def test_gradient_matches_finite_differences(seed):
    w, X, T = rand_problem(seed)
    mask = (np.random.default_rng(seed + 7).uniform(size=T.shape) > 0.3).astype(float)
    for msk in (None, mask):
        l, g = M.loss_grad(w, X, T, msk)
        assert l == pytest.approx(M.loss(w, X, T, msk), rel=1e-12)
        h = 1e-6
        num = np.empty_like(w)
        for p in range(M.NPARAM):
            e = np.zeros_like(w)
            e[p] = h
            num[p] = (M.loss(w + e, X, T, msk) - M.loss(w - e, X, T, msk)) / (2 * h)
        assert np.allclose(g, num, rtol=1e-5, atol=1e-8)


def test_forward_by_hand():
    w = np.zeros(M.NPARAM)
    W1, b1, W2, b2 = [a.copy() for a in M.unpack(w)]
    # hidden unit 0 = relu(x0 - x1 + 1), unit 1 = relu(-x0); output 3 = 2 h0 - h1 + 0.5
    W1[0, 0], W1[0, 1], b1[0] = 1, -1, 1
    W1[1, 0] = -1
    W2[3, 0], W2[3, 1], b2[3] = 2, -1, 0.5
    w = np.concatenate([W1.ravel(), b1, W2.ravel(), b2])
    X = np.zeros((2, M.IN))
    X[0, :2] = [0.25, 2.0]    # h0 = relu(-0.75) = 0, h1 = 0
    X[1, :2] = [-0.5, 0.0]    # h0 = 0.5, h1 = 0.5
    Y, H = M.forward(w, X)
    assert H[0, :2].tolist() == [0, 0] and H[1, :2].tolist() == [0.5, 0.5]
    assert Y[0, 3] == 0.5 and Y[1, 3] == 2 * 0.5 - 0.5 + 0.5
    assert np.all(Y[:, [0, 1, 2, 4, 5, 6, 7]] == 0)
    assert M.choose(w, X[1], 9) == 1          # ties -> first output
    w2 = w.copy()
    w2[M.OFF_B2 + 5] = -1
    assert M.choose(w2, X[0], 9) == 6
    # the plain (fallback) choice is kept unless beaten by more than margin
    # output 4 predicts 0.5 and output 6 predicts -1 for X[0]: a 1.5 gap
    assert M.choose(w2, X[0], 4) == 6 and M.choose(w2, X[0], 4, margin=1.4) == 6
    assert M.choose(w2, X[0], 4, margin=1.5) == 4 and M.choose(w2, X[0], 6, margin=2) == 6
    w2[M.OFF_B2] = np.nan
    assert M.choose(w2, X[0], 9) == 9


def test_adam_first_step_is_sign_step():
    w, X, T = rand_problem(4)
    lr = 1e-3
    w1, first, _ = M.adam(w, X, T, epochs=1, lr=lr)
    _, g = M.loss_grad(w, X, T)
    moved = np.abs(g) > 1e-3     # |g| >> eps: the step is lr * g / |g|
    assert np.allclose((w - w1)[moved], lr * np.sign(g[moved]), rtol=1e-4)
    assert first == pytest.approx(M.loss(w, X, T))


def test_mask_equals_dropping_targets():
    w, X, T = rand_problem(5, B=4)
    mask = np.ones_like(T)
    mask[:, 2] = 0
    l, g = M.loss_grad(w, X, T, mask)
    T2 = T.copy()
    Y, _ = M.forward(w, X)
    T2[:, 2] = Y[:, 2]            # a target equal to the prediction adds nothing
    l2, g2 = M.loss_grad(w, X, T2)
    assert l == pytest.approx(l2) and np.allclose(g, g2)


def test_adam_decreases_loss():
    w, X, T = rand_problem(6, B=64)
    _, first, last = M.adam(w, X, T, epochs=200, lr=1e-2)
    assert last < 0.5 * first


def test_features_match_definition():
    f = M.features(1 << 20, 16 << 20, 1023, 4095 * 16, 300, 7)
    assert f[0] == pytest.approx(10 / 24) and f[3] == pytest.approx(1023 / (1 << 20))
    assert f[4] == pytest.approx(4095 * 16 / (16 << 20)) and f[8] == pytest.approx(20 / 24)
    assert f[6] == pytest.approx(3 / 16) and f[7] == pytest.approx(np.log2(17) / 8)
    assert M.features(100, 0, 0, 0, 0, 0)[5] == 0
    assert f[9] == 0 and M.features(1 << 20, 16 << 20, 1023, 4095 * 16, 300, 7, True)[9] == 1
