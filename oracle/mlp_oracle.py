"""Float64 restatement of the scheduler network (test infrastructure only).

Network: x (10) -> h = relu(W1 x + b1) (32) -> y = W2 h + b2 (8); flat
parameter layout W1[32][10] | b1[32] | W2[8][32] | b2[8].  Loss: mean over
samples and outputs of mask * (y - t)^2.  Adam as published (Kingma & Ba,
ICLR 2015, Algorithm 1).  The reference has no learned scheduler; these
functions are pinned by central finite differences and closed forms in
tests/test_mlp_oracle.py, not by reference vectors.
"""
import math

import numpy as np

IN, HID, OUT = 10, 32, 8
OFF_W1 = 0
OFF_B1 = OFF_W1 + HID * IN
OFF_W2 = OFF_B1 + HID
OFF_B2 = OFF_W2 + OUT * HID
NPARAM = OFF_B2 + OUT


def unpack(w):
    w = np.asarray(w, np.float64)
    assert w.shape == (NPARAM,)
    return (w[OFF_W1:OFF_B1].reshape(HID, IN), w[OFF_B1:OFF_W2],
            w[OFF_W2:OFF_B2].reshape(OUT, HID), w[OFF_B2:])


def forward(w, X):
    W1, b1, W2, b2 = unpack(w)
    X = np.asarray(X, np.float64)
    A = X @ W1.T + b1
    H = np.maximum(A, 0.0)
    return H @ W2.T + b2, H


def loss(w, X, T, mask=None):
    Y, _ = forward(w, X)
    M = np.ones_like(Y) if mask is None else np.asarray(mask, np.float64)
    D = (Y - np.asarray(T, np.float64)) * M
    return float((D * D).sum() / D.size)


def loss_grad(w, X, T, mask=None):
    """(loss, gradient) by backpropagation."""
    W1, b1, W2, b2 = unpack(w)
    X = np.asarray(X, np.float64)
    Y, H = forward(w, X)
    M = np.ones_like(Y) if mask is None else np.asarray(mask, np.float64)
    D = (Y - np.asarray(T, np.float64)) * M
    dY = 2.0 * D * M / D.size
    dW2 = dY.T @ H
    db2 = dY.sum(0)
    dH = (dY @ W2) * (H > 0)
    dW1 = dH.T @ X
    db1 = dH.sum(0)
    g = np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2])
    return float((D * D).sum() / D.size), g


def choose(w, x, fallback, margin=0.0):
    """Variant the device scheduler runs for feature vector x (1-based):
    the argmin of the predictions, unless the fallback (plain) variant's
    prediction is within `margin` of it or a prediction is not finite."""
    y, _ = forward(w, np.asarray(x, np.float64)[None, :])
    y = y[0]
    if not np.all(np.isfinite(y)):
        return fallback
    best = int(np.argmin(y))
    if 1 <= fallback <= OUT and not (y[fallback - 1] - y[best] > margin):
        return fallback
    return best + 1


def adam(w, X, T, mask=None, epochs=10, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """Full-batch Adam; returns (weights, loss before step 1, loss after)."""
    w = np.asarray(w, np.float64).copy()
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    first = None
    for t in range(1, epochs + 1):
        l, g = loss_grad(w, X, T, mask)
        if first is None:
            first = l
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        w -= lr * mh / (np.sqrt(vh) + eps)
    return w, first, loss(w, X, T, mask)


def features(n, m, count, sum_outdeg, max_outdeg, it, settled=False):
    """The 10 scheduler features of a frontier (float64); the last is 1 for
    runs whose values settle (BFS from a unit start), else 0."""
    fn = float(n) if n else 1.0
    fm = float(m) if m else 1.0
    fc, fs = float(count), float(sum_outdeg)
    avg = fs / fc if count else 0.0
    return np.array([
        math.log2(1 + fc) / 24, math.log2(1 + fs) / 32, math.log2(1 + max_outdeg) / 24,
        fc / fn, fs / fm, math.log2(1 + avg) / 16, math.log2(1 + it) / 16,
        math.log2(1 + fm / fn) / 8, math.log2(fn) / 24, 1.0 if settled else 0.0])
