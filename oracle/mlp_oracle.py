"""FP64 numpy restatement of the learned proposer's MLP (n_in -> 64 -> 64 ->
1, tanh) and of its mean-squared-error gradient.  TEST INFRASTRUCTURE ONLY:
the checker for the FP32 HIP forward/backward kernels (gpe_mlp_*)."""
from __future__ import annotations

import numpy as np

H = 64


def unpack(p, nin):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * nin].reshape(H, nin); o += H * nin
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, X, nin):
    W1, b1, W2, b2, W3, b3 = unpack(p, nin)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    return h2 @ W3 + b3


def loss(p, X, Y, nin):
    return float(np.mean((forward(p, X, nin) - Y) ** 2))


def grad(p, X, Y, nin):
    """Analytic gradient of mean((f(x) - y)^2), packed like p."""
    W1, b1, W2, b2, W3, b3 = unpack(p, nin)
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[0]
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    out = h2 @ W3 + b3
    d = 2.0 * (out - Y) / n
    gW3 = h2.T @ d
    gb3 = d.sum()
    dz2 = (d[:, None] * W3[None, :]) * (1 - h2 ** 2)
    gW2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2) * (1 - h1 ** 2)
    gW1 = dz1.T @ X
    gb1 = dz1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])


def fd_grad(p, X, Y, nin, idx, h=1e-6):
    """Central finite differences of the loss for the parameter indices idx."""
    p = np.asarray(p, dtype=np.float64)
    out = []
    for i in idx:
        a = p.copy(); a[i] += h
        b = p.copy(); b[i] -= h
        out.append((loss(a, X, Y, nin) - loss(b, X, Y, nin)) / (2 * h))
    return np.array(out)
