"""numpy oracle of the J-interpolation network (float64).

TEST INFRASTRUCTURE ONLY (imported by tests/).  It states the same network,
loss and optimiser as the HIP kernels in quantum-dynamics_amd/csrc/qd_mlp.hpp
so that their gradients and training steps can be checked; the oracle's own
backward pass is checked against central finite differences
(tests/test_mlp_oracle.py).  The network has no counterpart in the reference.

Network: x (n_in) -> tanh(W1 x + b1) -> tanh(W2 h + b2) -> tanh(W3 h + b3)
         -> y = w4 . h + b4,  width 128.
Loss:    L = (1/n) sum_s (y_s - t_s)^2.
Adam:    m = b1 m + (1-b1) g;  v = b2 v + (1-b2) g^2;
         p -= lr * (m / (1 - b1^k)) / (sqrt(v / (1 - b2^k)) + eps).
Parameters are one flat vector in the order W1[o][i], b1, W2, b2, W3, b3, w4, b4.
"""
from __future__ import annotations

import numpy as np

WIDTH = 128
DEPTH = 3


def shapes(n_in, width=WIDTH):
    return [(width, n_in), (width,), (width, width), (width,), (width, width), (width,), (1, width), (1,)]


def n_params(n_in, width=WIDTH):
    return sum(int(np.prod(s)) for s in shapes(n_in, width))


def unpack(p, n_in, width=WIDTH):
    out, o = [], 0
    for s in shapes(n_in, width):
        k = int(np.prod(s))
        out.append(p[o:o + k].reshape(s))
        o += k
    return out


def init_params(n_in, seed=0, width=WIDTH):
    """Xavier-uniform weights, zero biases (the host draws these for the GPU too)."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in shapes(n_in, width):
        if len(s) == 2:
            lim = np.sqrt(6.0 / (s[0] + s[1]))
            parts.append(rng.uniform(-lim, lim, size=s).reshape(-1))
        else:
            parts.append(np.zeros(s).reshape(-1))
    return np.concatenate(parts)


def forward(p, X, n_in, width=WIDTH):
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p, n_in, width)
    h1 = np.tanh(X @ W1.T + b1)
    h2 = np.tanh(h1 @ W2.T + b2)
    h3 = np.tanh(h2 @ W3.T + b3)
    y = h3 @ w4[0] + b4[0]
    return y, (h1, h2, h3)


def loss_grad(p, X, T, n_in, width=WIDTH):
    """Mean squared error and its gradient with respect to the flat parameters."""
    W1, b1, W2, b2, W3, b3, w4, b4 = unpack(p, n_in, width)
    n = X.shape[0]
    y, (h1, h2, h3) = forward(p, X, n_in, width)
    r = y - T
    L = float(np.mean(r * r))
    dy = 2.0 * r / n                              # dL/dy
    gw4 = dy @ h3
    gb4 = np.array([dy.sum()])
    d3 = np.outer(dy, w4[0]) * (1 - h3 * h3)      # dL/d(pre-activation of layer 3)
    gW3 = d3.T @ h2
    gb3 = d3.sum(0)
    d2 = (d3 @ W3) * (1 - h2 * h2)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (d2 @ W2) * (1 - h1 * h1)
    gW1 = d1.T @ X
    gb1 = d1.sum(0)
    g = np.concatenate([gW1.reshape(-1), gb1, gW2.reshape(-1), gb2, gW3.reshape(-1), gb3, gw4, gb4])
    return L, g


class Adam:
    def __init__(self, n, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = np.zeros(n)
        self.v = np.zeros(n)
        self.k = 0
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps

    def step(self, p, g):
        self.k += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.k)
        vh = self.v / (1 - self.b2 ** self.k)
        return p - self.lr * mh / (np.sqrt(vh) + self.eps)


def train(p, X, T, n_in, iters, lr=1e-3, width=WIDTH):
    opt = Adam(p.size, lr)
    hist = []
    for _ in range(iters):
        L, g = loss_grad(p, X, T, n_in, width)
        hist.append(L)
        p = opt.step(p, g)
    return p, hist
