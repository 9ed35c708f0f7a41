"""The numpy network oracle (oracle/mlp_oracle.py) on the CPU: its gradient
against central finite differences of its own loss (so a dropped term, a
transposed weight or a wrong activation derivative in the backward pass
fails), Adam against the closed form of its first step, and the host-side
features and baseline interpolation of quantum-dynamics_amd/jinterp.py."""
import importlib

import numpy as np
import pytest


@pytest.fixture(scope="module")
def mo():
    import mlp_oracle
    return mlp_oracle


def test_gradient_matches_finite_differences(mo):
    rng = np.random.default_rng(3)
    n_in = 3
    p = mo.init_params(n_in, seed=1) + 0.05 * rng.standard_normal(mo.n_params(n_in))
    X = rng.uniform(-1, 1, size=(40, n_in))
    T = rng.uniform(0, 1, size=40)
    L, g = mo.loss_grad(p, X, T, n_in)
    shapes = mo.shapes(n_in)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in shapes])
    # a few entries of every parameter block
    idx = np.concatenate([o + rng.choice(int(np.prod(s)), size=min(6, int(np.prod(s))), replace=False)
                          for o, s in zip(offs[:-1], shapes)])
    h = 1e-6
    for q in idx:
        e = np.zeros_like(p)
        e[q] = h
        fd = (mo.loss_grad(p + e, X, T, n_in)[0] - mo.loss_grad(p - e, X, T, n_in)[0]) / (2 * h)
        assert abs(fd - g[q]) < 1e-7 + 1e-5 * abs(g[q]), (q, fd, g[q])


def test_loss_is_mean_squared_error(mo):
    rng = np.random.default_rng(5)
    p = mo.init_params(2, seed=4)
    X = rng.uniform(size=(7, 2))
    T = rng.uniform(size=7)
    y, _ = mo.forward(p, X, 2)
    L, _ = mo.loss_grad(p, X, T, 2)
    assert abs(L - np.mean((y - T) ** 2)) < 1e-15


def test_adam_first_step_is_signed_lr(mo):
    """With zero moments, the first bias-corrected Adam step is -lr g/(|g|+eps)."""
    g = np.array([0.3, -2.0, 1e-3, 0.0])
    p0 = np.ones(4)
    opt = mo.Adam(4, lr=0.01)
    p1 = opt.step(p0, g)
    assert np.allclose(p1 - p0, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def test_training_reduces_loss_on_a_smooth_target(mo):
    rng = np.random.default_rng(0)
    X = rng.uniform(-1, 1, size=(64, 2))
    T = 0.5 + 0.4 * np.tanh(2 * X[:, 0] - X[:, 1])
    p, hist = mo.train(mo.init_params(2, seed=2), X, T, 2, iters=150, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]


def test_host_features_and_baseline(qd):
    ji = importlib.import_module("quantum-dynamics_amd.jinterp")
    E = np.linspace(0.01, 0.05, 5)
    f = ji.Features(E, mu=1000.0, R_b=3.0, J_scale=40.0)
    F = f([0, 10], [0.01, 0.05])
    assert np.allclose(F[0], [-1.0, -1.0, 0.0])
    assert np.allclose(F[1], [-0.5, 1.0, 2 * 110.0 / (2 * 1000.0 * 9.0) / 0.04])
    P = {0: np.full(5, 0.8), 4: np.full(5, 0.4)}
    assert np.allclose(ji.linear_in_J([0, 4], P, 1), 0.7)
    assert np.allclose(ji.linear_in_J([0, 4], P, 9), 0.4)
    assert np.allclose(ji.partial_wave_sum(P), 0.8 + 9 * 0.4)
