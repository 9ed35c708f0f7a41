"""Learned J interpolation of initial-state-selected reaction probabilities.

The reference sums P^J(E) over propagated J only (every J costs a full
propagation of its K blocks).  Here the propagator runs every `stride`-th J,
and a small network (3 x 128 tanh, FP32; libqd_hip.so `qd_mlp_*`, trained on
the GPU by its own forward, backward and Adam kernels) predicts P^J(E) of the
skipped J from three features:

    2 J / J_scale - 1,   2 (E - E_lo) / (E_hi - E_lo) - 1,   2 B_J / (E_hi - E_lo),

where B_J = J(J+1) / (2 mu R_b^2) is the centrifugal barrier at a
representative atom-diatom distance R_b.  The method without the network is
linear interpolation of P^J(E) in J at fixed E (`linear_in_J`); the
propagator itself judges both on held-out J.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from typing import Dict, Iterable, Optional

import numpy as np

from . import _abi

N_IN = 3


def _lib():
    lib = _abi.load_library()
    if not getattr(lib, "_qd_mlp_typed", False):
        f32p = C.POINTER(C.c_float)
        lib.qd_mlp_n_params.argtypes = [C.c_int32, C.POINTER(C.c_int64)]
        lib.qd_mlp_create.argtypes = [C.c_int32, f32p, C.c_int32, C.POINTER(C.c_void_p)]
        lib.qd_mlp_set_data.argtypes = [C.c_void_p, f32p, f32p, C.c_int32]
        lib.qd_mlp_loss_grad.argtypes = [C.c_void_p, f32p, f32p]
        lib.qd_mlp_train.argtypes = [C.c_void_p, C.c_int32, C.c_float, C.c_int32, f32p]
        lib.qd_mlp_predict.argtypes = [C.c_void_p, f32p, C.c_int32, f32p]
        lib.qd_mlp_get_params.argtypes = [C.c_void_p, f32p]
        lib.qd_mlp_destroy.argtypes = [C.c_void_p]
        lib.qd_mlp_last_error.restype = C.c_char_p
        lib._qd_mlp_typed = True
    return lib


def _check(lib, rc):
    if rc != 0:
        raise _abi.QDError(rc, lib.qd_mlp_last_error().decode())


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def n_params(n_in: int = N_IN) -> int:
    lib = _lib()
    n = C.c_int64()
    _check(lib, lib.qd_mlp_n_params(n_in, C.byref(n)))
    return int(n.value)


class MLP:
    """Handle over the GPU network (parameters, Adam moments, training set)."""

    def __init__(self, params: np.ndarray, n_in: int = N_IN, device: int = 0):
        self.lib = _lib()
        self.n_in = n_in
        p = _f32(params)
        if p.size != n_params(n_in):
            raise ValueError("parameter vector has %d entries, expected %d" % (p.size, n_params(n_in)))
        h = C.c_void_p()
        _check(self.lib, self.lib.qd_mlp_create(n_in, _ptr(p), device, C.byref(h)))
        self.h = h

    def set_data(self, X, T):
        self._X, self._T = _f32(X), _f32(T)
        _check(self.lib, self.lib.qd_mlp_set_data(self.h, _ptr(self._X), _ptr(self._T), int(self._T.size)))

    def loss_grad(self):
        g = np.empty(n_params(self.n_in), np.float32)
        L = C.c_float()
        _check(self.lib, self.lib.qd_mlp_loss_grad(self.h, C.byref(L), _ptr(g)))
        return float(L.value), g

    def train(self, iters: int, lr: float, use_graph: bool = True) -> float:
        L = C.c_float()
        _check(self.lib, self.lib.qd_mlp_train(self.h, int(iters), float(lr), 1 if use_graph else 0, C.byref(L)))
        return float(L.value)

    def predict(self, X) -> np.ndarray:
        x = _f32(X)
        y = np.empty(x.shape[0], np.float32)
        _check(self.lib, self.lib.qd_mlp_predict(self.h, _ptr(x), int(x.shape[0]), _ptr(y)))
        return y

    def params(self) -> np.ndarray:
        p = np.empty(n_params(self.n_in), np.float32)
        _check(self.lib, self.lib.qd_mlp_get_params(self.h, _ptr(p)))
        return p

    def close(self):
        if self.h:
            self.lib.qd_mlp_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def init_params(n_in: int = N_IN, seed: int = 0, width: int = 128) -> np.ndarray:
    """Xavier-uniform weights, zero biases, in the library's flat order."""
    rng = np.random.default_rng(seed)
    parts = []
    for s in [(width, n_in), (width,), (width, width), (width,), (width, width), (width,), (1, width), (1,)]:
        if len(s) == 2:
            lim = math.sqrt(6.0 / (s[0] + s[1]))
            parts.append(rng.uniform(-lim, lim, size=s).reshape(-1))
        else:
            parts.append(np.zeros(s).reshape(-1))
    return np.concatenate(parts).astype(np.float32)


def centrifugal_barrier(J, mu: float, R_b: float):
    """J(J+1) / (2 mu R_b^2) in hartree."""
    J = np.asarray(J, dtype=np.float64)
    return J * (J + 1.0) / (2.0 * mu * R_b * R_b)


class Features:
    def __init__(self, energies, mu: float, R_b: float, J_scale: float):
        self.e_lo, self.e_hi = float(np.min(energies)), float(np.max(energies))
        self.mu, self.R_b, self.J_scale = mu, R_b, J_scale

    def __call__(self, J, E) -> np.ndarray:
        """Centred on [-1, 1] for the tanh layers; E and B_J share one scale,
        so E - B_J (the barrier-shifted energy) is a linear combination."""
        J = np.asarray(J, dtype=np.float64)
        E = np.asarray(E, dtype=np.float64)
        span = self.e_hi - self.e_lo
        return np.stack([2.0 * J / self.J_scale - 1.0, 2.0 * (E - self.e_lo) / span - 1.0,
                         2.0 * centrifugal_barrier(J, self.mu, self.R_b) / span], axis=-1)


def linear_in_J(J_train: Iterable[int], P_train: Dict[int, np.ndarray], J: int) -> np.ndarray:
    """The method without the network: P^J(E) linearly interpolated in J at fixed E."""
    js = sorted(J_train)
    if J <= js[0]:
        return P_train[js[0]].copy()
    if J >= js[-1]:
        return P_train[js[-1]].copy()
    k = np.searchsorted(js, J)
    a, b = js[k - 1], js[k]
    w = (J - a) / (b - a)
    return (1 - w) * P_train[a] + w * P_train[b]


def partial_wave_sum(PJ: Dict[int, np.ndarray]) -> np.ndarray:
    """sum_J (2J+1) P^J(E); the cross section is pi/k^2 times this."""
    return sum((2 * J + 1) * P for J, P in PJ.items())


class JInterpolator:
    """Fit the network to P^J(E) of the propagated J and predict the others."""

    def __init__(self, energies, mu: float, R_b: float, J_max: int, seed: int = 0, device: int = 0):
        self.E = np.asarray(energies, dtype=np.float64)
        self.feat = Features(self.E, mu, R_b, max(1.0, float(J_max)))
        self.net = MLP(init_params(N_IN, seed), N_IN, device)
        self.train_seconds = 0.0
        self.loss = float("nan")

    def fit(self, P_train: Dict[int, np.ndarray], iters: int = 4000, lr: float = 2e-3, lr_final: float = 2e-4):
        J = np.concatenate([np.full(self.E.size, j) for j in sorted(P_train)])
        E = np.concatenate([self.E for _ in sorted(P_train)])
        T = np.concatenate([P_train[j] for j in sorted(P_train)])
        self.net.set_data(self.feat(J, E), T)
        t0 = time.perf_counter()
        # piecewise-constant learning-rate decay in four stages
        n = max(1, iters // 4)
        for k in range(4):
            lr_k = lr * (lr_final / lr) ** (k / 3.0)
            self.loss = self.net.train(n, lr_k)
        self.train_seconds = time.perf_counter() - t0
        return self.loss

    def predict(self, J: int) -> np.ndarray:
        y = self.net.predict(self.feat(np.full(self.E.size, J), self.E)).astype(np.float64)
        return np.clip(y, 0.0, 1.0)

    def close(self):
        self.net.close()
