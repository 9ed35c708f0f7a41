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
propagator itself checks both on held-out J.
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
        lib.qd_mlp_predict_noalloc.argtypes = [C.c_void_p, f32p, C.c_int32, f32p]
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

    def predict_noalloc(self, X) -> np.ndarray:
        """predict() through qd_mlp_predict_noalloc: no device allocation per
        call, at most QD_MLP_NOALLOC_MAX (65536) samples."""
        x = _f32(X)
        y = np.empty(x.shape[0], np.float32)
        _check(self.lib, self.lib.qd_mlp_predict_noalloc(self.h, _ptr(x), int(x.shape[0]), _ptr(y)))
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


def propagate_pj(qd, base, system, Js, jmax: int, steps: int, chunk: int = 1000, log=None, **jk_kw):
    """P^J(E) of every J in Js from one batched mixed-precision propagation
    (all K blocks of every J, Coriolis-coupled; each J's flux normalised by
    the packet's |eta_J(E)|^2 in partial wave J), and the seconds it took:
    problem set-up, handle creation (upload, tables, autotuning) and the
    propagation itself."""
    from . import problem as pbm
    from .propagator import JKPropagator

    t0 = time.perf_counter()
    jk = pbm.make_jk_problem(base, system, list(Js), jmax=jmax, **jk_kw)
    t1 = time.perf_counter()
    with JKPropagator(jk, precision="mixed", autotune=1) as prop:
        t2 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            prop.propagate(n)
            done += n
            if log:
                log("  J %s: %d / %d steps, %.1f s" % (_jrange(Js), done, steps, time.perf_counter() - t2))
        PJ = prop.crp_by_J()
    t3 = time.perf_counter()
    points = int(jk.psi.size)
    del jk
    return PJ, {"setup_s": t1 - t0, "create_s": t2 - t1, "propagate_s": t3 - t2, "total_s": t3 - t0,
                "grid_points": points, "blocks": sum(min(J, jmax) + 1 for J in Js)}


def _jrange(Js):
    Js = sorted(Js)
    return "%d..%d (%d values)" % (Js[0], Js[-1], len(Js))


def learned_partial_waves(qd, base, system, J_max: int = 40, stride: int = 4, n_held: int = 5, jmax: int = 60,
                          steps: Optional[int] = None, R_b: float = 3.5, iters: int = 20000, reference: str = "all",
                          log=None, seed: int = 0) -> dict:
    """The learned J interpolation end to end.

    Propagates every `stride`-th J of 0..J_max, trains the network on their
    P^J(E) on the GPU, predicts the skipped J, and checks the result with the
    propagator itself: `reference` = "all" propagates every J (the partial-wave
    sum without the network and its wall time), "held" only `n_held` skipped J
    spread over the range.  Reports the per-energy error of the partial-wave
    sum sum_J (2J+1) P^J(E) (the cross section up to the common factor pi/k^2)
    for the network and for linear interpolation in J, the error on the
    held-out J, and the wall times."""
    steps = int(steps if steps is not None else base.total_steps)
    J_all = list(range(J_max + 1))
    J_train = J_all[::stride]
    if J_train[-1] != J_max:
        J_train.append(J_max)
    skipped = [J for J in J_all if J not in J_train]
    held = skipped[::max(1, len(skipped) // n_held)][:n_held]
    E = np.asarray(base.energies, dtype=np.float64)

    PJ_train, t_train = propagate_pj(qd, base, system, J_train, jmax, steps, log=log)
    t0 = time.perf_counter()
    model = JInterpolator(E, mu=system.mu_R, R_b=R_b, J_max=J_max, seed=seed)
    loss = model.fit(PJ_train, iters=iters, lr=3e-3, lr_final=1e-4)
    pred = {J: (PJ_train[J] if J in PJ_train else model.predict(J)) for J in J_all}
    fit_s = time.perf_counter() - t0
    model.close()
    lin = {J: (PJ_train[J] if J in PJ_train else linear_in_J(J_train, PJ_train, J)) for J in J_all}

    if reference == "all":
        PJ_ref, t_ref = propagate_pj(qd, base, system, J_all, jmax, steps, log=log)
    else:
        PJ_ref, t_ref = propagate_pj(qd, base, system, held, jmax, steps, log=log)
        PJ_ref.update(PJ_train)
    out = {
        "J_max": J_max, "stride": stride, "J_propagated": J_train, "J_held_out": held, "steps": steps,
        "energies_hartree": [float(E[0]), float(E[-1]), int(E.size)],
        "train": t_train, "fit_seconds": fit_s, "training_loss": loss, "reference": reference,
        "reference_run": t_ref,
    }
    d_net = {J: np.abs(pred[J] - PJ_ref[J]) for J in held}
    d_lin = {J: np.abs(lin[J] - PJ_ref[J]) for J in held}
    out["held_out_max_abs_dP"] = {"network": {J: float(d_net[J].max()) for J in held},
                                  "linear_in_J": {J: float(d_lin[J].max()) for J in held}}
    if reference == "all":
        # every skipped J judged by the propagator, not only the held-out ones
        s_net = {J: float(np.max(np.abs(pred[J] - PJ_ref[J]))) for J in skipped}
        s_lin = {J: float(np.max(np.abs(lin[J] - PJ_ref[J]))) for J in skipped}
        out["skipped_max_abs_dP"] = {
            "network": s_net, "linear_in_J": s_lin,
            "bound": {"network": max(s_net.values()), "linear_in_J": max(s_lin.values())},
            "mean": {"network": float(np.mean(list(s_net.values()))),
                     "linear_in_J": float(np.mean(list(s_lin.values())))},
            "network_better": int(sum(s_net[J] < s_lin[J] for J in skipped)), "skipped": len(skipped)}
        ref_sum = partial_wave_sum(PJ_ref)
        scale = float(np.max(np.abs(ref_sum)))
        e_net = np.abs(partial_wave_sum(pred) - ref_sum)
        e_lin = np.abs(partial_wave_sum(lin) - ref_sum)
        out["partial_wave_sum_max"] = scale
        out["cross_section_rel_error"] = {
            "network": {"max": float(e_net.max() / scale), "rms": float(np.sqrt(np.mean(e_net ** 2)) / scale)},
            "linear_in_J": {"max": float(e_lin.max() / scale), "rms": float(np.sqrt(np.mean(e_lin ** 2)) / scale)},
        }
        out["per_energy"] = {"E_hartree": E.tolist(), "reference": ref_sum.tolist(),
                             "network": partial_wave_sum(pred).tolist(), "linear_in_J": partial_wave_sum(lin).tolist(),
                             "PJ_reference": {int(J): PJ_ref[J].tolist() for J in J_all},
                             "PJ_network": {int(J): pred[J].tolist() for J in J_all}}
        learned = t_train["total_s"] + fit_s
        out["wall_seconds"] = {"learned": learned, "all_J": t_ref["total_s"], "speedup": t_ref["total_s"] / learned}
    return out
