"""ctypes mirror of include/qd.h and the loader of the HIP library.

The structs are laid out field for field as in qd.h; ``problem_struct`` fills
one from a :class:`Problem` and keeps the numpy arrays it points into alive.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libqd_hip.so")

QD_OK = 0
QD_PREC_FP64, QD_PREC_MIXED = 0, 1
QD_PIPE_FUSED, QD_PIPE_PLAIN = 0, 1
QD_VARIANT_CPU, QD_VARIANT_CUDA = 0, 1

_dp = C.POINTER(C.c_double)


class qd_radial(C.Structure):
    _fields_ = [("n", C.c_int32), ("r", _dp), ("dr", C.c_double), ("mass", C.c_double)]


class qd_angle(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("x", _dp), ("w", _dp), ("legendre", _dp)]


class qd_time(C.Structure):
    _fields_ = [("total_steps", C.c_int32), ("steps", C.c_int32), ("time_step", C.c_double)]


class qd_crp(C.Structure):
    _fields_ = [("n_dividing_surface", C.c_int32), ("n_gradient_points", C.c_int32),
                ("n_energies", C.c_int32), ("calculate_CRP", C.c_int32),
                ("energies", _dp), ("eta_sq", _dp), ("CRP", _dp)]


class qd_problem(C.Structure):
    _fields_ = [("r1", qd_radial), ("r2", qd_radial), ("theta", qd_angle), ("pot", _dp),
                ("psi", _dp), ("time", qd_time), ("dump1", _dp), ("dump2", _dp), ("crp", qd_crp)]


class qd_options(C.Structure):
    _fields_ = [("precision", C.c_int32), ("pipeline", C.c_int32), ("variant", C.c_int32),
                ("flux_block", C.c_int32), ("crp_every", C.c_int32), ("use_graph", C.c_int32),
                ("device", C.c_int32), ("profile", C.c_int32),
                ("theta_split", C.c_int32), ("autotune", C.c_int32), ("fuse_theta", C.c_int32),
                ("diagnostics_every", C.c_int32)]


class qd_packet(C.Structure):
    _fields_ = [("R0", C.c_double), ("delta", C.c_double), ("k0", C.c_double), ("v0", C.c_int32),
                ("j0", C.c_int32), ("J", C.c_int32), ("K", C.c_int32), ("vdiat", _dp)]


class qd_jk_block(C.Structure):
    _fields_ = [("J", C.c_int32), ("K", C.c_int32), ("pot", _dp), ("psi", _dp), ("eta_sq", _dp)]


class qd_jk_problem(C.Structure):
    _fields_ = [("base", qd_problem), ("jmax", C.c_int32), ("coriolis", C.c_int32),
                ("n_blocks", C.c_int32), ("reserved", C.c_int32), ("blocks", C.POINTER(qd_jk_block))]


QD_STEP_CB = C.CFUNCTYPE(C.c_int, C.c_int32, _dp, C.c_void_p)


def _ptr(a):
    if a is None:
        return _dp()
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "arrays must be contiguous float64"
    return a.ctypes.data_as(_dp)


class ProblemStruct:
    """A qd_problem pointing into the arrays of a Problem (kept alive here)."""

    def __init__(self, p, total_steps=None, psi=None, crp=None):
        self.keep = [np.ascontiguousarray(p.r1, dtype=np.float64),
                     np.ascontiguousarray(p.r2, dtype=np.float64),
                     np.ascontiguousarray(p.x, dtype=np.float64),
                     np.ascontiguousarray(p.w, dtype=np.float64),
                     p.legendre_colmajor(),
                     np.ascontiguousarray(p.pot, dtype=np.float64),
                     None if p.dump1 is None else np.ascontiguousarray(p.dump1, dtype=np.float64),
                     None if p.dump2 is None else np.ascontiguousarray(p.dump2, dtype=np.float64),
                     np.ascontiguousarray(p.energies, dtype=np.float64),
                     np.ascontiguousarray(p.eta_sq, dtype=np.float64)]
        r1, r2, x, w, leg, pot, d1, d2, en, eta = self.keep
        self.psi = p.psi if psi is None else psi
        self.crp = p.crp if crp is None else crp
        s = qd_problem()
        s.r1 = qd_radial(p.n1, _ptr(r1), p.dr1, p.mass1)
        s.r2 = qd_radial(p.n2, _ptr(r2), p.dr2, p.mass2)
        s.theta = qd_angle(p.ntheta, p.m, _ptr(x), _ptr(w), _ptr(leg))
        s.pot = _ptr(pot)
        s.psi = _ptr(self.psi)
        s.time = qd_time(p.total_steps if total_steps is None else total_steps, p.steps, p.dt)
        s.dump1 = _ptr(d1)
        s.dump2 = _ptr(d2)
        s.crp = qd_crp(p.n_dividing_surface, p.n_gradient_points, p.energies.size,
                       p.calculate_crp, _ptr(en), _ptr(eta), _ptr(self.crp))
        self.s = s

    @property
    def steps(self):
        return self.s.time.steps


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libqd_hip.so.  There is no fallback: a missing library is an error."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError("libqd_hip.so is not built (%s); run __graft_entry__.build()" % path)
    lib = C.CDLL(path)
    P = C.POINTER
    lib.qd_time_evolution.argtypes = [P(qd_problem), P(qd_options)]
    lib.qd_create.argtypes = [P(qd_problem), P(qd_options), P(C.c_void_p)]
    lib.qd_propagate.argtypes = [C.c_void_p, C.c_int32]
    lib.qd_set_callback.argtypes = [C.c_void_p, QD_STEP_CB, C.c_void_p]
    lib.qd_get_psi.argtypes = [C.c_void_p, _dp]
    lib.qd_get_crp.argtypes = [C.c_void_p, _dp]
    lib.qd_get_norm.argtypes = [C.c_void_p, _dp]
    lib.qd_get_steps.argtypes = [C.c_void_p, P(C.c_int32)]
    lib.qd_get_timing.argtypes = [C.c_void_p, C.c_int32, _dp]
    lib.qd_destroy.argtypes = [C.c_void_p]
    lib.qd_set_psi.argtypes = [C.c_void_p, _dp]
    lib.qd_create_jk.argtypes = [P(qd_jk_problem), P(qd_options), P(C.c_void_p)]
    lib.qd_jk_theta_grid.argtypes = [C.c_int32, C.c_int32, _dp, _dp]
    lib.qd_get_crp_blocks.argtypes = [C.c_void_p, _dp]
    lib.qd_debug_stage.argtypes = [C.c_void_p, C.c_int32]
    lib.qd_debug_select.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    lib.qd_debug_get_fbr.argtypes = [C.c_void_p, _dp]
    lib.qd_debug_set_fbr.argtypes = [C.c_void_p, _dp]
    lib.qd_last_error.restype = C.c_char_p
    for f in ("qd_time_evolution", "qd_create", "qd_propagate", "qd_set_callback", "qd_get_psi",
              "qd_get_crp", "qd_get_norm", "qd_get_steps", "qd_get_timing", "qd_destroy", "qd_set_psi",
              "qd_create_jk", "qd_jk_theta_grid", "qd_get_crp_blocks", "qd_debug_stage", "qd_debug_get_fbr",
              "qd_debug_set_fbr", "qd_debug_select"):
        getattr(lib, f).restype = C.c_int
    _lib = lib
    return lib


class QDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("qd error %d: %s" % (code, msg))
        self.code = code


def check(code):
    if code != QD_OK:
        msg = _lib.qd_last_error().decode() if _lib is not None else ""
        raise QDError(code, msg)
