"""Host-side mirror of the reference's propagator entry points.

``TimeEvolutionMex(data)`` and ``TimeEvolutionMexCUDA(data)`` take the same
struct the MATLAB drivers build (``H3/main.m:122-140``): a dict with fields
``r1, r2, theta, pot, psi, time, options, dump1, dump2, CRP`` carrying the
same sub-fields (``common/TimeEvolutionMex.C:30-68``).  As in the reference
they return nothing and mutate ``psi``, ``time['steps']`` and ``CRP['CRP']``
in place.  Both run on the GPU through libqd_hip.so; the second selects the
CUDA propagator's semantics (dump before the step, 0-based surface index,
Zdotc sign; common-cuda/evolutionCUDA2.cu:487-552,898-994).

``Propagator`` is the handle API (``qd_create`` / ``qd_propagate``): it keeps
the flux accumulators across calls.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _abi
from .problem import Problem

PRECISIONS = {"fp64": _abi.QD_PREC_FP64, "mixed": _abi.QD_PREC_MIXED}
PIPELINES = {"fused": _abi.QD_PIPE_FUSED, "plain": _abi.QD_PIPE_PLAIN}
VARIANTS = {"cpu": _abi.QD_VARIANT_CPU, "cuda": _abi.QD_VARIANT_CUDA}


def make_options(precision="fp64", pipeline="fused", variant="cpu", flux_block=0, crp_every=0,
                 device=0, profile=0, theta_split=0, use_graph=0, autotune=0, fuse_theta=True,
                 diagnostics_every=0):
    return _abi.qd_options(PRECISIONS[precision], PIPELINES[pipeline], VARIANTS[variant], flux_block,
                           crp_every, use_graph, device, profile, theta_split, autotune,
                           0 if fuse_theta else -1, diagnostics_every)


def _field(d, name, required=True):
    if name not in d:
        if required:
            # the reference dereferences a missing field (matutils.h:23-31); here it is an error
            raise KeyError("missing struct field %r" % name)
        return None
    return d[name]


def _f64(a):
    a = np.asarray(a)
    if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
        raise TypeError("arrays must be contiguous float64")
    return a


def problem_from_struct(data: dict) -> Problem:
    """Read the reference struct into a Problem that shares psi and CRP.CRP."""
    r1, r2, th = _field(data, "r1"), _field(data, "r2"), _field(data, "theta")
    tm, crp = _field(data, "time"), _field(data, "CRP")
    d1 = _field(data, "dump1", required=False) or {}
    d2 = _field(data, "dump2", required=False) or {}
    n = int(th["n"])
    m = int(th["m"])
    leg = np.asarray(th["legendre"], dtype=np.float64)
    if leg.size != (m + 1) * n:
        raise ValueError("theta.legendre must hold (m+1) x n values")
    # legendre is (m+1) x n (MATLAB layout, l fastest); keep it as the (m+1, n) matrix
    leg = leg.reshape((m + 1, n), order="F") if leg.ndim == 1 else leg.reshape((m + 1, n))
    psi = data["psi"]
    if psi.dtype != np.float64 or not psi.flags["C_CONTIGUOUS"] and not psi.flags["F_CONTIGUOUS"]:
        raise TypeError("psi must be a contiguous float64 array (interleaved re/im)")
    crp_out = crp["CRP"]
    return Problem(
        name="struct", r1=_f64(r1["r"]).reshape(-1), r2=_f64(r2["r"]).reshape(-1), dr1=float(r1["dr"]),
        dr2=float(r2["dr"]), mass1=float(r1["mass"]), mass2=float(r2["mass"]),
        x=_f64(th["x"]).reshape(-1), w=_f64(th["w"]).reshape(-1), m=m, legendre=np.ascontiguousarray(leg),
        pot=np.ascontiguousarray(np.asarray(data["pot"], dtype=np.float64).reshape(-1, order="F")),
        psi=psi.reshape(-1, order="A"), dt=float(tm["time_step"]), total_steps=int(tm["total_steps"]),
        steps=int(tm["steps"]),
        dump1=None if d1.get("dump") is None else _f64(d1["dump"]).reshape(-1),
        dump2=None if d2.get("dump") is None else _f64(d2["dump"]).reshape(-1),
        n_dividing_surface=int(crp["n_dividing_surface"]), n_gradient_points=int(crp["n_gradient_points"]),
        energies=_f64(crp["energies"]).reshape(-1), eta_sq=_f64(crp["eta_sq"]).reshape(-1),
        crp=crp_out.reshape(-1), calculate_crp=int(crp.get("calculate_CRP", 1)))


def _run_struct(data: dict, variant: str, **opts):
    lib = _abi.load_library()
    p = problem_from_struct(data)
    st = _abi.ProblemStruct(p)
    o = make_options(variant=variant, **opts)
    _abi.check(lib.qd_time_evolution(C.byref(st.s), C.byref(o)))
    data["time"]["steps"] = int(st.s.time.steps)


def TimeEvolutionMex(data: dict, **opts) -> None:
    """GPU drop-in for the CPU propagator TimeEvolutionMex (common/TimeEvolutionMex.C:18-78)."""
    _run_struct(data, "cpu", **opts)


def TimeEvolutionMexCUDA(data: dict, **opts) -> None:
    """GPU drop-in for TimeEvolutionMexCUDA (common-cuda/TimeEvolutionMexCUDA.C:20-83)."""
    _run_struct(data, "cuda", **opts)


def problem_to_struct(p: Problem) -> dict:
    """The MATLAB-style struct for a Problem (fields as H3/main.m builds them)."""
    return {
        "r1": {"n": p.n1, "r": p.r1, "dr": p.dr1, "mass": p.mass1},
        "r2": {"n": p.n2, "r": p.r2, "dr": p.dr2, "mass": p.mass2},
        "theta": {"n": p.ntheta, "m": p.m, "x": p.x, "w": p.w, "legendre": p.legendre},
        "pot": p.pot,
        "psi": p.psi,
        "time": {"total_steps": p.total_steps, "time_step": p.dt, "steps": p.steps},
        "options": {},
        "dump1": {"dump": p.dump1},
        "dump2": {"dump": p.dump2},
        "CRP": {"n_dividing_surface": p.n_dividing_surface, "n_gradient_points": p.n_gradient_points,
                "n_energies": p.energies.size, "calculate_CRP": p.calculate_crp,
                "energies": p.energies, "eta_sq": p.eta_sq, "CRP": p.crp},
    }


class Propagator:
    """Handle API over qd_create / qd_propagate / qd_get_* / qd_destroy."""

    def __init__(self, problem: Problem, precision="fp64", pipeline="fused", variant="cpu",
                 flux_block=0, crp_every=0, device=0, profile=0, theta_split=0, use_graph=0, autotune=0,
                 fuse_theta=True, diagnostics_every=0):
        self.lib = _abi.load_library()
        self.problem = problem
        self._st = _abi.ProblemStruct(problem)
        self._opt = make_options(precision, pipeline, variant, flux_block, crp_every, device, profile,
                                 theta_split, use_graph, autotune, fuse_theta, diagnostics_every)
        h = C.c_void_p()
        _abi.check(self.lib.qd_create(C.byref(self._st.s), C.byref(self._opt), C.byref(h)))
        self.h = h
        self._cb = None

    def propagate(self, n_steps: int):
        _abi.check(self.lib.qd_propagate(self.h, int(n_steps)))

    def set_callback(self, fn):
        """fn(steps, crp_or_None) -> bool (True stops)."""
        ne = self.problem.energies.size

        def tramp(steps, crp_ptr, user):
            crp = None
            if crp_ptr:
                crp = np.ctypeslib.as_array(crp_ptr, shape=(ne,)).copy()
            return 1 if fn(int(steps), crp) else 0

        self._cb = _abi.QD_STEP_CB(tramp)
        _abi.check(self.lib.qd_set_callback(self.h, self._cb, None))

    def psi(self) -> np.ndarray:
        out = np.empty(2 * self.problem.grid_points)
        _abi.check(self.lib.qd_get_psi(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def crp(self) -> np.ndarray:
        out = np.empty(self.problem.energies.size)
        _abi.check(self.lib.qd_get_crp(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def norm(self) -> float:
        v = C.c_double()
        _abi.check(self.lib.qd_get_norm(self.h, C.byref(v)))
        return v.value

    def diagnostics(self) -> dict:
        """The reference's per-step diagnostics of the last step that evaluated
        them (options diagnostics_every; qd_get_diagnostics)."""
        v = (C.c_double * 5)()
        _abi.check(self.lib.qd_get_diagnostics(self.h, v))
        return {"steps": int(v[0]), "module": v[1], "e_kin": v[2], "e_rot": v[3], "e_pot": v[4],
                "e_tot": v[2] + v[3] + v[4]}

    def checkpoint(self) -> np.ndarray:
        """The whole propagation state (qd_checkpoint_save) as a byte array."""
        n = C.c_int64()
        _abi.check(self.lib.qd_checkpoint_size(self.h, C.byref(n)))
        buf = np.empty(n.value, dtype=np.uint8)
        _abi.check(self.lib.qd_checkpoint_save(self.h, buf.ctypes.data_as(C.c_void_p), C.c_int64(n.value)))
        return buf

    def restore(self, buf: np.ndarray):
        """Continue from a checkpoint of a handle of the same problem and options."""
        b = np.ascontiguousarray(buf, dtype=np.uint8)
        _abi.check(self.lib.qd_checkpoint_load(self.h, b.ctypes.data_as(C.c_void_p), C.c_int64(b.size)))

    def steps(self) -> int:
        v = C.c_int32()
        _abi.check(self.lib.qd_get_steps(self.h, C.byref(v)))
        return v.value

    def tuning(self) -> dict:
        """Kernel configuration in use (qd_get_tuning)."""
        v = (C.c_int32 * 15)()
        _abi.check(self.lib.qd_get_tuning(self.h, v))
        keys = ("autotuned", "theta_ksteps", "theta_tiles_per_group", "rpass_register_fft",
                "rline_register_fft", "rpass_lines_per_group", "mixing_bundle_lines", "rline_lines_per_group",
                "rpass_prefetch_mix_bundle_lines", "rpass_prefetch_mix_group_bucket", "fused_theta",
                "fused_theta_tiles_per_group", "rpass_prefetch_mix_two_groups", "fused_theta_threads",
                "rpass_split_mix")
        return dict(zip(keys, list(v)))

    def timing(self, which: int) -> float:
        v = C.c_double()
        _abi.check(self.lib.qd_get_timing(self.h, int(which), C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self.lib.qd_destroy(self.h)
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


def run(problem: Problem, steps: Optional[int] = None, **opts):
    """Propagate a copy of problem.psi for `steps` steps; returns (psi, crp)."""
    p = problem.copy()
    if steps is not None:
        p.total_steps = steps
    with Propagator(p, **opts) as prop:
        prop.propagate(p.total_steps)
        return prop.psi(), (prop.crp() if p.calculate_crp else None)


def _debug_methods(cls):
    def debug_stage(self, stage: int):
        _abi.check(self.lib.qd_debug_stage(self.h, int(stage)))

    def get_fbr(self) -> np.ndarray:
        out = np.empty(2 * self.n_fbr_points)
        _abi.check(self.lib.qd_debug_get_fbr(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def set_fbr(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == 2 * self.n_fbr_points
        _abi.check(self.lib.qd_debug_set_fbr(self.h, a.ctypes.data_as(C.POINTER(C.c_double))))

    def set_psi(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == 2 * self.n_points
        _abi.check(self.lib.qd_set_psi(self.h, a.ctypes.data_as(C.POINTER(C.c_double))))

    def debug_select(self, what: int, value: int):
        _abi.check(self.lib.qd_debug_select(self.h, int(what), int(value)))

    cls.debug_stage = debug_stage
    cls.debug_select = debug_select
    cls.get_fbr = get_fbr
    cls.set_fbr = set_fbr
    cls.set_psi = set_psi
    return cls


_debug_methods(Propagator)
Propagator.n_points = property(lambda self: self.problem.grid_points)
Propagator.n_fbr_points = property(lambda self: self.problem.n1 * self.problem.n2 * (self.problem.m + 1))


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_radial(n: int, lo: float, hi: float):
    """(r, dr) of qd_make_radial: linspace(lo, hi, n)."""
    lib = _abi.load_library()
    r = np.empty(int(n))
    dr = C.c_double()
    _abi.check(lib.qd_make_radial(int(n), C.c_double(lo), C.c_double(hi), _dptr(r), C.byref(dr)))
    return r, dr.value


def make_theta(n: int, m: int):
    """(x, w, legendre) of qd_make_theta; legendre as the (m+1, n) matrix P_l(x_k)."""
    lib = _abi.load_library()
    x, w = np.empty(n), np.empty(n)
    leg = np.empty((n, m + 1))          # column k holds P_0..P_m at x_k (l fastest)
    _abi.check(lib.qd_make_theta(int(n), int(m), _dptr(x), _dptr(w), _dptr(leg)))
    return x, w, np.ascontiguousarray(leg.T)


def dividing_surface(rd: float, rmin: float, dr: float) -> int:
    lib = _abi.load_library()
    v = C.c_int32()
    _abi.check(lib.qd_dividing_surface(C.c_double(rd), C.c_double(rmin), C.c_double(dr), C.byref(v)))
    return int(v.value)


def init_packet(R, r, mass1, mass2, vdiat, x=None, R0=8.0, delta=0.3, k0=0.0, v0=0, j0=0, J=0, K=0,
                energies=None, want_psi=True):
    """qd_init_packet: (psi interleaved or None, e_v, eta_sq or None) for the
    packet g(R) phi_{v0 j0}(r) Theta(x) on the grids R, r, x."""
    lib = _abi.load_library()
    R = np.ascontiguousarray(R, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    vd = np.ascontiguousarray(vdiat, dtype=np.float64)
    r1 = _abi.qd_radial(R.size, _dptr(R), float(R[1] - R[0]), float(mass1))
    r2 = _abi.qd_radial(r.size, _dptr(r), float(r[1] - r[0]), float(mass2))
    th = _abi.qd_angle()
    xs = None
    if x is not None:
        xs = np.ascontiguousarray(x, dtype=np.float64)
        th.n = xs.size
        th.x = _dptr(xs)
    pk = _abi.qd_packet(float(R0), float(delta), float(k0), int(v0), int(j0), int(J), int(K), _dptr(vd))
    E = np.ascontiguousarray(energies if energies is not None else np.zeros(0), dtype=np.float64)
    eta = np.empty(E.size)
    psi = np.empty(2 * R.size * r.size * xs.size) if (want_psi and xs is not None) else None
    ev = C.c_double()
    _abi.check(lib.qd_init_packet(C.byref(r1), C.byref(r2), C.byref(th), C.byref(pk), int(E.size),
                                  _dptr(E) if E.size else None, _dptr(psi) if psi is not None else None,
                                  C.byref(ev), _dptr(eta) if E.size else None))
    return psi, ev.value, (eta if E.size else None)


def jk_theta_grid(K: int, jmax: int):
    """The library's theta points and weights of block K (qd_jk_theta_grid)."""
    lib = _abi.load_library()
    n = jmax - K + 1
    x = np.empty(n)
    w = np.empty(n)
    _abi.check(lib.qd_jk_theta_grid(int(K), int(jmax), x.ctypes.data_as(C.POINTER(C.c_double)),
                                    w.ctypes.data_as(C.POINTER(C.c_double))))
    return x, w


@_debug_methods
class JKPropagator(Propagator):
    """Handle over a batch of (J, K) blocks (qd_create_jk)."""

    def __init__(self, jk, precision="fp64", pipeline="fused", variant="cpu", flux_block=0, crp_every=0,
                 device=0, profile=0, theta_split=0, use_graph=0, autotune=0, fuse_theta=True,
                 eta="partial_wave", diagnostics_every=0):
        """eta: the |eta(E)|^2 that normalises each J's flux (qd_jk_block.eta_sq):
        "partial_wave" (default) the packet's projection onto the incoming wave
        of partial wave J (JKProblem.eta_sq_J), "wkb" its leading WKB term, or
        "plane" the J = 0 plane-wave |eta|^2 of base.eta_sq for every J."""
        if eta not in ("plane", "partial_wave", "wkb"):
            raise ValueError("eta must be 'plane', 'partial_wave' or 'wkb'")
        self.lib = _abi.load_library()
        self.jk = jk
        self.eta = eta
        base = jk.base
        self.problem = base
        self._st = _abi.ProblemStruct(base)
        self._psi_blocks = []
        blocks = (_abi.qd_jk_block * len(jk.blocks))()
        n12 = base.n1 * base.n2
        psi = np.ascontiguousarray(jk.psi)
        self._eta = {}
        if eta != "plane" and base.calculate_crp:
            for J in sorted({b.J for b in jk.blocks}):
                self._eta[J] = np.ascontiguousarray(jk.eta_sq_J(J, kind=eta), dtype=np.float64)
        for i, b in enumerate(jk.blocks):
            # a view (the library copies psi to the device while creating the handle)
            z = psi[b.offset:b.offset + n12 * b.n].view(np.float64)
            self._psi_blocks.append(z)
            blocks[i].J = b.J
            blocks[i].K = b.K
            blocks[i].pot = jk.pot[b.K].ctypes.data_as(C.POINTER(C.c_double))
            blocks[i].psi = z.ctypes.data_as(C.POINTER(C.c_double))
            if b.J in self._eta:
                blocks[i].eta_sq = self._eta[b.J].ctypes.data_as(C.POINTER(C.c_double))
        self._blocks = blocks
        q = _abi.qd_jk_problem()
        q.base = self._st.s
        q.jmax = jk.jmax
        q.coriolis = 1 if jk.coriolis else 0
        q.n_blocks = len(jk.blocks)
        q.blocks = C.cast(blocks, C.POINTER(_abi.qd_jk_block))
        self._q = q
        self._opt = make_options(precision, pipeline, variant, flux_block, crp_every, device, profile, theta_split,
                                 use_graph, autotune, fuse_theta, diagnostics_every)
        h = C.c_void_p()
        _abi.check(self.lib.qd_create_jk(C.byref(q), C.byref(self._opt), C.byref(h)))
        self.h = h
        self._cb = None

    @property
    def n_points(self):
        return int(self.jk.psi.size)

    @property
    def n_fbr_points(self):
        return int(self.jk.psi.size)

    def psi(self) -> np.ndarray:
        out = np.empty(2 * self.n_points)
        _abi.check(self.lib.qd_get_psi(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def crp_blocks(self) -> np.ndarray:
        out = np.empty((len(self.jk.blocks), self.problem.energies.size))
        _abi.check(self.lib.qd_get_crp_blocks(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def crp_by_J(self) -> dict:
        """P^J(E) = sum over the K blocks of J, each block's flux normalised by
        the |eta(E)|^2 chosen at creation (`eta`)."""
        pb = self.crp_blocks()
        out = {}
        for i, b in enumerate(self.jk.blocks):
            out[b.J] = out.get(b.J, 0) + pb[i]
        return out
