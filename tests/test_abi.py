"""The C ABI (include/qd.h) without a GPU: the library loads, exports every
declared entry point, and rejects bad input with the documented error codes
before any device work (the reference aborts through insist()/MCrash instead:
fftwinterface.C:112, Gradient3.F:25, evolutionCUDA2.cu:915)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "qd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qd_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol(qd):
    lib = qd._abi.load_library()
    names = declared_symbols()
    assert "qd_time_evolution" in names and "qd_create" in names and len(names) >= 11
    for n in names:
        assert hasattr(lib, n), n


def test_abi_struct_layout_matches_header(qd):
    abi = qd._abi
    # x86-64 natural alignment of the structs declared in qd.h
    assert C.sizeof(abi.qd_radial) == 32
    assert C.sizeof(abi.qd_angle) == 32
    assert C.sizeof(abi.qd_time) == 16
    assert C.sizeof(abi.qd_crp) == 40
    assert C.sizeof(abi.qd_options) == 48
    assert C.sizeof(abi.qd_jk_block) == 32
    assert C.sizeof(abi.qd_problem) == 32 + 32 + 32 + 8 + 8 + 16 + 8 + 8 + 40


def _create(qd, p, **opt):
    lib = qd._abi.load_library()
    st = qd._abi.ProblemStruct(p)
    o = qd.propagator.make_options(**opt)
    h = C.c_void_p()
    rc = lib.qd_create(C.byref(st.s), C.byref(o), C.byref(h))
    return rc, lib.qd_last_error().decode()


@pytest.fixture()
def small(qd):
    return qd.problem.config(0, nR=32, nr=24, ntheta=6, steps=2)


def test_odd_grid_rejected(qd, small):
    p = small.copy()
    p.r1 = p.r1[:-1]
    rc, msg = _create(qd, p)
    assert rc == 2 and "even" in msg


def test_missing_psi_rejected(qd, small):
    lib = qd._abi.load_library()
    st = qd._abi.ProblemStruct(small)
    st.s.psi = C.POINTER(C.c_double)()
    h = C.c_void_p()
    assert lib.qd_create(C.byref(st.s), None, C.byref(h)) == 1


@pytest.mark.parametrize("npoints", [4, 1, 63])
def test_bad_stencil_rejected(qd, small, npoints):
    p = small.copy()
    p.n_gradient_points = npoints
    rc, msg = _create(qd, p)
    assert rc == 3 and "NPoints" in msg


def test_surface_outside_grid_rejected(qd, small):
    p = small.copy()
    p.n_dividing_surface = p.n2 - 2
    rc, _ = _create(qd, p)
    assert rc == 4


def test_fft_length_with_large_prime_rejected(qd):
    p = qd.problem.config(0, nR=14, nr=24, ntheta=6, steps=2)
    rc, msg = _create(qd, p)
    assert rc == 7 and "prime" in msg


def test_bad_options_rejected(qd, small):
    rc, _ = _create(qd, small, theta_split=5)
    assert rc == 7


def test_struct_mirror_reads_reference_fields(qd, small):
    """TimeEvolutionMex(struct) reads the same fields as common/TimeEvolutionMex.C:30-68."""
    d = qd.propagator.problem_to_struct(small)
    p = qd.propagator.problem_from_struct(d)
    assert p.n1 == small.n1 and p.n2 == small.n2 and p.ntheta == small.ntheta
    assert np.shares_memory(p.psi, small.psi) and np.shares_memory(p.crp, small.crp)
    del d["r1"]
    with pytest.raises(KeyError):
        qd.propagator.problem_from_struct(d)
