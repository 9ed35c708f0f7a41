"""Batched (J, K) propagation on the GPU against the numpy oracle
(oracle/jk_oracle.py), and the invariants that pin the J > 0 extension where
the reference has nothing to compare with (it propagates J = 0 only):

* the J = 0 block of a batch reproduces the single-block J = 0 propagator;
* the norm summed over the K blocks of each J is conserved without absorbers;
* the centrifugal-sudden limit (Coriolis off) agrees with the oracle too.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def jkcase(qd):
    pb = qd.problem
    base = pb.config(0, nR=48, nr=32, ntheta=8, steps=40, e0_ev=0.6)
    return pb.make_jk_problem(base, pb.H_H2, [0, 1, 3], jmax=7, R0=8.0, delta=0.3, e0_ev=0.6)


@pytest.mark.parametrize("coriolis", [True, False])
def test_jk_fp64_matches_numpy_oracle(qd, jkcase, coriolis):
    import jk_oracle

    jk = jkcase
    jk.coriolis = coriolis
    steps = jk.base.total_steps
    print("case: jk fp64 coriolis", coriolis, "blocks", len(jk.blocks), flush=True)
    psi_o, PJ, _ = jk_oracle.propagate(jk, steps)
    with qd.propagator.JKPropagator(jk, precision="fp64", eta="plane") as prop:
        prop.propagate(steps)
        psi_g = prop.psi().view(np.complex128)
        PJg = prop.crp_by_J()
    assert rel(psi_g, psi_o) < 1e-10
    scale = max(max(np.max(np.abs(v)) for v in PJ.values()), 1e-2)
    for J in PJ:
        assert np.max(np.abs(PJg[J] - PJ[J])) < 1e-8 * scale, J
    jk.coriolis = True


def test_jk_block_zero_equals_single_block_path(qd, jkcase):
    pb = qd.problem
    jk = pb.make_jk_problem(jkcase.base, pb.H_H2, [0], jmax=jkcase.jmax, R0=8.0, delta=0.3, e0_ev=0.6)
    steps = 25
    p = jk.base.copy()
    p.x, p.w = pb.gauss_legendre(jk.jmax + 1)
    p.m = jk.jmax
    p.legendre = pb.legendre_p2(p.m, p.x)
    p.pot = jk.pot[0]
    p.psi = jk.psi.view(np.float64).copy()
    with qd.propagator.JKPropagator(jk, precision="fp64", eta="plane") as a:
        a.propagate(steps)
        pa, ca = a.psi(), a.crp()
    with qd.Propagator(p, precision="fp64") as b:
        b.propagate(steps)
        pbb, cb = b.psi(), b.crp()
    # same kernels and tables; only the Gauss-Legendre grid is rebuilt in C++
    assert rel(pa, pbb) < 1e-12
    assert np.max(np.abs(ca - cb)) < 1e-12 * max(np.max(np.abs(cb)), 1e-2)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_jk_norm_conserved_without_absorbers(qd, jkcase, prec):
    import jk_oracle

    jk = jkcase
    base = jk.base.copy()
    base.dump1 = None
    base.dump2 = None
    jk2 = qd.problem.make_jk_problem(base, qd.problem.H_H2, jk.Js, jmax=jk.jmax, R0=8.0, delta=0.3, e0_ev=0.6)
    n0 = jk_oracle.norm(jk2, jk2.psi)
    with qd.propagator.JKPropagator(jk2, precision=prec) as prop:
        prop.propagate(60)
        n1 = prop.norm()
    # mixed: the per-step bound of test_mixed_norm_drift, 5e-8; on these 5..8-point
    # theta bases the FP32 rounding alone moves the norm by about 1e-8 per step
    # (-0.5e-6 .. -1.1e-6 after 60 steps, -2e-6 .. -3e-6 after 240, for every
    # theta kernel and FFT table form measured), with no trend beyond that
    assert abs(n1 - n0) < (1e-12 if prec == "fp64" else 60 * 5e-8) * n0


def test_jk_prefetch_mixing_mixed_within_1e4_of_fp64(qd):
    """Mixed-precision batch on a 64-point R grid (prefetching FP32 mixing R
    pass, with one and with two work groups per CU) against the FP64 batch
    over 1500 steps: per-energy |dP^J(E)| within 1e-4, and the norm of a
    no-absorber run within 1e-5 of its start."""
    pb = qd.problem
    base = pb.config(0, nR=64, nr=32, ntheta=12, steps=1500, e0_ev=0.6)
    jk = pb.make_jk_problem(base, pb.H_H2, [0, 4, 9], jmax=11, R0=8.0, delta=0.3, e0_ev=0.6)
    print("case: jk mixed vs fp64, prefetch mixing, 1500 steps", flush=True)
    out = {}
    for prec in ("fp64", "mixed", "mixed_two"):
        with qd.propagator.JKPropagator(jk, precision=prec[:5] if prec != "fp64" else prec) as prop:
            if prec != "fp64":
                assert prop.tuning()["rpass_prefetch_mix_bundle_lines"] == 32
                prop.debug_select(0, 1 if prec == "mixed_two" else 0)
            prop.propagate(base.total_steps)
            out[prec] = prop.crp_by_J()
    assert max(float(np.max(out["fp64"][J])) for J in out["fp64"]) > 1e-3  # a non-trivial P
    for m in ("mixed", "mixed_two"):
        worst = max(float(np.max(np.abs(out[m][J] - out["fp64"][J]))) for J in out["fp64"])
        print("max |dP^J(E)| %s vs fp64: %.2e" % (m, worst))
        assert worst < 1e-4

    b2 = base.copy()
    b2.dump1 = None
    b2.dump2 = None
    jk2 = pb.make_jk_problem(b2, pb.H_H2, [0, 4, 9], jmax=11, R0=8.0, delta=0.3, e0_ev=0.6)
    import jk_oracle
    n0 = jk_oracle.norm(jk2, jk2.psi)
    with qd.propagator.JKPropagator(jk2, precision="mixed") as prop:
        prop.propagate(1500)
        n1 = prop.norm()
    print("norm drift over 1500 steps: %.2e" % ((n1 - n0) / n0))
    assert abs(n1 - n0) < 1e-5 * n0


@pytest.mark.parametrize("path", ["mixed-prefetch48", "mixed-split", "fp64-split"])
def test_jk_groups_above_32_lines_match_oracle(qd, path):
    """J = 33: 34 Coriolis-coupled K blocks per (J, j >= 33), against the FP64
    numpy oracle: mixed precision with the prefetching mixing R pass (48-line
    bundles) or the split pass, and FP64, whose only R pass for groups above
    32 lines is the split one."""
    import jk_oracle

    pb = qd.problem
    base = pb.config(0, nR=64, nr=32, ntheta=8, steps=12, e0_ev=0.6)
    jk = pb.make_jk_problem(base, pb.H_H2, [33], jmax=35, R0=8.0, delta=0.3, e0_ev=0.6)
    print("case: jk J=33 groups of 34 lines vs oracle,", path, flush=True)
    psi_o, PJ, _ = jk_oracle.propagate(jk, 12)
    prec = path.split("-")[0]
    with qd.propagator.JKPropagator(jk, precision=prec) as prop:
        if path == "mixed-prefetch48":
            assert prop.tuning()["rpass_prefetch_mix_bundle_lines"] == 48
        elif path == "mixed-split":
            prop.debug_select(2, 1)
        assert prop.tuning()["rpass_split_mix"] == (0 if path == "mixed-prefetch48" else 1)
        prop.propagate(12)
        psi_g = prop.psi().view(np.complex128)
    assert rel(psi_g, psi_o) < (1e-10 if prec == "fp64" else 2e-5)


def test_autotuned_batch_keeps_its_wavefunction(qd):
    """The autotuner times every R-pass candidate on the handle's arrays before
    the first step; the split pass's trials take a scratch array of their own,
    so the state the handle starts from is untouched: an autotuned mixed batch
    (on a grid where the split pass is a candidate) propagates like an
    untuned one, to the mixed path's rounding."""
    pb = qd.problem
    base = pb.config(0, nR=128, nr=16, ntheta=8, steps=30, e0_ev=0.6)
    jk = pb.make_jk_problem(base, pb.H_H2, [0, 3, 9], jmax=11, R0=8.0, delta=0.3, e0_ev=0.6)
    print("case: autotuned batch vs untuned, 128 x 16 grid", flush=True)
    out = {}
    for tune in (0, 1):
        with qd.propagator.JKPropagator(jk, precision="mixed", autotune=tune) as prop:
            prop.propagate(30)
            out[tune] = (prop.psi().view(np.complex128), prop.crp_by_J())
    assert rel(out[1][0], out[0][0]) < 1e-5
    for J in out[0][1]:
        assert np.max(np.abs(out[1][1][J] - out[0][1][J])) < 1e-5
