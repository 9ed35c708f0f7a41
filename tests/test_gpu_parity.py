"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars: FP64 psi and P(E) within 1e-8 relative of the oracle
(|dP_e| <= 1e-8 * max(max_e |P|, 1e-2)); the mixed path within 1e-4 of the
FP64 path on the same grid.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel_psi(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def crp_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-2))


@pytest.fixture(scope="module")
def hh2_small(qd):
    """H+H2 J=0 on 64x48x40, 2000 x 10 au."""
    return qd.problem.config(0)


@pytest.mark.parametrize("pipeline", ["fused", "plain"])
@pytest.mark.parametrize("variant", ["cpu", "cuda"])
def test_fp64_short_matches_oracle(qd, oracle, hh2_small, pipeline, variant):
    steps = 30
    print("case: hh2_small fp64", pipeline, variant, steps, flush=True)
    psi_g, crp_g = qd.run(hh2_small, steps=steps, precision="fp64", pipeline=pipeline, variant=variant)
    psi_o, crp_o, _, _ = oracle.time_evolution(hh2_small, total_steps=steps, variant=1 if variant == "cuda" else 0)
    assert rel_psi(psi_g, psi_o) < 1e-10
    assert crp_err(crp_g, crp_o) < 1e-8


def test_fp64_full_hh2_small_matches_oracle(qd, oracle, hh2_small):
    print("case: hh2_small fp64 full 2000 steps", flush=True)
    psi_g, crp_g = qd.run(hh2_small, precision="fp64")
    psi_o, crp_o, _, _ = oracle.time_evolution(hh2_small)
    print("dpsi", rel_psi(psi_g, psi_o), "dP", crp_err(crp_g, crp_o))
    assert crp_err(crp_g, crp_o) < 1e-8


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_full_FH2_192x128x80_matches_cached_oracle(qd, prec):
    """F+H2 J=0 on 192x128x80 for the whole 10000 x 5 au against the oracle's
    result cached in tests/golden (about 40 min of the CPU oracle;
    make_oracle_pe.py): FP64 P(E) within 1e-8 and the final |psi|^2 within
    1e-10; mixed P(E) within 1e-4 of it."""
    ref = np.load(os.path.join(GOLDEN, "oracle_pe_FH2_192x128x80.npz"))
    p = qd.problem.config(1)
    assert int(ref["steps"]) == p.total_steps
    print("case: F+H2 192x128x80 full", prec, flush=True)
    psi, crp = qd.run(p, precision=prec)
    z = psi.view(np.complex128)
    a2 = float(np.sum(np.abs(z) ** 2))
    print("dP", np.max(np.abs(crp - ref["crp"])), "d|psi|^2", abs(a2 - float(ref["psi_abs2"])) / float(ref["psi_abs2"]))
    assert float(np.max(ref["crp"])) > 0.5  # a non-trivial P(E)
    if prec == "fp64":
        assert crp_err(crp, ref["crp"]) < 1e-8
        assert abs(a2 - float(ref["psi_abs2"])) < 1e-10 * float(ref["psi_abs2"])
    else:
        assert np.max(np.abs(crp - ref["crp"])) < 1e-4


@pytest.mark.parametrize("split", [2, 3])
def test_mixed_short_close_to_fp64(qd, hh2_small, split):
    print("case: hh2_small mixed split", split, flush=True)
    psi_d, crp_d = qd.run(hh2_small, steps=200, precision="fp64")
    psi_m, crp_m = qd.run(hh2_small, steps=200, precision="mixed", theta_split=split)
    print("dpsi", rel_psi(psi_m, psi_d), "dP", np.max(np.abs(crp_m - crp_d)))
    assert rel_psi(psi_m, psi_d) < 1e-3
    assert np.max(np.abs(crp_m - crp_d)) < 1e-3


def test_handle_resume_and_norm(qd, hh2_small):
    p = hh2_small.copy()
    p.dump1 = None
    p.dump2 = None
    with qd.Propagator(p, precision="fp64") as prop:
        n0 = prop.norm()
        prop.propagate(10)
        prop.propagate(15)
        n1 = prop.norm()
        assert prop.steps() == 25
    assert abs(n1 - n0) < 1e-12
    psi_a, crp_a = qd.run(p, steps=25, precision="fp64")
    with qd.Propagator(p, precision="fp64") as prop:
        prop.propagate(10)
        prop.propagate(15)
        assert rel_psi(prop.psi(), psi_a) < 1e-14
        assert crp_err(prop.crp(), crp_a) < 1e-12


@pytest.mark.parametrize("prec,variant", [("fp64", "cpu"), ("mixed", "cpu"), ("mixed", "cuda")])
@pytest.mark.parametrize("crp_every", [0, 20])
def test_graph_replay_equals_eager(qd, hh2_small, prec, variant, crp_every):
    """Whole flux blocks replayed as one captured graph give the same psi and
    P(E) as the same kernels launched one by one.  Three propagate calls with
    host work between them replay the graph several times; with crp_every = 20
    some blocks hold a P(E) point and run eagerly, so both kinds interleave.
    Mixed precision runs the fused theta step, in either propagator order."""
    print("case: graph replay", prec, variant, "crp_every", crp_every, flush=True)
    out = []
    for graph in (0, 1):
        with qd.Propagator(hh2_small.copy(), precision=prec, variant=variant, flux_block=16, crp_every=crp_every,
                           use_graph=graph) as prop:
            seen = []
            for n in (40, 33, 55):
                prop.propagate(n)
                seen.append(prop.crp().copy())       # eager work and a host sync between replays
            out.append((prop.psi(), seen, prop.steps()))
    (pa, ca, sa), (pb, cb, sb) = out
    assert sa == sb == 128
    assert np.array_equal(pa, pb)
    for a, b in zip(ca, cb):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_autotuned_handle_matches(qd, oracle, hh2_small, prec):
    """The autotuner only picks among equivalent kernels: an autotuned FP64
    handle still matches the oracle, an autotuned mixed one the default."""
    print("case: autotune", prec, flush=True)
    steps = 40
    with qd.Propagator(hh2_small.copy(), precision=prec, autotune=1) as prop:
        tun = prop.tuning()
        prop.propagate(steps)
        psi_t, crp_t = prop.psi(), prop.crp()
    print("tuning", tun)
    assert tun["autotuned"] == 1
    if prec == "fp64":
        psi_o, crp_o, _, _ = oracle.time_evolution(hh2_small, total_steps=steps)
        assert rel_psi(psi_t, psi_o) < 1e-10
        assert crp_err(crp_t, crp_o) < 1e-8
    else:
        psi_d, crp_d = qd.run(hh2_small, steps=steps, precision="mixed")
        assert rel_psi(psi_t, psi_d) < 1e-5
        assert np.max(np.abs(crp_t - crp_d)) < 1e-5


@pytest.mark.parametrize("variant", ["cpu", "cuda"])
@pytest.mark.parametrize("kind", ["single", "batched", "wide"])
def test_fused_theta_step_matches_unfused(qd, kind, variant):
    """The fused theta step (theta-backward, dump, flux rows, V phase and the
    next theta-forward in one kernel; the state between steps is the FBR
    array) against the unfused mixed step, across a resume with a norm read in
    between: psi (rebuilt from the FBR state, or in the CUDA propagator's
    order written by the step before the read) and P(E) agree to the rounding
    level of the mixed path, and are no further from the FP64 path than the
    unfused ones.  Each fused kernel is checked: the dense transforms, and
    their parity halves with 512 and 768 threads per work group; on a
    101-point theta grid (half bases of 51 and an odd middle node) the parity
    halves with two k-steps per half, one wave per tile or a pair of waves
    (one per component) per tile."""
    pb = qd.problem
    print("case: fused theta vs unfused,", kind, variant, flush=True)
    kernels = (1, 3) if kind == "wide" else (0, 1, 2)
    if kind in ("single", "wide"):
        prob = pb.config(0, steps=60) if kind == "single" else pb.config(0, ntheta=101, steps=60)
        mk = lambda fuse: qd.Propagator(prob.copy(), precision="mixed", variant=variant, fuse_theta=fuse)
    else:
        base = pb.config(0, nR=64, nr=32, ntheta=12, steps=60, e0_ev=0.6)
        jk = pb.make_jk_problem(base, pb.H_H2, [0, 3, 7], jmax=11, R0=8.0, delta=0.3, e0_ev=0.6)
        mk = lambda fuse: qd.propagator.JKPropagator(jk, precision="mixed", variant=variant, fuse_theta=fuse)
    if kind in ("single", "wide"):
        mk64 = lambda: qd.Propagator(prob.copy(), precision="fp64", variant=variant)
    else:
        mk64 = lambda: qd.propagator.JKPropagator(jk, precision="fp64", variant=variant)
    out = {}
    for fuse in (False,) + kernels + ("fp64",):
        with (mk64() if fuse == "fp64" else mk(fuse is not False)) as prop:
            if fuse is False:
                assert prop.tuning()["fused_theta"] == 0
            elif fuse != "fp64":
                # parity halves by default, as wave pairs for half bases of 33..64
                assert prop.tuning()["fused_theta"] == (3 if kind == "wide" else 2)
                prop.debug_select(3, fuse)
                assert prop.tuning()["fused_theta"] == {0: 1, 1: 2, 2: 2, 3: 3}[fuse]
            prop.propagate(25)
            n_mid = prop.norm()
            prop.propagate(35)
            out[fuse] = (prop.psi(), prop.crp(), n_mid, prop.norm())
    psi_u, crp_u, nm_u, n_u = out[False]
    psi_d, crp_d = out["fp64"][0], out["fp64"][1]
    nm_d, n_d = out["fp64"][2], out["fp64"][3]
    scale = np.max(np.abs(psi_d))
    cs = max(np.max(np.abs(crp_d)), 1e-2)
    for kernel in kernels:
        psi_f, crp_f, nm_f, n_f = out[kernel]
        d_fu, d_ud, d_fd = (np.max(np.abs(a - b)) / scale for a, b in ((psi_f, psi_u), (psi_u, psi_d), (psi_f, psi_d)))
        print("fused kernel %d psi: |fused - unfused| %.2e, |unfused - fp64| %.2e, |fused - fp64| %.2e"
              % (kernel, d_fu, d_ud, d_fd))
        assert d_fu <= 1e-5 and d_fd <= 2 * d_ud + 1e-6
        assert np.max(np.abs(crp_f - crp_u)) <= 1e-6 * cs
        assert np.max(np.abs(crp_f - crp_d)) <= 2 * np.max(np.abs(crp_u - crp_d)) + 1e-7 * cs
        print("norm after 25 / 60 steps: fp64 %.9f %.9f, unfused %.9f %.9f, fused %.9f %.9f"
              % (nm_d, n_d, nm_u, n_u, nm_f, n_f))
        assert abs(nm_f - nm_d) <= 2 * abs(nm_u - nm_d) + 1e-6 * nm_d
        assert abs(n_f - n_d) <= 2 * abs(n_u - n_d) + 1e-6 * n_d


@pytest.mark.parametrize("variant", ["cpu", "cuda"])
def test_wave_pair_meeting_through_lds_same_bits(qd, variant):
    """The wave-pair theta kernel with the pair meeting through LDS counters
    (the default; qd_debug_select what 6) against its work-group-barrier form: the same
    arithmetic, so psi and P(E) bit for bit, on a 101-point theta grid (half
    bases of 51, odd middle node) across a replayed 64-step flux block, a
    P(E) read and a resume."""
    pb = qd.problem
    prob = pb.config(0, ntheta=101, steps=80)
    out = {}
    for meet in (0, 1):
        print("case: wave-pair theta, LDS meeting", meet, variant, flush=True)
        with qd.Propagator(prob.copy(), precision="mixed", variant=variant, use_graph=1) as prop:
            assert prop.tuning()["fused_theta"] == 3
            prop.debug_select(6, meet)
            prop.propagate(70)
            c1 = prop.crp()
            prop.propagate(10)
            out[meet] = (prop.psi(), c1, prop.crp())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("variant", ["cpu", "cuda"])
@pytest.mark.parametrize("kind", ["single", "batched"])
def test_flux_extraction_in_r_pass_same_bits(qd, kind, variant):
    """The fused step's flux-surface extraction run by the next step's r pass
    (the default) against a launch of its own after each step: the same psi
    and P(E) bit for bit, across a replayed 64-step flux block (graphs on), a
    P(E) read inside a block, a resume and (single grid) a diagnostics step."""
    pb = qd.problem
    print("case: extraction in the r pass,", kind, variant, flush=True)
    if kind == "single":
        prob = pb.config(0, steps=100)
        mk = lambda: qd.Propagator(prob.copy(), precision="mixed", variant=variant, use_graph=1, diagnostics_every=50)
    else:
        base = pb.config(0, nR=64, nr=32, ntheta=12, steps=100, e0_ev=0.6)
        jk = pb.make_jk_problem(base, pb.H_H2, [0, 3, 7], jmax=11, R0=8.0, delta=0.3, e0_ev=0.6)
        mk = lambda: qd.propagator.JKPropagator(jk, precision="mixed", variant=variant, use_graph=1)
    out = []
    for deferred in (1, 0):
        with mk() as prop:
            assert prop.tuning()["fused_theta"] >= 1
            prop.debug_select(5, deferred)
            prop.propagate(7)
            c1 = prop.crp().copy()
            prop.propagate(90)
            out.append((c1, prop.crp().copy(), prop.psi().copy(), prop.norm()))
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.fixture(scope="module")
def wide_r(qd):
    """H+H2 on a finer r grid (192 points over 0.4..6 a0) with the surface at
    1.6 a0 (row 41), inside the H2 wavefunction, so that stencils up to 61
    points (30 rows each side) fit and the surface carries flux from the first
    steps (|P| ~ 0.2 after 30 steps, differing by 1 % between 3 and 31 points)."""
    return qd.problem.config(0, nr=192, r_range=(0.4, 6.0), rs=1.6, steps=30)


@pytest.mark.parametrize("npoints", [3, 5, 7, 9, 31, 61])
@pytest.mark.parametrize("variant", ["cpu", "cuda"])
def test_flux_stencils_match_oracle(qd, oracle, wide_r, npoints, variant):
    """The surface gradient with every stencil width the reference propagators
    take (3..11 points in Gradient3.F:15-26; odd 3..61 in gradients2.cu:566-569;
    31 in HO2/DMBEIV/main.m:105), in both propagators' conventions, FP64
    against the oracle."""
    p = wide_r.copy()
    p.n_gradient_points = npoints
    print("case: stencil", npoints, variant, flush=True)
    psi_g, crp_g = qd.run(p, precision="fp64", variant=variant)
    psi_o, crp_o, _, _ = oracle.time_evolution(p, variant=1 if variant == "cuda" else 0)
    assert rel_psi(psi_g, psi_o) < 1e-10
    assert crp_err(crp_g, crp_o) < 1e-8
    assert np.max(np.abs(crp_o)) > 0.1  # the surface sees flux
