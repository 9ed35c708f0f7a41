"""The reference's per-step diagnostics on the GPU (options.diagnostics_every,
qd_get_diagnostics): module, e_kin, e_rot and e_pot evaluated at the points
of the step where TimeEvolution::evolution_dt and time_evolution evaluate them
(common/timeEvol.C:619-647,671; module_for_psi :220-243, potential_energy
:256-283, rotational_energy :323-360, kinetic_energy_for_legendre_psi
:390-436), against the oracle's diagnostics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NAMES = ("module", "e_kin", "e_rot", "e_pot")


def oracle_diag(oracle, p, steps):
    _, _, _, d = oracle.time_evolution(p, total_steps=steps, diagnostics=1)
    return {"module": d.module, "e_kin": d.e_kin, "e_rot": d.e_rot, "e_pot": d.e_pot}


def test_fp64_diagnostics_match_oracle(qd, oracle):
    p = qd.problem.config(0)
    print("case: fp64 diagnostics cfg0 steps 1, 2, 7, 30", flush=True)
    got = {}
    with qd.Propagator(p.copy(), precision="fp64", diagnostics_every=1, use_graph=1) as prop:
        done = 0
        for k in (1, 2, 7, 30):
            prop.propagate(k - done)
            done = k
            got[k] = prop.diagnostics()
        psi, crp = prop.psi(), prop.crp()
    for k, g in got.items():
        ref = oracle_diag(oracle, p, k)
        scale = ref["e_kin"] + ref["e_rot"] + ref["e_pot"]
        assert g["steps"] == k
        assert abs(g["module"] - ref["module"]) < 1e-10 * ref["module"], (k, g, ref)
        for n in ("e_kin", "e_rot", "e_pot"):
            assert abs(g[n] - ref[n]) < 1e-10 * scale, (k, n, g[n], ref[n])
    # the diagnostic steps propagate exactly like the others
    psi_o, crp_o, _, _ = oracle.time_evolution(p, total_steps=30)
    assert np.max(np.abs(psi - psi_o)) < 1e-10 * np.max(np.abs(psi_o))
    assert np.max(np.abs(crp - crp_o)) < 1e-8 * max(np.max(np.abs(crp_o)), 1e-2)


def test_mixed_fused_theta_with_diagnostics(qd, oracle):
    """A mixed-precision handle on the fused theta step: every 5th step runs the
    operator-by-operator step with the reductions and returns to the fused
    state; P(E) and psi stay within the mixed path's rounding of the run
    without diagnostics, and the diagnostics within 1e-5 of the oracle's."""
    p = qd.problem.config(0)
    print("case: mixed diagnostics every 5 steps, 40 steps", flush=True)
    with qd.Propagator(p.copy(), precision="mixed") as a:
        assert a.tuning()["fused_theta"] >= 1
        a.propagate(40)
        psi_a, crp_a = a.psi(), a.crp()
    with qd.Propagator(p.copy(), precision="mixed", diagnostics_every=5) as b:
        b.propagate(40)
        d = b.diagnostics()
        psi_b, crp_b = b.psi(), b.crp()
    assert d["steps"] == 40
    ref = oracle_diag(oracle, p, 40)
    scale = ref["e_kin"] + ref["e_rot"] + ref["e_pot"]
    for n in NAMES:
        assert abs(d[n] - ref[n]) < 1e-5 * (1.0 if n == "module" else scale), (n, d[n], ref[n])
    assert np.max(np.abs(psi_b - psi_a)) < 1e-5 * np.max(np.abs(psi_a))
    assert np.max(np.abs(crp_b - crp_a)) < 1e-5


def test_cuda_order_fused_theta_with_diagnostics(qd):
    """The fused theta step in the CUDA propagator's order, whose FBR state
    holds the dumped psi: with diagnostics every 5 steps (each needs psi from
    the step before it) and a P(E) callback every 7 steps, psi and P(E) stay
    within the mixed path's rounding of the run without them, and the
    diagnostics within 1e-5 of the unfused handle's."""
    p = qd.problem.config(0)
    print("case: mixed cuda-order diagnostics every 5 steps, 40 steps", flush=True)
    with qd.Propagator(p.copy(), precision="mixed", variant="cuda") as a:
        assert a.tuning()["fused_theta"] >= 1
        a.propagate(40)
        psi_a, crp_a = a.psi(), a.crp()
    out = {}
    for fuse in (True, False):
        with qd.Propagator(p.copy(), precision="mixed", variant="cuda", diagnostics_every=5, crp_every=7,
                           fuse_theta=fuse) as b:
            assert (b.tuning()["fused_theta"] >= 1) == bool(fuse)
            seen = []
            b.set_callback(lambda steps, crp: seen.append(steps) or False)
            b.propagate(23)
            d_mid = b.diagnostics()
            b.propagate(17)
            out[fuse] = (b.diagnostics(), d_mid, b.psi(), b.crp(), seen)
    d_f, dm_f, psi_f, crp_f, seen_f = out[True]
    d_u, dm_u, psi_u, crp_u, _ = out[False]
    assert seen_f == [7, 14, 21, 28, 35]
    assert d_f["steps"] == 40 and dm_f["steps"] == 20
    for x, y in ((d_f, d_u), (dm_f, dm_u)):
        scale = abs(y["e_kin"]) + abs(y["e_rot"]) + abs(y["e_pot"])
        for n in NAMES:
            assert abs(x[n] - y[n]) < 1e-5 * (y["module"] if n == "module" else scale), (n, x[n], y[n])
    assert np.max(np.abs(psi_f - psi_a)) < 1e-5 * np.max(np.abs(psi_a))
    assert np.max(np.abs(psi_f - psi_u)) < 1e-5 * np.max(np.abs(psi_u))
    assert np.max(np.abs(crp_f - crp_a)) < 1e-5
    assert np.max(np.abs(crp_f - crp_u)) < 1e-5


def test_total_energy_conserved_without_absorbers(qd, oracle):
    """With the absorbers off the reference's e_tot = e_kin + e_rot + e_pot
    stays constant to the splitting error over the first 50 steps of cfg 0
    (0.2 % at dt = 10 au; later the products wrap round the periodic r grid
    into the capped region of the surface), and the norm to rounding."""
    p = qd.problem.config(0)
    p.dump1 = None
    p.dump2 = None
    print("case: e_tot conservation, 50 steps without absorbers", flush=True)
    e = []
    with qd.Propagator(p.copy(), precision="fp64", diagnostics_every=10) as prop:
        for _ in range(5):
            prop.propagate(10)
            d = prop.diagnostics()
            e.append((d["e_tot"], d["module"]))
    etot = np.array([v for v, _ in e])
    mod = np.array([m for _, m in e])
    assert np.max(np.abs(etot - etot[0])) < 3e-3 * etot[0]
    assert np.max(np.abs(mod - 1.0)) < 1e-10
    ref = oracle_diag(oracle, p, 50)
    assert abs(etot[-1] - (ref["e_kin"] + ref["e_rot"] + ref["e_pot"])) < 1e-10 * etot[-1]


def test_diagnostics_rejected_for_batches(qd):
    pb = qd.problem
    base = pb.config(0, nR=32, nr=24, ntheta=6, steps=2)
    jk = pb.make_jk_problem(base, pb.H_H2, [0, 2], jmax=5)
    with pytest.raises(qd._abi.QDError) as ei:
        qd.propagator.JKPropagator(jk, precision="fp64", diagnostics_every=1)
    assert ei.value.code == 7  # QD_ERR_ARG
