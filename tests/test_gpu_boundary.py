"""The drop-in entry points on the GPU, against the oracle.

* propagator.TimeEvolutionMex / TimeEvolutionMexCUDA (qd_time_evolution, the
  mexFunction of common/TimeEvolutionMex.C:18-78 and
  common-cuda/TimeEvolutionMexCUDA.C:20-83) mutate psi, time.steps and
  CRP.CRP in place.  A second call with time.steps != 0 skips the first -dt/2
  phase and starts the flux accumulators from zero, as the reference does
  (timeEvol.C:664-665, setup_CRP_data :762-805).
* qd_set_callback: called every crp_every steps with P(E) so far; a nonzero
  return stops the propagation (the options.wave_to_matlab hook,
  wave2matlab.C:27-34).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def crp_err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-2))


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("variant", ["cpu", "cuda"])
def test_time_evolution_mex_in_place_and_resume(qd, oracle, variant):
    pv = qd.problem.config(0)
    ov = 1 if variant == "cuda" else 0
    entry = qd.propagator.TimeEvolutionMexCUDA if variant == "cuda" else qd.propagator.TimeEvolutionMex
    print("case: TimeEvolutionMex%s cfg0 30 + 20 steps" % ("CUDA" if variant == "cuda" else ""), flush=True)

    # first call: 30 steps from steps = 0
    p = pv.copy()
    p.total_steps = 30
    data = qd.propagator.problem_to_struct(p)
    psi_buf, crp_buf = data["psi"], data["CRP"]["CRP"]
    psi_o, crp_o, steps_o, _ = oracle.time_evolution(p, total_steps=30, variant=ov)
    entry(data)
    assert data["psi"] is psi_buf and data["CRP"]["CRP"] is crp_buf      # the caller's arrays, in place
    assert data["time"]["steps"] == 30 == steps_o
    assert rel(psi_buf, psi_o) < 1e-10
    assert crp_err(crp_buf, crp_o) < 1e-8

    # second call: 20 more steps from steps = 30 (no half phase, accumulators restart)
    p2 = pv.copy()
    p2.psi = psi_o.copy()
    p2.steps = 30
    p2.total_steps = 20
    psi_o2, crp_o2, steps_o2, _ = oracle.time_evolution(p2, total_steps=20, variant=ov)
    data["time"]["total_steps"] = 20
    entry(data)
    assert data["time"]["steps"] == 50 == steps_o2
    assert rel(psi_buf, psi_o2) < 1e-10
    assert crp_err(crp_buf, crp_o2) < 1e-8
    # the restart really restarted: the accumulated P(E) of 50 steps differs
    psi_o50, crp_o50, _, _ = oracle.time_evolution(pv, total_steps=50, variant=ov)
    assert rel(psi_buf, psi_o50) < 1e-10
    assert np.max(np.abs(crp_o50 - crp_o2)) > 1e-6 * max(np.max(np.abs(crp_o50)), 1e-12)


def test_callback_cadence_and_early_stop(qd, oracle):
    p = qd.problem.config(0)
    print("case: callback every 5 steps, stop at 15", flush=True)
    seen = []

    def cb(steps, crp):
        seen.append((steps, None if crp is None else crp.copy()))
        return steps >= 15

    with qd.Propagator(p.copy(), precision="fp64", crp_every=5, use_graph=1) as prop:
        prop.set_callback(cb)
        prop.propagate(100)
        assert prop.steps() == 15
        psi = prop.psi()
    assert [s for s, _ in seen] == [5, 10, 15]
    for s, crp in seen:
        _, crp_o, _, _ = oracle.time_evolution(p, total_steps=s)
        assert crp is not None and crp_err(crp, crp_o) < 1e-8, s
    psi_o, _, _, _ = oracle.time_evolution(p, total_steps=15)
    assert rel(psi, psi_o) < 1e-10


def test_callback_without_crp_gets_null(qd):
    p = qd.problem.config(0)
    p.calculate_crp = 0
    seen = []
    with qd.Propagator(p.copy(), precision="mixed", crp_every=4) as prop:
        prop.set_callback(lambda s, c: seen.append((s, c)) and False)
        prop.propagate(10)
        assert prop.steps() == 10
    assert seen == [(4, None), (8, None)]
