"""Checkpoints (qd_checkpoint_save / qd_checkpoint_load): the propagation state
-- wavefunction, energy-resolved flux accumulators, the open time block's
surface slices, step counts -- saved mid-block and loaded into a fresh handle
continues bit for bit like the uninterrupted run.  The reference keeps this
state only in memory (timeEvol.C:762-805) and restarts it on every call."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec,graph,variant", [("fp64", 0, "cpu"), ("mixed", 0, "cpu"), ("mixed", 1, "cpu"),
                                                ("mixed", 0, "cuda"), ("mixed", 1, "cuda")])
def test_checkpoint_resumes_bit_for_bit(qd, prec, graph, variant):
    """Mixed precision runs the fused theta step; in the CUDA propagator's
    order its checkpoint carries psi beside the FBR state, and psi reads back
    right after a load."""
    p = qd.problem.config(0)
    print("case: checkpoint", prec, "graph", graph, variant, flush=True)
    with qd.Propagator(p.copy(), precision=prec, use_graph=graph, variant=variant) as a:
        a.propagate(100)                      # 64-step block closed, 36 slices open
        ck = a.checkpoint()
        psi_ck = a.psi()
        a.propagate(150)
        psi_a, crp_a, n_a = a.psi(), a.crp(), a.steps()
    with qd.Propagator(p.copy(), precision=prec, use_graph=graph, variant=variant) as b:
        b.restore(ck)
        assert b.steps() == 100
        assert np.array_equal(b.psi(), psi_ck)
        b.propagate(150)
        psi_b, crp_b, n_b = b.psi(), b.crp(), b.steps()
    assert n_a == n_b == 250
    assert np.array_equal(psi_a, psi_b)
    assert np.array_equal(crp_a, crp_b)


def test_checkpoint_before_the_first_step(qd, oracle):
    """A checkpoint taken at steps = 0 carries the pending first half step."""
    p = qd.problem.config(0)
    with qd.Propagator(p.copy(), precision="fp64") as a:
        ck = a.checkpoint()
    with qd.Propagator(p.copy(), precision="fp64") as b:
        b.restore(ck)
        b.propagate(30)
        psi, crp = b.psi(), b.crp()
    psi_o, crp_o, _, _ = oracle.time_evolution(p, total_steps=30)
    assert np.max(np.abs(psi - psi_o)) < 1e-10 * np.max(np.abs(psi_o))
    assert np.max(np.abs(crp - crp_o)) < 1e-8 * max(np.max(np.abs(crp_o)), 1e-2)


def test_checkpoint_refused_by_another_problem(qd):
    p = qd.problem.config(0)
    q = qd.problem.config(0, nR=32)
    with qd.Propagator(p.copy(), precision="fp64") as a:
        ck = a.checkpoint()
    with qd.Propagator(q.copy(), precision="fp64") as b:
        with pytest.raises(qd._abi.QDError) as ei:
            b.restore(ck)
        assert ei.value.code == 7
    with qd.Propagator(p.copy(), precision="mixed") as c:
        with pytest.raises(qd._abi.QDError):
            c.restore(ck)
