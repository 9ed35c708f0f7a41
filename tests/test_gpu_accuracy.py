"""Mixed precision against the FP64 path over a whole propagation.

Bar: per-energy |P_mixed(E) - P_fp64(E)| <= 1e-4 on the F+H2 J=0
256x192x120 grid, 10000 steps of 5 au (the throughput workload), with the
three-piece BF16 split of the theta GEMM.  Also: the mixed step conserves the
norm without absorbers to better than 5e-8 per step (the compensated FP32
tables remove the coherent drift of rounded twiddles and phase factors)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_mixed_full_propagation_within_1e4(qd):
    p = qd.problem.config(2)
    print("case: F+H2 256x192x120, 10000 steps, fp64 vs mixed", flush=True)
    _, crp_d = qd.run(p, precision="fp64")
    _, crp_m = qd.run(p, precision="mixed", theta_split=3)
    d = np.abs(crp_m - crp_d)
    print("max |dP| = %.3e at %.3f eV (max P %.3f)" % (d.max(), p.collision_energies_ev[d.argmax()], crp_d.max()))
    assert d.max() <= 1e-4


@pytest.mark.parametrize("split,limit", [(3, 5e-8)])
def test_mixed_norm_drift(qd, split, limit):
    p = qd.problem.config(2, steps=100)
    p.dump1 = None
    p.dump2 = None
    p.calculate_crp = 0
    with qd.Propagator(p, precision="mixed", theta_split=split) as prop:
        n0 = prop.norm()
        prop.propagate(100)
        n1 = prop.norm()
    assert abs(n1 - n0) / n0 / 100 < limit
