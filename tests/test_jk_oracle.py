"""The numpy (J, K) oracle (oracle/jk_oracle.py) and the body-fixed helpers it
uses, pinned by facts fixed outside them.  The reference propagates J = 0
only (common/timeEvol.C:465-491 uses l(l+1) alone), so nothing in it covers
J > 0; the checks here come from the mathematics:

* the Coriolis matrix (J - j)^2 in the parity-adapted K basis has the exact
  spectrum l(l+1), l = |J-j|..J+j with J + j + l even;
* its half-step exponential equals a brute-force matrix exponential;
* the associated-Legendre DVR nodes are the Gauss-Gegenbauer(K + 1/2) nodes
  and its basis values are the normalised associated Legendre functions
  (scipy.special, independent code);
* the batched step is unitary when the absorbers are off;
* with only J = 0 the oracle is the reference step, checked against the C++
  oracle (itself pinned by tests/test_oracle.py).
"""
import math

import numpy as np
import pytest
import scipy.linalg
import scipy.special


@pytest.fixture(scope="module")
def jko():
    import jk_oracle
    return jk_oracle


@pytest.mark.parametrize("J,j", [(1, 1), (2, 5), (5, 2), (7, 7), (20, 13), (3, 0), (0, 4)])
def test_coriolis_spectrum_is_l_l_plus_1(qd, J, j):
    pb = qd.problem
    B = pb.coriolis_matrix(J, j)
    assert np.allclose(B, B.T)
    lam = np.sort(np.linalg.eigvalsh(B))
    ls = [l for l in range(abs(J - j), J + j + 1) if (J + j + l) % 2 == 0]
    want = np.sort(np.array([l * (l + 1.0) for l in ls]))
    assert lam.size == want.size == min(J, j) + 1
    assert np.allclose(lam, want, atol=1e-9 * max(1.0, want.max()))
    U, l = pb.coriolis_eigen(J, j)
    assert sorted(l.tolist()) == sorted(ls)
    assert np.allclose(U.T @ U, np.eye(U.shape[0]), atol=1e-12)


def test_rotational_half_step_equals_matrix_exponential(qd):
    """U exp(-i dt/2 l(l+1) I) U^T (as the oracle and the R pass apply it) is
    exp(-i dt/2 (J-j)^2 I) computed by scipy's Pade exponential."""
    pb = qd.problem
    dt, inertia = 5.0, 1.0 / (2 * 1500.0 * 2.3 ** 2)
    for J, j in [(4, 6), (9, 9), (12, 3)]:
        B = pb.coriolis_matrix(J, j)
        U, l = pb.coriolis_eigen(J, j)
        lam = pb.rot_lambda(J, j, l)
        ours = U @ np.diag(np.exp(-0.5j * dt * lam * inertia)) @ U.T
        ref = scipy.linalg.expm(-0.5j * dt * inertia * B)
        assert np.max(np.abs(ours - ref)) < 1e-12


@pytest.mark.parametrize("K", [1, 2, 5, 9])
def test_assoc_legendre_dvr_nodes_and_basis(qd, K):
    pb = qd.problem
    jmax = 14
    x, om, Pb = pb.assoc_legendre_dvr(K, jmax)
    n = jmax - K + 1
    # nodes: zeros of Pt_{jmax+1}^K / (1-x^2)^(K/2), a Gegenbauer polynomial
    xg, _ = scipy.special.roots_gegenbauer(n, K + 0.5)
    assert np.allclose(np.sort(x), np.sort(xg), atol=1e-13)
    # basis values: normalised associated Legendre functions (sign fixed by Pt_K^K > 0)
    for jj in range(K, jmax + 1):
        norm = math.sqrt((2 * jj + 1) / 2.0 * math.exp(math.lgamma(jj - K + 1) - math.lgamma(jj + K + 1)))
        val = norm * scipy.special.lpmv(K, jj, x) * (-1) ** K
        assert np.allclose(Pb[jj - K], val, atol=1e-11), jj
    # quadrature is exact for the products: Pb diag(om) Pb^T = I
    assert np.allclose(Pb @ np.diag(om) @ Pb.T, np.eye(n), atol=1e-12)


@pytest.fixture(scope="module")
def small_jk(qd):
    pb = qd.problem
    base = pb.config(0, nR=32, nr=24, ntheta=6, steps=4)
    return pb, base


@pytest.mark.parametrize("coriolis", [True, False])
def test_jk_oracle_unitary_without_absorbers(qd, jko, small_jk, coriolis):
    pb, base = small_jk
    b = base.copy()
    b.dump1 = None
    b.dump2 = None
    jk = pb.make_jk_problem(b, pb.H_H2, [0, 2, 3], jmax=5, j0=1, K0=1, coriolis=coriolis)
    # the first step's half phase is not unitary for |V dt| > pi (sqrt_ref);
    # start from a packet already past it
    n0 = jko.norm(jk, jk.psi)
    psi, _, _ = jko.propagate(jk, 6, first=False)
    assert abs(jko.norm(jk, psi) / n0 - 1) < 1e-12


def test_jk_oracle_j0_is_the_reference_step(qd, oracle, jko, small_jk):
    """Only J = 0: the batched oracle and the C++ restatement of timeEvol.C
    give the same psi and P(E) on the same inputs."""
    pb, base = small_jk
    jk = pb.make_jk_problem(base, pb.H_H2, [0], jmax=5)
    steps = 6
    psi, PJ, _ = jko.propagate(jk, steps)
    p = jk.base.copy()
    assert p.ntheta == jk.jmax + 1 and p.m == jk.jmax
    p.pot = jk.pot[0].copy()
    p.psi = np.ascontiguousarray(jk.psi).view(np.float64).copy()
    ref_psi, ref_crp, nsteps, _ = oracle.time_evolution(p, total_steps=steps)
    assert nsteps == steps
    got = psi.view(np.float64)
    assert np.max(np.abs(got - ref_psi)) < 1e-11 * np.max(np.abs(ref_psi))
    scale = max(np.max(np.abs(ref_crp)), 1e-2)
    assert np.max(np.abs(PJ[0] - ref_crp)) < 1e-9 * scale
