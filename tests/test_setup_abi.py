"""The set-up entry points of the C ABI (qd_make_radial, qd_make_theta,
qd_dividing_surface, qd_init_packet), host code that needs no GPU, against
the reference's own data and closed forms:

* Gauss-Legendre points and weights against the reference's tabulated grids
  (GaussLegendre/<n>-LegendreGauss.grids, fixture gausslegendre.npz) and
  P_l(x) against scipy;
* the diatomic sine-DVR eigenvalues against the 25 H-Cl levels recorded in
  OHCl/HClVibRotTest.m:33-57, on the KSG H-Cl curve of OHCl/OHClKSG.F
  (fixture hcl_ksg_levels.npz), and against the analytic Morse levels;
* the packet, e_v and |eta|^2 against the numpy restatement problem.py
  (itself pinned in test_oracle.py and test_eta_partial_wave.py), for J = 0
  (InitWavePacket.m, EtaSq.m) and for a J > 0 block.
"""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def pr(qd):
    return qd.propagator


def test_make_theta_matches_reference_tables(qd, pr):
    from scipy.special import eval_legendre

    g = np.load(os.path.join(GOLD, "gausslegendre.npz"))
    for n in (40, 80, 120, 300):
        x, w, leg = pr.make_theta(n, n - 1)
        half = g["g_x_%d" % n]
        assert np.allclose(x[-half.size:], half, rtol=0, atol=2e-15), n
        assert np.allclose(w[-half.size:], g["g_w_%d" % n], rtol=1e-13, atol=0), n
        for l in (0, 1, 7, n - 1):
            assert np.allclose(leg[l], eval_legendre(l, x), atol=1e-12), (n, l)
    x, w, _ = pr.make_theta(7, 3)
    assert abs(w.sum() - 2.0) < 1e-14 and abs(x[3]) == 0.0


def test_make_radial_and_dividing_surface(qd, pr):
    r, dr = pr.make_radial(256, 0.4, 10.0)
    assert np.array_equal(r, np.linspace(0.4, 10.0, 256))
    assert r[-1] == 10.0 and abs(dr - (r[1] - r[0])) == 0
    pb = qd.problem
    for rd, rmin, d in ((5.0, 0.4, 0.1), (0.5, 0.25, 0.5), (1.5, 0.25, 0.5), (0.6, 0.25, 0.5), (-0.5, 0.25, 0.5)):
        assert pr.dividing_surface(rd, rmin, d) == pb.n_dividing_surface(rd, rmin, d)
    assert pr.dividing_surface(1.5, 0.25, 0.5) == 3   # 2.5 rounds away from zero


def test_diatomic_levels_match_reference_hcl_record(qd, pr):
    """OHCl/HClVibRotTest.m records 25 levels of the H-Cl curve; the recorded
    values are those of the KSG curve itself (HClPES.m now also subtracts
    its minimum, -0.1697 Eh, which shifts every level by that constant)."""
    g = np.load(os.path.join(GOLD, "hcl_ksg_levels.npz"))
    r, v, mu, levels = g["r"], g["v"], float(g["mu"]), g["levels"]
    R = np.linspace(5.0, 10.0, 4)
    for v0 in range(levels.size):
        _, e, _ = pr.init_packet(R, r, 1.0, mu, v, x=None, v0=v0)
        assert abs(e - levels[v0]) < 1e-12, (v0, e, levels[v0])


def test_diatomic_levels_match_morse(qd, pr):
    pb = qd.problem
    m = pb.H2_MORSE
    mu = pb.H_H2.mu_r
    r = np.linspace(0.4, 6.0, 200)
    V = m.De * (1 - np.exp(-m.beta * (r - m.re))) ** 2
    om = m.beta * np.sqrt(2 * m.De / mu)
    for v0 in range(4):
        _, e, _ = pr.init_packet(np.linspace(5, 10, 4), r, 1.0, mu, V, v0=v0)
        ref = om * (v0 + 0.5) - (om * (v0 + 0.5)) ** 2 / (4 * m.De)
        assert abs(e - ref) < 1e-7, (v0, e, ref)


def test_init_packet_j0_matches_problem(qd, pr):
    pb = qd.problem
    p = pb.config(0)
    sysm = pb.H_H2
    k0 = np.sqrt(2 * p.mass1 * 0.5 / pb.H_EV)
    psi, ev, eta = pr.init_packet(p.r1, p.r2, p.mass1, p.mass2, sysm.diatomic_potential(p.r2), x=p.x, R0=8.0,
                                  delta=0.3, k0=k0, energies=p.energies)
    assert abs(ev - p.e_diatomic) < 1e-12
    assert np.max(np.abs(psi - p.psi)) < 1e-12 * np.max(np.abs(p.psi))
    assert np.max(np.abs(eta - p.eta_sq) / p.eta_sq) < 1e-12


@pytest.mark.parametrize("v0,j0", [(1, 0), (0, 3)])
def test_init_packet_excited_states(qd, pr, v0, j0):
    pb = qd.problem
    p = pb.config(0)
    V = pb.H_H2.diatomic_potential(p.r2)
    psi, ev, _ = pr.init_packet(p.r1, p.r2, p.mass1, p.mass2, V, x=p.x, k0=5.0, v0=v0, j0=j0)
    e_ref, phi = pb.diatomic_state(p.r2, p.dr2, p.mass2, V, j0, v0)
    ref = pb.init_wave_packet(p.r1, 0.3, 8.0, 5.0, phi, j0, p.x)
    assert abs(ev - e_ref) < 1e-12
    assert np.max(np.abs(psi - ref)) < 1e-12 * np.max(np.abs(ref))


def test_init_packet_jk_block_matches_problem(qd, pr):
    """J = 6, K = 0 block (j0 = 0): the block's packet on its own theta grid and
    |eta_J|^2 in partial wave l = J (JKProblem.eta_sq_J)."""
    pb = qd.problem
    base = pb.config(0, nR=32, nr=24, ntheta=6, steps=2)
    jk = pb.make_jk_problem(base, pb.H_H2, [6], jmax=5)
    b0 = [b for b in jk.blocks if b.K == 0][0]
    x, _ = pr.jk_theta_grid(0, jk.jmax)
    pk = jk.packet
    psi, ev, eta = pr.init_packet(base.r1, base.r2, base.mass1, base.mass2, pb.H_H2.diatomic_potential(base.r2),
                                  x=x, R0=pk["R0"], delta=pk["delta"], k0=pk["k0"], J=6, K=0,
                                  energies=jk.base.energies)
    ref = jk.psi[b0.offset:b0.offset + base.n1 * base.n2 * b0.n]
    assert np.max(np.abs(psi.view(np.complex128) - ref)) < 1e-12 * np.max(np.abs(ref))
    want = jk.eta_sq_J(6)
    ok = want > 1e-6 * want.max()
    assert np.max(np.abs(eta[ok] - want[ok]) / want[ok]) < 1e-10


def test_init_packet_rejects_bad_input(qd):
    import ctypes as C
    abi = qd._abi
    lib = abi.load_library()
    r = np.linspace(0.5, 5, 8)
    v = np.zeros(8)
    dp = C.POINTER(C.c_double)
    r1 = abi.qd_radial(8, r.ctypes.data_as(dp), r[1] - r[0], 1000.0)
    th = abi.qd_angle()
    pk = abi.qd_packet(5.0, 0.3, 1.0, 0, 1, 4, 2, v.ctypes.data_as(dp))   # K = 2 > j0 = 1
    e = C.c_double()
    assert lib.qd_init_packet(C.byref(r1), C.byref(r1), C.byref(th), C.byref(pk), 0, None, None, C.byref(e),
                              None) == 7
    pk = abi.qd_packet(5.0, 0.3, 1.0, 8, 0, 0, 0, v.ctypes.data_as(dp))   # v0 beyond the grid
    assert lib.qd_init_packet(C.byref(r1), C.byref(r1), C.byref(th), C.byref(pk), 0, None, None, C.byref(e),
                              None) == 7
