"""Pins the CPU oracle (oracle/qd_oracle.cpp) before it is trusted as the checker.

Each reference routine the oracle restates is checked against something the
oracle's own formula cannot fake:
  * golden vectors of the reference's compiled Fortran (tests/golden/*.npz):
    LegTransform.F and CRPAux.F must match bit for bit;
  * the known answers the reference embeds (FFT convention,
    common/fftwinterface.C:132-188; stencil table, common-cuda/gradients2.cu);
  * closed forms: brute-force DFT, polynomial exactness of the stencils,
    Gauss-Legendre exactness, the free Gaussian, the infinite-mass phase (with
    numpy's principal square root for the first half step), and the geometric
    series of the flux accumulation.
"""
import ctypes as C
import json
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")
_dp = C.POINTER(C.c_double)


def ptr(a):
    return a.ctypes.data_as(_dp)


# ---------------------------------------------------------------- golden vectors
def test_legtransform_matches_reference_fortran(oracle):
    g = np.load(os.path.join(GOLD, "legtransform.npz"))
    nr, nth, nleg = int(g["nr"]), int(g["nth"]), int(g["nleg"])
    lp = np.zeros(2 * nr * nleg)
    oracle.lib().qdo_forward_legendre(ptr(g["cpsi"]), ptr(lp), nr, nth, nleg, ptr(g["wlegp"]))
    assert np.array_equal(lp, g["lpsi_fwd"])
    cp = np.zeros(2 * nr * nth)
    oracle.lib().qdo_backward_legendre(ptr(cp), ptr(g["lpsi_in"]), nr, nth, nleg, ptr(g["legp"]))
    assert np.array_equal(cp, g["cpsi_bwd"])


def test_crpaux_matches_reference_fortran(oracle):
    g = np.load(os.path.join(GOLD, "crpaux.npz"))
    n1, nt, ne = int(g["n1"]), int(g["nt"]), int(g["ne"])
    fai, dfai = g["fai0"].copy(), g["dfai0"].copy()
    oracle.lib().qdo_psi_time_to_energy(n1, nt, ne, ptr(g["e"]), ptr(g["psit"]), ptr(g["dpsit"]), ptr(fai), ptr(dfai))
    assert np.array_equal(fai, g["fai"]) and np.array_equal(dfai, g["dfai"])
    crp = np.zeros(ne)
    oracle.lib().qdo_calculate_crp(n1, nt, ne, float(g["dr1"]), float(g["mu2"]), ptr(g["w"]), ptr(g["eta"]),
                                   ptr(g["fai"]), ptr(g["dfai"]), ptr(crp))
    assert np.allclose(crp, g["crp"], rtol=1e-15, atol=0)


def test_gauss_legendre_matches_reference_tables(qd):
    g = np.load(os.path.join(GOLD, "gausslegendre.npz"))
    for n in (1, 2, 3, 7, 40, 80, 120, 199):
        x, w = qd.problem.gauss_legendre(n)
        assert np.allclose(x, g["f_x_%d" % n], rtol=0, atol=2e-15), n
        assert np.allclose(w, g["f_w_%d" % n], rtol=1e-13, atol=0), n
    for n in (40, 80, 120, 300):
        x, w = qd.problem.gauss_legendre(n)
        half = g["g_x_%d" % n]
        # the .grids files hold the x >= 0 half, ascending (GaussLegendre2.m:24-30)
        assert np.allclose(x[-half.size:], half, rtol=0, atol=2e-15), n
        assert np.allclose(w[-half.size:], g["g_w_%d" % n], rtol=1e-13, atol=0), n


def test_leps_matches_reference_ksg_surface(qd):
    """OHClKSG.F is a LEPS surface whose r_e and beta depend on the bond angles
    (KSG_re, KSG_beta, :27-78).  With those per-geometry parameters restated here,
    the project's Coulomb/exchange integrals and LEPS combination must reproduce
    the compiled reference surface."""
    pb = qd.problem
    g = np.load(os.path.join(GOLD, "ksg_leps.npz"))
    a = np.array([[1.705, 0.7823, 0.4463], [3.273, -0.9621, -3.116], [2.306, 0.4932, 2.131]])
    b = np.array([[1.313, -1.206, 1.763], [1.268, -2.934, 4.700], [1.239, -2.840, 6.012]])
    r0 = np.array([1.833, 2.967, 2.409])
    beta0 = np.array([1.214, 1.212, 0.9885])
    S = np.array([0.0355, 0.307, 0.153])
    Dm = np.array([0.1699, 0.1030, 0.1697])
    for (r1, r2, r3), vref in zip(g["geo"], g["v"]):
        th1 = math.acos((r1 * r1 + r2 * r2 - r3 * r3) / (2 * r1 * r2))
        th3 = math.acos((r2 * r2 + r3 * r3 - r1 * r1) / (2 * r2 * r3))
        th = [th1, th1 + th3, th3]
        s1 = 0.5 * (1.0 + math.tanh(2.0 * (9.0 - r2)))
        q, j = [], []
        for i, r in enumerate((r1, r2, r3)):
            t2 = th[i] ** 2
            re = r0[i] + (a[i, 0] + a[i, 1] * t2 + a[i, 2] * t2 * t2 - r0[i]) * math.exp(-6 * abs(th[i]) ** 2) * s1
            be = beta0[i] + (b[i, 0] + b[i, 1] * t2 + b[i, 2] * t2 * t2 - beta0[i]) * math.exp(-3 * abs(th[i]) ** 2) * s1
            qi, ji = pb.coulomb_exchange(r, Dm[i], re, be, S[i])
            q.append(float(qi))
            j.append(float(ji))
        v = float(pb.leps_qj(*q, *j))
        assert abs(v - vref) < 1e-13, (r1, r2, r3, v, vref)


def test_stencil_table_matches_reference(oracle):
    table = json.load(open(os.path.join(GOLD, "stencils.json")))
    for h_str, ref in table.items():
        h = int(h_str)
        c = np.zeros(h)
        assert oracle.lib().qdo_gradient_coefficients(2 * h + 1, ptr(c)) == 0
        assert np.allclose(c, ref, rtol=1e-14, atol=1e-18), h


# ---------------------------------------------------------------- FFT and momenta
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 12, 15, 16, 30, 40, 48, 64, 96, 128, 192, 256, 7, 14])
def test_fft_matches_bruteforce_dft(oracle, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    for sign in (-1, 1):
        ref = (W if sign < 0 else W.conj()) @ x
        got = oracle.fft_1d(x, sign)
        assert np.max(np.abs(got - ref)) < 1e-12 * max(1, n)


def test_fft2d_matches_numpy(oracle):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((48, 20)) + 1j * rng.standard_normal((48, 20))
    assert np.allclose(oracle.fft_2d(a, -1), np.fft.fft2(a), atol=1e-11)
    assert np.allclose(oracle.fft_2d(a, +1), np.fft.ifft2(a) * a.size, atol=1e-11)


def test_fft_known_answer_harmonic_oscillator(oracle):
    """common/fftwinterface.C:132-188: HO ground state on linspace(-10, 10, 1024):
    norm 1, <V> = 0.25; after the forward FFT, sum|f|^2/n dx = 1 and
    sum |f|^2 k^2 / n dx / 2 = 0.25 with k = 2 pi/L [0:N/2, -N/2+1:-1], L = n dx."""
    n = 1024
    x = np.linspace(-10, 10, n)
    dx = x[1] - x[0]
    f = (1 / np.pi ** 0.25 * np.exp(-0.5 * x * x)).astype(np.complex128)
    assert abs(np.sum(np.abs(f) ** 2) * dx - 0.999999999999998) < 1e-14
    assert abs(np.sum(np.abs(f) ** 2 * 0.5 * x * x) * dx - 0.25) < 1e-14
    F = oracle.fft_1d(f, -1)
    assert abs(np.sum(np.abs(F) ** 2) / n * dx - 0.999999999999999) < 1e-14
    k = oracle.momentum(n, n * dx)
    kref = (2 * np.pi / (n * dx)) * np.concatenate([np.arange(0, n // 2 + 1), np.arange(-n // 2 + 1, 0)])
    assert np.allclose(k, kref, rtol=1e-12, atol=1e-11)  # running sum, fftwinterface.C:116-124
    assert abs(np.sum(np.abs(F) ** 2 * k * k) / n * dx / 2 - 0.25) < 1e-12


def test_momentum_rejects_odd_n(oracle):
    with pytest.raises(ValueError):
        oracle.momentum(7, 1.0)


# ---------------------------------------------------------------- stencils
@pytest.mark.parametrize("npoints", [3, 5, 7, 9, 11, 13, 21, 31])
def test_gradient_exact_on_polynomials(oracle, npoints):
    h = npoints // 2
    n1, n2, n3 = 3, 40, 2
    dr = 0.05
    r = 1.0 + dr * np.arange(n2)
    s = 20
    # degree 2h polynomial: central differences of this width are exact
    coef = np.random.default_rng(npoints).standard_normal(2 * h + 1) / 10
    p = np.polynomial.Polynomial(coef)
    f = np.zeros((n3, n2, n1), dtype=np.complex128)
    for i in range(n1):
        f[:, :, i] = (1 + 0.5j * i) * p(r - r[s])
    v = np.zeros(2 * n1 * n3)
    g = np.zeros(2 * n1 * n3)
    fl = np.ascontiguousarray(f).view(np.float64).reshape(-1)
    rc = oracle.lib().qdo_gradient3d(n1, n2, n3, s + 1, 1, dr, ptr(fl), ptr(v), ptr(g), npoints)
    assert rc == 0
    vg = v.view(np.complex128).reshape(n3, n1)
    gg = g.view(np.complex128).reshape(n3, n1)
    dp = p.deriv()(0.0)
    for i in range(n1):
        assert np.allclose(vg[:, i], (1 + 0.5j * i) * p(0.0), rtol=1e-13)
        assert np.allclose(gg[:, i], (1 + 0.5j * i) * dp, rtol=1e-7, atol=1e-9)


def test_gradient_rejects_bad_stencil(oracle):
    f = np.zeros(2 * 4 * 20)
    v = np.zeros(8)
    g = np.zeros(8)
    assert oracle.lib().qdo_gradient3d(4, 20, 1, 10, 1, 0.1, ptr(f), ptr(v), ptr(g), 4) == 3
    assert oracle.lib().qdo_gradient3d(4, 20, 1, 2, 1, 0.1, ptr(f), ptr(v), ptr(g), 11) == 4


# ---------------------------------------------------------------- theta transform
def test_legendre_matrix_and_roundtrip(qd):
    from scipy.special import eval_legendre

    n = 40
    x, w = qd.problem.gauss_legendre(n)
    L = qd.problem.legendre_p2(n - 1, x)
    for l in (0, 1, 5, 39):
        assert np.allclose(L[l], eval_legendre(l, x), atol=1e-13)
    Wf = (np.arange(n)[None, :] + 0.5) * w[:, None] * L.T   # [k, l]
    # Gauss-Legendre with M = n integrates P_l P_l' exactly: P_b W_f = I
    assert np.allclose(L.T @ Wf.T, np.eye(n), atol=1e-12)


# ---------------------------------------------------------------- whole-path closed forms
def _tiny_problem(qd, n1=32, n2=24, nth=6, V=None, mass=None, dump=False, steps=10, dt=10.0):
    p = qd.problem.config(0, nR=n1, nr=n2, ntheta=nth, steps=steps, dt=dt)
    if V is not None:
        p.pot = np.full(p.grid_points, float(V)) if np.isscalar(V) else V
    if mass is not None:
        p.mass1 = mass
        p.mass2 = mass
    if not dump:
        p.dump1 = None
        p.dump2 = None
    return p


def test_free_gaussian_matches_analytic(qd, oracle):
    """V = 0 and an l = 0 packet: the step is exact free motion in R and r."""
    pb = qd.problem
    n1, n2, nth = 128, 96, 4
    # short enough that neither packet spreads to the edge of the periodic box
    p = _tiny_problem(qd, n1=n1, n2=n2, nth=nth, V=0.0, steps=20, dt=2.5)
    R, r = p.r1, p.r2
    mu1, mu2 = p.mass1, p.mass2
    d1, d2, R0, r0, k0 = 0.6, 0.3, 6.0, 3.0, 3.0

    def gauss(x, x0, d, k, mu, t):
        a = 1 + 1j * t / (mu * d * d)
        return (np.pi * d * d) ** -0.25 / np.sqrt(a) * np.exp(
            -(x - x0 - k * t / mu) ** 2 / (2 * d * d * a) + 1j * k * (x - x0) - 1j * k * k * t / (2 * mu))

    def packet(t):
        psi = gauss(R, R0, d1, -k0, mu1, t)[:, None] * gauss(r, r0, d2, 0.0, mu2, t)[None, :]
        full = np.repeat(psi[:, :, None], nth, axis=2) * math.sqrt(0.5)
        c = np.ascontiguousarray(np.transpose(full, (2, 1, 0)))
        out = np.empty(c.shape[:2] + (2 * n1,))
        out[..., 0::2] = c.real
        out[..., 1::2] = c.imag
        return out.reshape(-1)

    p.psi = packet(0.0)
    p.calculate_crp = 0
    psi, _, _, _ = oracle.time_evolution(p)
    ref = packet(p.total_steps * p.dt)
    assert np.max(np.abs(psi - ref)) < 1e-10 * np.max(np.abs(ref))


def test_static_phase_and_first_half_step(qd, oracle):
    """Infinite masses: only the potential phase acts.  psi_n = psi_0 e^{-iV dt n}
    / sqrt(e^{-iV dt}) with the principal square root, which flips sign where
    |V dt| > pi (timeEvol.C:519-538 with complex.h:60-68,148-159)."""
    p = _tiny_problem(qd, mass=1e30, steps=7, dt=10.0)
    rng = np.random.default_rng(3)
    p.pot = rng.uniform(0.0, 0.9, p.grid_points)   # V dt up to 9 rad: both branches
    p.calculate_crp = 0
    psi, _, _, _ = oracle.time_evolution(p)
    z0 = p.psi.view(np.complex128)
    e = np.exp(-1j * p.pot * p.dt)
    ref = z0 * e ** p.total_steps / np.sqrt(e)
    assert np.max(np.abs(psi.view(np.complex128) - ref)) < 1e-11
    # CUDA propagator: exact e^{+iV dt/2} (evolutionCUDA2.cu:507)
    psi_c, _, _, _ = oracle.time_evolution(p, variant=1)
    ref_c = z0 * e ** p.total_steps * np.exp(0.5j * p.pot * p.dt)
    assert np.max(np.abs(psi_c.view(np.complex128) - ref_c)) < 1e-11


def test_dump_closed_form_in_both_orders(qd, oracle):
    """Infinite masses with the absorbers on: psi_n = psi_0 e^{iV dt/2}
    e^{-iV n dt} (d1 d2)^n in both propagators -- the CPU one dumps after each
    step (timeEvol.C:667, dump_psi :713-735), the CUDA one before it
    (evolutionCUDA2.cu:503) -- and, after n steps, both have applied n dumps.
    V dt stays below pi so the first half phase is e^{iV dt/2} in both."""
    p = _tiny_problem(qd, mass=1e30, dump=True, steps=9, dt=5.0)
    rng = np.random.default_rng(5)
    p.pot = rng.uniform(0.0, 0.6, p.grid_points)
    p.dump1 = rng.uniform(0.5, 1.0, p.n1)
    p.dump2 = rng.uniform(0.5, 1.0, p.n2)
    p.calculate_crp = 0
    z0 = p.psi.view(np.complex128)
    d = (p.dump2[:, None] * p.dump1[None, :]).reshape(-1)       # (j, i) -> i + n1 j
    dd = np.tile(d, p.ntheta)
    n = p.total_steps
    ref = z0 * np.exp(0.5j * p.pot * p.dt) * np.exp(-1j * p.pot * p.dt * n) * dd ** n
    for variant in (0, 1):
        psi, _, _, _ = oracle.time_evolution(p, variant=variant)
        assert np.max(np.abs(psi.view(np.complex128) - ref)) < 1e-12 * np.max(np.abs(ref)), variant
    # one dump fewer would be off by a factor d: the check sees the count
    assert np.max(np.abs(ref / dd - ref)) > 1e-3 * np.max(np.abs(ref))


def test_flux_accumulation_geometric_series(qd, oracle):
    """Infinite masses and constant V0: psi_s(t_n) = psi_s(0) e^{iV0 dt/2} e^{-iV0 (n+1) dt},
    so fai(E) = G(E) psi_s, G = sum_n dt e^{i (E - V0)(n+1) dt}, and
    P(E) = |G|^2 Im(sum_k w_k sum_i conj(psi_s) dpsi_s) / eta^2 dr1/mu2."""
    V0 = 0.013
    p = _tiny_problem(qd, V=V0, mass=1e30, steps=25, dt=10.0)
    rng = np.random.default_rng(4)
    z = (rng.standard_normal(p.grid_points) + 1j * rng.standard_normal(p.grid_points)) * 0.1
    p.psi = np.ascontiguousarray(z).view(np.float64).copy()
    _, crp, steps, _ = oracle.time_evolution(p)
    assert steps == p.total_steps
    n1, n2, nth = p.n1, p.n2, p.ntheta
    zz = z.reshape(nth, n2, n1)
    s = p.n_dividing_surface - 1
    c = np.array([2100, -600, 150, -25, 2]) / (2520 * p.dr2)
    ps = zz[:, s, :]
    dps = sum(c[j - 1] * (zz[:, s + j, :] - zz[:, s - j, :]) for j in range(1, 6))
    base = np.sum(p.w[:, None] * np.conj(ps) * dps)
    n = np.arange(p.total_steps)
    G = np.array([np.sum(p.dt * np.exp(1j * (E - V0) * (n + 1) * p.dt)) for E in p.energies])
    ref = np.abs(G) ** 2 * np.imag(base) / p.eta_sq * p.dr1 / p.mass2
    assert np.allclose(crp, ref, rtol=1e-9, atol=1e-14)
    # the CUDA propagator reads the surface at the 0-based index and flips the sign
    _, crp_c, _, _ = oracle.time_evolution(p, variant=1)
    ps = zz[:, s + 1, :]
    dps = sum(c[j - 1] * (zz[:, s + 1 + j, :] - zz[:, s + 1 - j, :]) for j in range(1, 6))
    base_c = np.sum(p.w[:, None] * np.conj(ps) * dps)
    ref_c = -np.abs(G) ** 2 * np.imag(base_c) / p.eta_sq * p.dr1 / p.mass2
    assert np.allclose(crp_c, ref_c, rtol=1e-9, atol=1e-14)


def test_norm_conserved_without_dumps(qd, oracle):
    p = qd.problem.config(0, steps=50)
    p.dump1 = None
    p.dump2 = None
    _, _, _, d0 = oracle.time_evolution(p, total_steps=1, diagnostics=1)
    _, _, _, d1 = oracle.time_evolution(p, diagnostics=1)
    assert abs(d1.module - 1.0) < 1e-10 and abs(d1.module - d0.module) < 1e-12


# ---------------------------------------------------------------- setup helpers
def test_morse_levels_from_sine_dvr(qd):
    """The diatomic sine-DVR eigenvalues of a Morse curve against the analytic
    levels E_v = w(v+1/2) - (w(v+1/2))^2/(4 De), w = beta sqrt(2 De / mu)."""
    pb = qd.problem
    m = pb.H2_MORSE
    r = np.linspace(0.4, 6.0, 160)
    dr = r[1] - r[0]
    mu = pb.M_H / 2
    V = m.De * (1 - np.exp(-m.beta * (r - m.re))) ** 2
    om = m.beta * math.sqrt(2 * m.De / mu)
    for v in (0, 1, 2):
        e, phi = pb.diatomic_state(r, dr, mu, V, 0, v)
        ref = om * (v + 0.5) - (om * (v + 0.5)) ** 2 / (4 * m.De)
        assert abs(e - ref) < 1e-7, (v, e, ref)
        assert abs(np.sum(phi ** 2) * dr - 1) < 1e-12


def test_eta_sq_is_energy_normalised_projection(qd):
    """EtaSq (common/EtaSq.m) equals (mu/k) |int g(R) e^{ikR} dR|^2 for the
    incoming Gaussian of InitWavePacket.m."""
    pb = qd.problem
    mu, d, R0, k0 = 1224.8, 0.3, 8.0, 6.0
    R = np.linspace(2, 14, 20001)
    g = (np.pi * d * d) ** -0.25 * np.exp(-(R - R0) ** 2 / (2 * d * d) - 1j * k0 * R)
    for E in (0.005, 0.0147, 0.03):
        k = math.sqrt(2 * mu * E)
        proj = np.trapezoid(g * np.exp(1j * k * R), R)
        ref = mu / k * abs(proj) ** 2
        assert abs(pb.eta_sq(d, k0, mu, np.array([E]))[0] - ref) < 1e-10 * ref


def test_dividing_surface_rounding(qd):
    pb = qd.problem
    assert pb.n_dividing_surface(5.0, 0.4, 0.1) == 46
    assert pb.n_dividing_surface(0.5, 0.25, 0.5) == 1    # 0.5 rounds away from zero
    assert pb.n_dividing_surface(1.5, 0.25, 0.5) == 3    # 2.5 -> 3 (numpy's round gives 2)
    assert pb.n_dividing_surface(0.6, 0.25, 0.5) == 1
