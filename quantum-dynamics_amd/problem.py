"""Problem construction: grids, LEPS surface, initial packet, absorbers, eta^2.

This is the setup side that the reference does in MATLAB before it calls the
propagator (``H3/main.m:24-140``).  Each helper restates one reference script:

* ``gauss_legendre``       common/GaussLegendre.F:4-38, common/GaussLegendre2.m:4-32
* ``legendre_p2``          common/LegendreP2.m:4-23 (stored transposed, H3/main.m:101-104)
* ``legendre_p``           common/LegendreP.m:4-23
* ``woods_saxon``          common/WoodsSaxon.m:4-6
* ``diatomic_state``       H3/H2VibRotWaveFunction.m:4-47 (H3/H2PES.m:3-9 for the potential)
* ``init_wave_packet``     H3/InitWavePacket.m:2-31
* ``eta_sq``               common/EtaSq.m:4-18
* ``eta_sq_partial_wave``  the same projection in partial wave J > 0 (reduces to EtaSq.m at J = 0)
* ``jacobi_to_bonds``      H3/H3PESJacobi.m:6-29
* ``leps``                 OHCl/OHClKSG.F:89-150 (Q/J form), constant r_e and beta
* ``n_dividing_surface``   H3/main.m:77 (MATLAB int32() rounds half away from zero)

The workloads (``config``) are the H+H2 and F+H2 LEPS systems used for parity
and throughput.  Everything is float64 numpy; nothing here touches the GPU.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

H_EV = 27.21138505          # H3/main.m:34
M_H = 1837.15               # electron masses
M_F = 34631.0


# --------------------------------------------------------------------------
# angular grid
# --------------------------------------------------------------------------
def gauss_legendre(n: int):
    """Gauss-Legendre nodes (ascending) and weights on [-1, 1].

    Newton iteration on P_n from the Chebyshev-like initial guess; weights
    2 / ((1 - x^2) P_n'(x)^2).  The reference
    reads the same nodes from a table (common/GaussLegendre.F:13-31) and sorts
    them ascending (Sort2Double).
    """
    if n < 1:
        raise ValueError("n must be >= 1")
    if n == 1:
        return np.array([0.0]), np.array([2.0])
    # Newton on P_n in extended precision (x86 long double) so that the
    # weights keep ~1e-15 relative accuracy for n in the hundreds
    ld = np.longdouble
    k = np.arange(1, n + 1, dtype=ld)
    x = np.cos(ld(np.pi) * (k - ld(0.25)) / (n + ld(0.5)))
    for _ in range(100):
        p0 = np.ones_like(x)
        p1 = x.copy()
        for j in range(2, n + 1):
            p0, p1 = p1, ((2 * j - 1) * x * p1 - (j - 1) * p0) / j
        dp = n * (x * p1 - p0) / (x * x - 1)
        dx = p1 / dp
        x = x - dx
        if np.max(np.abs(dx)) < ld(1e-19):
            break
    p0 = np.ones_like(x)
    p1 = x.copy()
    for j in range(2, n + 1):
        p0, p1 = p1, ((2 * j - 1) * x * p1 - (j - 1) * p0) / j
    dp = n * (x * p1 - p0) / (x * x - 1)
    w = (2 / ((1 - x * x) * dp * dp)).astype(np.float64)
    x = x.astype(np.float64)
    order = np.argsort(x)
    x = x[order]
    w = w[order]
    # exact symmetry, as the tabulated half-grids give it
    x = 0.5 * (x - x[::-1])
    w = 0.5 * (w + w[::-1])
    if n % 2:
        x[n // 2] = 0.0
    return x, w


def legendre_p2(m: int, x: np.ndarray) -> np.ndarray:
    """P_l(x_k) for l = 0..m, returned as the (m+1) x n matrix the propagator
    takes (LegendreP2(m, x)' in H3/main.m:101-104).  Recursion
    P_i = (2 - 1/i) x P_{i-1} - (1 - 1/i) P_{i-2}  (LegendreP2.m:18-20)."""
    x = np.asarray(x, dtype=np.float64).reshape(-1)
    v = np.zeros((x.size, m + 1))
    v[:, 0] = 1.0
    if m >= 1:
        v[:, 1] = x
    for i in range(2, m + 1):
        v[:, i] = (2 - 1 / i) * x * v[:, i - 1] - (1 - 1 / i) * v[:, i - 2]
    return np.ascontiguousarray(v.T)


def legendre_p(n: int, x: np.ndarray) -> np.ndarray:
    """P_n(x) by the same recursion (common/LegendreP.m)."""
    x = np.asarray(x, dtype=np.float64)
    if n == 0:
        return np.ones_like(x)
    p0, p1 = np.ones_like(x), x.copy()
    for i in range(2, n + 1):
        p0, p1 = p1, (2 - 1 / i) * x * p1 - (1 - 1 / i) * p0
    return p1


def woods_saxon(cd: float, xd: float, x: np.ndarray) -> np.ndarray:
    """1 / (1 + exp(Cd (x - xd)))  (common/WoodsSaxon.m:5)."""
    return 1.0 / (1.0 + np.exp(cd * (np.asarray(x) - xd)))


def quadratic_absorber_dump(x: np.ndarray, width: float, strength: float, dt: float) -> np.ndarray:
    """Multiplicative dump exp(-V_cap dt) of a quadratic imaginary absorber
    V_cap = strength ((x - x0)/width)^2 over the last `width` of the grid.
    The reference applies absorbers only as such per-step dumps (timeEvol.C:713-735)."""
    x = np.asarray(x, dtype=np.float64)
    x0 = x[-1] - width
    u = np.clip((x - x0) / width, 0.0, None)
    return np.exp(-strength * u * u * dt)


def n_dividing_surface(rd: float, rmin: float, dr: float) -> int:
    """int32((rd - min(r)) / dr) with MATLAB rounding (H3/main.m:77)."""
    v = (rd - rmin) / dr
    return int(math.floor(abs(v) + 0.5)) * (1 if v >= 0 else -1)


# --------------------------------------------------------------------------
# LEPS potential
# --------------------------------------------------------------------------
@dataclasses.dataclass(frozen=True)
class Morse:
    De: float      # Eh
    re: float      # a0
    beta: float    # 1/a0


def coulomb_exchange(r, De, re, beta, sato):
    """Coulomb and exchange integrals of one pair (OHClKSG.F:89-91,107-108):
      Q = D/4 ((3+S) e^{-2x} - (2+6S) e^{-x}) / (1+S)
      J = D/4 ((1+3S) e^{-2x} - (6+2S) e^{-x}) / (1+S),   x = beta (r - r_e)."""
    e = np.exp(-beta * (np.asarray(r, dtype=np.float64) - re))
    q = 0.25 * De * ((3.0 + sato) * e * e - (2.0 + 6.0 * sato) * e) / (1.0 + sato)
    j = 0.25 * De * ((1.0 + 3.0 * sato) * e * e - (6.0 + 2.0 * sato) * e) / (1.0 + sato)
    return q, j


def leps_qj(q1, q2, q3, j1, j2, j3):
    """V = Q1 + Q2 + Q3 - sqrt(J1^2 + J2^2 + J3^2 - J1 J2 - J2 J3 - J3 J1)  (OHClKSG.F:148-150)."""
    return q1 + q2 + q3 - np.sqrt(np.maximum(j1 * j1 + j2 * j2 + j3 * j3 - j1 * j2 - j2 * j3 - j3 * j1, 0.0))


def leps(rab: np.ndarray, rbc: np.ndarray, rac: np.ndarray, m_ab: Morse, m_bc: Morse,
         m_ac: Morse, sato) -> np.ndarray:
    """London-Eyring-Polanyi-Sato surface with constant Morse parameters per pair;
    `sato` is one Sato parameter or a (ab, bc, ac) triple."""
    s = sato if isinstance(sato, (tuple, list)) else (sato, sato, sato)
    q1, j1 = coulomb_exchange(rab, m_ab.De, m_ab.re, m_ab.beta, s[0])
    q2, j2 = coulomb_exchange(rbc, m_bc.De, m_bc.re, m_bc.beta, s[1])
    q3, j3 = coulomb_exchange(rac, m_ac.De, m_ac.re, m_ac.beta, s[2])
    return leps_qj(q1, q2, q3, j1, j2, j3)


def jacobi_to_bonds(R, r, x, masses):
    """Bond lengths r12, r13, r23 of A + BC in Jacobi (R, r, cos theta) (H3PESJacobi.m:6-29)."""
    f2 = masses[2] / (masses[1] + masses[2])
    f3 = 1.0 - f2
    r12 = np.sqrt((f2 * r) ** 2 + R ** 2 - 2 * f2 * r * R * x)
    r13 = np.sqrt((f3 * r) ** 2 + R ** 2 + 2 * f3 * r * R * x)
    r23 = np.broadcast_to(r, r12.shape)
    return r12, r13, r23


@dataclasses.dataclass(frozen=True)
class LepsSystem:
    """A + BC on a LEPS surface. pairs: (A-B, B-C, A-C) Morse curves."""
    name: str
    masses: tuple           # (m_A, m_B, m_C) in electron masses
    ab: Morse
    bc: Morse
    ac: Morse
    sato: float
    vmax: float = 2.5       # H3PESJacobi.m:13
    rmin: float = 0.4       # H3PESJacobi.m:14

    def potential(self, R, r, x):
        """V(R, r, theta) shifted so the BC asymptote minimum is 0, capped at vmax
        where a bond is shorter than rmin (H3PESJacobi.m:16-27)."""
        r12, r13, r23 = jacobi_to_bonds(R, r, x, self.masses)
        v = leps(r12, r23, r13, self.ab, self.bc, self.ac, self.sato) + self.bc.De
        bad = (r12 <= self.rmin) | (r13 <= self.rmin) | (r23 <= self.rmin)
        v = np.where(bad, self.vmax, v)
        return np.minimum(v, self.vmax)

    def diatomic_potential(self, r):
        """BC curve from the surface at a far collinear geometry (H3/H2PES.m:3-9)."""
        r = np.asarray(r, dtype=np.float64)
        far = np.full_like(r, 100.0)
        return leps(far, r, far + r, self.ab, self.bc, self.ac, self.sato) + self.bc.De

    @property
    def mu_R(self):
        mA, mB, mC = self.masses
        return mA * (mB + mC) / (mA + mB + mC)

    @property
    def mu_r(self):
        _, mB, mC = self.masses
        return mB * mC / (mB + mC)


H2_MORSE = Morse(0.1745, 1.401, 1.028)
HF_MORSE = Morse(0.2250, 1.733, 1.174)
H_H2 = LepsSystem("H+H2", (M_H, M_H, M_H), H2_MORSE, H2_MORSE, H2_MORSE, 0.10)
F_H2 = LepsSystem("F+H2", (M_F, M_H, M_H), HF_MORSE, H2_MORSE, HF_MORSE, 0.16)


# --------------------------------------------------------------------------
# initial state
# --------------------------------------------------------------------------
def diatomic_state(r: np.ndarray, dr: float, mu: float, vdiat: np.ndarray, j: int, v: int):
    """Eigenpair (e_v, psi_v(r)) of the diatomic on the r grid.

    Colbert-Miller radial sine-DVR kinetic matrix, indexed from 1 as the
    reference does on its offset grid (H2VibRotWaveFunction.m:19-27):
      T_ii = pi^2/3 - 1/(2 i^2),  T_ij = (-1)^(i-j) (2/(i-j)^2 - 2/(i+j)^2),  / (2 mu dr^2)
    plus V(r) + j(j+1)/(2 mu r^2) on the diagonal; psi normalised to sum |psi|^2 dr = 1.
    """
    n = r.size
    i = np.arange(1, n + 1, dtype=np.float64)
    I, J = np.meshgrid(i, i, indexing="ij")
    with np.errstate(divide="ignore", invalid="ignore"):
        H = np.where(I == J, np.pi ** 2 / 3 - 1.0 / (2 * I * I),
                     (-1.0) ** (I - J) * (2.0 / (I - J) ** 2 - 2.0 / (I + J) ** 2))
    H = H / (2 * mu * dr * dr)
    H[np.diag_indices(n)] += vdiat + j * (j + 1) / (2 * mu * r * r)
    e, vecs = np.linalg.eigh(H)
    psi = vecs[:, v] / math.sqrt(dr)
    if psi[np.argmax(np.abs(psi))] < 0:
        psi = -psi
    return float(e[v]), psi


def eta_sq(delta: float, k0: float, mass: float, E: np.ndarray) -> np.ndarray:
    """|eta(E)|^2 of the incoming Gaussian (common/EtaSq.m:8-16)."""
    E = np.asarray(E, dtype=np.complex128)
    kE = np.sqrt(2 * mass * E)
    eta = np.pi ** 0.25 * np.sqrt(2 * mass * delta / kE) * np.exp(-delta * delta / 2 * (kE - k0) ** 2)
    return np.abs(eta) ** 2


def riccati_hankel_plus(l: int, x) -> np.ndarray:
    """w_l(x) = x h_l^(1)(x) = x (j_l(x) + i y_l(x)), the outgoing Riccati-Hankel
    function, by the upward recursion w_{l+1} = (2l+1)/x w_l - w_{l-1} from
    w_0 = -i e^{ix}, w_1 = -(1 + i/x) e^{ix} (stable: h_l is dominated by the
    irregular y_l).  |w_l(x)| -> 1 for x >> l."""
    x = np.asarray(x, dtype=np.float64)
    e = np.exp(1j * x)
    w0 = -1j * e
    if l == 0:
        return w0
    w1 = -(1.0 + 1j / x) * e
    for n in range(1, l):
        w0, w1 = w1, (2 * n + 1) / x * w1 - w0
    return w1


def eta_sq_partial_wave(delta: float, R0: float, k0: float, mass: float, E, J: int,
                        n_quad: int = 4001, width: float = 10.0) -> np.ndarray:
    """|eta_J(E)|^2 of the packet g(R) = (pi delta^2)^(-1/4) exp(-(R-R0)^2/2delta^2 - i k0 R)
    (H3/InitWavePacket.m:13-17) in partial wave J:

        |eta_J(E)|^2 = (mass / k) |int w_J(k R) g(R) dR|^2,   k = sqrt(2 mass E),

    the projection onto the energy-normalised incoming free wave of orbital
    angular momentum J (Riccati-Hankel w_J^- = conj(w_J)), in the normalisation
    of common/EtaSq.m:8-16, to which it reduces exactly at J = 0 (w_0 = -i e^{ikR}).
    For a packet far outside the centrifugal barrier it tends to the plane-wave
    form evaluated at the local momentum k_J(R0) = sqrt(k^2 - J(J+1)/R0^2),
    i.e. at E - J(J+1)/(2 mass R0^2) (the leading WKB term).  E <= 0 gives 0.
    Quadrature: trapezoid over R0 +- width*delta (spectrally accurate for the
    Gaussian-weighted integrand)."""
    E = np.asarray(E, dtype=np.float64).reshape(-1)
    out = np.zeros(E.size)
    ok = E > 0
    if not np.any(ok):
        return out
    R = np.linspace(R0 - width * delta, R0 + width * delta, n_quad)
    R = R[R > 0]
    dR = R[1] - R[0]
    g = (1.0 / (np.pi * delta ** 2)) ** 0.25 * np.exp(-(R - R0) ** 2 / (2 * delta * delta) - 1j * k0 * R)
    k = np.sqrt(2 * mass * E[ok])
    w = riccati_hankel_plus(J, k[:, None] * R[None, :])
    f = w * g[None, :]
    I = dR * (f.sum(axis=1) - 0.5 * (f[:, 0] + f[:, -1]))
    out[ok] = mass / k * np.abs(I) ** 2
    return out


def init_wave_packet(R, delta, R0, k0, phi_r, j, x):
    """psi(R_i, r_j, theta_k) = g(R_i) phi(r_j) sqrt(j+1/2) P_j(x_k) as interleaved
    (2 n1, n2, ntheta) doubles (H3/InitWavePacket.m:13-31), returned flat."""
    g = (1.0 / (np.pi * delta ** 2)) ** 0.25 * np.exp(-(R - R0) ** 2 / (2 * delta * delta) - 1j * k0 * R)
    P = math.sqrt(j + 0.5) * legendre_p(j, x)
    psi = g[:, None, None] * phi_r[None, :, None] * P[None, None, :]
    # column-major with i fastest: (ntheta, n2, n1) in C order, re/im interleaved
    c = np.ascontiguousarray(np.transpose(psi, (2, 1, 0)))
    inter = np.empty(c.shape[:2] + (2 * R.size,))
    inter[..., 0::2] = c.real
    inter[..., 1::2] = c.imag
    return inter.reshape(-1)


# --------------------------------------------------------------------------
# a complete problem
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Problem:
    """Every array of the reference struct (common/TimeEvolutionMex.C:30-68)."""
    name: str
    r1: np.ndarray
    r2: np.ndarray
    dr1: float
    dr2: float
    mass1: float
    mass2: float
    x: np.ndarray
    w: np.ndarray
    m: int
    legendre: np.ndarray          # (m+1) x n, C-contiguous -> column-major (l fastest) when flattened as .T
    pot: np.ndarray               # flat, n1*n2*ntheta (i fastest)
    psi: np.ndarray               # flat, 2*n1*n2*ntheta interleaved
    dt: float
    total_steps: int
    steps: int
    dump1: Optional[np.ndarray]
    dump2: Optional[np.ndarray]
    n_dividing_surface: int
    n_gradient_points: int
    energies: np.ndarray          # total energies, Eh
    eta_sq: np.ndarray
    crp: np.ndarray
    calculate_crp: int = 1
    e_diatomic: float = 0.0
    collision_energies_ev: Optional[np.ndarray] = None

    @property
    def n1(self):
        return self.r1.size

    @property
    def n2(self):
        return self.r2.size

    @property
    def ntheta(self):
        return self.x.size

    @property
    def grid_points(self):
        return self.n1 * self.n2 * self.ntheta

    def legendre_colmajor(self):
        """legendre[l + (m+1) k] = P_l(x_k)."""
        return np.ascontiguousarray(self.legendre.T).reshape(-1)

    def copy(self):
        return dataclasses.replace(self, psi=self.psi.copy(), crp=self.crp.copy())


def make_problem(system: LepsSystem, *, R_range, nR, r_range, nr, ntheta, m=None, dt, steps,
                 R0, delta, e0_ev, rs, absorber_width=2.0, absorber_strength=0.02,
                 energies_ev=(0.3, 1.5), n_energies=200, j0=0, v0=0, n_gradient_points=11,
                 name="") -> Problem:
    """Build a reference-struct problem on uniform R, r grids and a GL theta grid."""
    R = np.linspace(R_range[0], R_range[1], nR)
    r = np.linspace(r_range[0], r_range[1], nr)
    dR = R[1] - R[0]
    dr = r[1] - r[0]
    x, w = gauss_legendre(ntheta)
    if m is None:
        m = ntheta - 1
    leg = legendre_p2(m, x)
    muR, mur = system.mu_R, system.mu_r
    pot = system.potential(R[:, None, None], r[None, :, None], x[None, None, :])
    pot_flat = np.ascontiguousarray(np.transpose(pot, (2, 1, 0))).reshape(-1)
    e_v, phi = diatomic_state(r, dr, mur, system.diatomic_potential(r), j0, v0)
    k0 = math.sqrt(2 * muR * e0_ev / H_EV)
    psi = init_wave_packet(R, delta, R0, k0, phi, j0, x)
    d1 = quadratic_absorber_dump(R, absorber_width, absorber_strength, dt)
    d2 = quadratic_absorber_dump(r, absorber_width, absorber_strength, dt)
    ec = np.linspace(energies_ev[0], energies_ev[1], n_energies)
    energies = ec / H_EV + e_v
    eta2 = eta_sq(delta, k0, muR, energies - e_v)
    return Problem(name=name or system.name, r1=R, r2=r, dr1=dR, dr2=dr, mass1=muR, mass2=mur,
                   x=x, w=w, m=m, legendre=leg, pot=pot_flat, psi=psi, dt=dt, total_steps=steps,
                   steps=0, dump1=d1, dump2=d2,
                   n_dividing_surface=n_dividing_surface(rs, r[0], dr),
                   n_gradient_points=n_gradient_points, energies=energies, eta_sq=eta2,
                   crp=np.zeros(n_energies), e_diatomic=e_v, collision_energies_ev=ec)


def config(cfg: int, **overrides) -> Problem:
    """The workloads: 0 = H+H2 J=0 (64x48x40, 2000 x 10 au), 1 = F+H2 J=0
    (192x128x80, 10000 x 5 au), 2 = F+H2 J=0 (256x192x120, 10000 x 5 au).
    Ranges not fixed by the workload statement (F+H2 grids, energy window)
    are chosen here and kept fixed."""
    if cfg == 0:
        kw = dict(R_range=(0.5, 12.0), nR=64, r_range=(0.4, 6.0), nr=48, ntheta=40, dt=10.0,
                  steps=2000, R0=8.0, delta=0.3, e0_ev=0.5, rs=3.5, energies_ev=(0.3, 1.5),
                  name="H+H2 J=0 64x48x40")
        system = H_H2
    elif cfg in (1, 2):
        big = cfg == 2
        kw = dict(R_range=(0.5, 12.0), nR=256 if big else 192, r_range=(0.4, 7.5),
                  nr=192 if big else 128, ntheta=120 if big else 80, dt=5.0, steps=10000,
                  R0=8.0, delta=0.3, e0_ev=0.2, rs=4.0, energies_ev=(0.05, 0.6),
                  name="F+H2 J=0 %s" % ("256x192x120" if big else "192x128x80"))
        system = F_H2
    else:
        raise ValueError("unknown config %r" % (cfg,))
    kw.update(overrides)
    return make_problem(system, **kw)


# --------------------------------------------------------------------------
# J > 0: body-fixed K blocks with Coriolis coupling (no reference counterpart;
# the reference propagates J = 0 only, common/timeEvol.C:487)
# --------------------------------------------------------------------------
def assoc_legendre_dvr(K: int, jmax: int):
    """DVR of the normalised associated Legendre functions Pt_j^K(x), j = K..jmax.

    The points x_k are the eigenvalues of x = cos(theta) in that basis
    (tridiagonal, <jK|x|j+1 K> = sqrt(((j+1)^2 - K^2) / ((2j+1)(2j+3)))); they
    are the Gauss-Jacobi(K, K) nodes.  With omega_k the matching weights,
      Pb[j, k] = Pt_j^K(x_k),   Wf[k, j] = omega_k Pt_j^K(x_k),   Pb @ Wf = I.
    For K = 0 these are the Gauss-Legendre nodes and weights.
    Returns (x, omega, Pb) with Pb of shape (n, n), n = jmax - K + 1."""
    n = jmax - K + 1
    j = np.arange(K, jmax, dtype=np.float64)
    off = np.sqrt(((j + 1) ** 2 - K * K) / ((2 * j + 1) * (2 * j + 3)))
    X = np.diag(off, 1) + np.diag(off, -1)
    x, U = np.linalg.eigh(X)
    # Pt_K^K(x) = c_K (1 - x^2)^(K/2),  c_K = sqrt((2K+1)/2) sqrt((2K)!) / (2^K K!)
    logc = 0.5 * math.log((2 * K + 1) / 2.0) + 0.5 * math.lgamma(2 * K + 1) - K * math.log(2.0) - math.lgamma(K + 1)
    pkk = np.exp(logc) * (1 - x * x) ** (K / 2.0)
    sgn = np.sign(U[0, :] * pkk)
    U = U * sgn[None, :]
    omega = (U[0, :] / pkk) ** 2
    Pb = U / np.sqrt(omega)[None, :]
    return x, omega, Pb


def coriolis_matrix(J: int, j: int) -> np.ndarray:
    """(J - j)^2 in the parity-adapted body-fixed basis K = 0..min(J, j):
      diagonal  J(J+1) + j(j+1) - 2K^2,
      K, K+1    -lambda+(J,K) lambda+(j,K) sqrt(1 + delta_K0),  lambda+(a,K) = sqrt(a(a+1) - K(K+1)).
    Its eigenvalues are l(l+1), l = |J-j|..J+j with J + j + l even."""
    nk = min(J, j) + 1
    B = np.zeros((nk, nk))
    for K in range(nk):
        B[K, K] = J * (J + 1) + j * (j + 1) - 2 * K * K
        if K + 1 < nk:
            c = -math.sqrt(J * (J + 1) - K * (K + 1)) * math.sqrt(j * (j + 1) - K * (K + 1))
            if K == 0:
                c *= math.sqrt(2.0)
            B[K, K + 1] = B[K + 1, K] = c
    return B


def coriolis_eigen(J: int, j: int):
    """Eigenvectors U (columns) and the exact orbital quantum numbers l of
    coriolis_matrix(J, j), ascending."""
    B = coriolis_matrix(J, j)
    lam, U = np.linalg.eigh(B)
    l = np.rint((-1 + np.sqrt(1 + 4 * np.maximum(lam, 0))) / 2).astype(int)
    return U, l


@dataclasses.dataclass
class JKBlock:
    J: int
    K: int
    n: int                 # theta DVR points = FBR functions (j = K..jmax)
    offset: int            # first complex element of the block in the batched psi


@dataclasses.dataclass
class JKProblem:
    """A batch of body-fixed (J, K) blocks sharing the R, r grids.

    psi is one complex array with the blocks back to back, each
    n1 x n2 x n_K column-major (R fastest).  Block (J=0, K=0) uses the
    reference's theta matrices (Wf = (l+1/2) w P_l, Pb = P_l, timeEvol.C:737-760)
    so that it reproduces the J = 0 propagator exactly; blocks with J > 0 use
    the normalised associated Legendre FBR so the Coriolis mixing is orthogonal."""
    base: Problem                  # grids, masses, dt, dumps, energies, surface
    jmax: int
    Js: list
    blocks: list
    x: dict                        # K -> DVR points
    omega: dict                    # K -> weights
    Pb: dict                       # key -> (n, n) Pb[j, k]
    Wf: dict                       # key -> (n, n) Wf[k, j]
    pot: dict                      # K -> V (n1*n2*n_K flat)
    psi: np.ndarray                # complex, all blocks
    coriolis: bool = True
    packet: Optional[dict] = None  # R0, delta, k0, e_v, mu of the initial packet

    def eta_sq_J(self, J: int, kind: str = "partial_wave") -> np.ndarray:
        """|eta_J(E)|^2 of the packet in partial wave J (its initial state has
        j0 = 0, so the orbital angular momentum is l = J).

        kind = "partial_wave": the projection onto the energy-normalised
        incoming wave of the centrifugal potential J(J+1)/2mu R^2
        (eta_sq_partial_wave); equal to base.eta_sq at J = 0.
        kind = "wkb": its leading WKB term, the plane-wave |eta|^2 taken at the
        kinetic energy E - e_v - J(J+1)/(2 mu R0^2) left at the packet centre
        (zero below that shift).

        The plane-wave |eta(E)|^2 of J = 0 (base.eta_sq) ignores the
        centrifugal energy the packet starts with; at the top of the energy
        window it then overstates the incoming flux's deficit and lets P^J(E)
        of large J exceed 1."""
        pk = self.packet
        Ek = np.asarray(self.base.energies, dtype=np.float64) - pk["e_v"]
        if kind == "partial_wave":
            return eta_sq_partial_wave(pk["delta"], pk["R0"], pk["k0"], pk["mu"], Ek, J)
        if kind != "wkb":
            raise ValueError("kind must be 'partial_wave' or 'wkb'")
        Ek = Ek - J * (J + 1.0) / (2.0 * pk["mu"] * pk["R0"] ** 2)
        out = np.zeros(Ek.size)
        ok = Ek > 0
        out[ok] = eta_sq(pk["delta"], pk["k0"], pk["mu"], Ek[ok])
        return out

    @property
    def grid_points(self):
        return int(self.psi.size)

    def matrix_key(self, b: JKBlock):
        return "ref0" if b.J == 0 else b.K

    def mixing(self, J: int):
        """For each j: (U_j, l-values) of the K coupling at this J."""
        out = {}
        for j in range(self.jmax + 1):
            if self.coriolis:
                out[j] = coriolis_eigen(J, j)
            else:
                nk = min(J, j) + 1
                U = np.eye(nk)
                # centrifugal-sudden: diagonal (J-j)^2 -> J(J+1) + j(j+1) - 2K^2 = l(l+1) is not
                # an integer l in general; carry the diagonal value itself
                lam = np.array([J * (J + 1) + j * (j + 1) - 2 * K * K for K in range(nk)], dtype=float)
                out[j] = (U, lam)
        return out


def rot_lambda(J: int, j: int, lvals):
    """Eigenvalues of (J-j)^2 per eigen slot: l(l+1) for integer l, else the value."""
    lv = np.asarray(lvals)
    if lv.dtype.kind in "iu":
        return (lv * (lv + 1)).astype(float)
    return lv.astype(float)


def make_jk_problem(base: Problem, system: LepsSystem, Js, jmax: int, j0: int = 0, K0: int = 0,
                    v0: int = 0, R0=None, delta=None, e0_ev=None, coriolis=True) -> JKProblem:
    """Blocks K = 0..min(J, jmax) for every J in Js on base's R, r grids; the
    initial packet g(R) phi_v0j0(r) Pt_j0^K0(x) sits in block K0 of every J."""
    R, r = base.r1, base.r2
    n1, n2 = R.size, r.size
    x, om, Pbm, Wfm, pot = {}, {}, {}, {}, {}
    Kmax = min(max(Js), jmax)
    for K in range(Kmax + 1):
        if K == 0:
            # the Gauss-Legendre grid of the J = 0 propagator, so block (0, 0)
            # and every K = 0 block see the same potential values
            xk, ok = gauss_legendre(jmax + 1)
            Pk = np.sqrt(np.arange(jmax + 1) + 0.5)[:, None] * legendre_p2(jmax, xk)
        else:
            xk, ok, Pk = assoc_legendre_dvr(K, jmax)
        x[K], om[K] = xk, ok
        Pbm[K] = Pk
        Wfm[K] = (ok[:, None] * Pk.T)
        V = system.potential(R[:, None, None], r[None, :, None], xk[None, None, :])
        pot[K] = np.ascontiguousarray(np.transpose(V, (2, 1, 0))).reshape(-1)
    if 0 in Js:
        xg, wg = gauss_legendre(jmax + 1)
        L = legendre_p2(jmax, xg)
        Pbm["ref0"] = L
        Wfm["ref0"] = (wg[:, None] * ((np.arange(jmax + 1) + 0.5)[None, :] * L.T))
    # initial state
    delta = delta if delta is not None else 0.3
    R0 = R0 if R0 is not None else 8.0
    muR, mur = system.mu_R, system.mu_r
    k0 = math.sqrt(2 * muR * (e0_ev if e0_ev is not None else 0.2) / H_EV)
    e_v, phi = diatomic_state(r, r[1] - r[0], mur, system.diatomic_potential(r), j0, v0)
    g = (1.0 / (np.pi * delta ** 2)) ** 0.25 * np.exp(-(R - R0) ** 2 / (2 * delta * delta) - 1j * k0 * R)
    blocks = []
    off = 0
    for J in Js:
        for K in range(min(J, jmax) + 1):
            n = jmax - K + 1
            blocks.append(JKBlock(J, K, n, off))
            off += n1 * n2 * n
    psi = np.zeros(off, dtype=np.complex128)
    for b in blocks:
        if b.K != K0 or j0 < K0:
            continue
        key = "ref0" if b.J == 0 else b.K
        # DVR values of Pt_j0^K0: row j0 - K of Pb (normalised) or sqrt(j0+1/2) P_j0 for ref0
        if key == "ref0":
            th = math.sqrt(j0 + 0.5) * Pbm["ref0"][j0]
        else:
            th = Pbm[K0][j0 - K0]
        blk = g[:, None, None] * phi[None, :, None] * th[None, None, :]
        psi[b.offset:b.offset + blk.size] = np.transpose(blk, (2, 1, 0)).reshape(-1)
    # P(E) is normalised by |eta(E)|^2 of this packet (not of base's)
    jbase = base.copy()
    jbase.eta_sq = eta_sq(delta, k0, muR, base.energies - e_v)
    return JKProblem(base=jbase, jmax=jmax, Js=list(Js), blocks=blocks, x=x, omega=om, Pb=Pbm, Wf=Wfm,
                     pot=pot, psi=psi, coriolis=coriolis,
                     packet=dict(R0=R0, delta=delta, k0=k0, e_v=e_v, mu=muR))
