"""|eta_J(E)|^2 for J > 0: the packet's projection onto the energy-normalised
incoming free wave of orbital angular momentum J (problem.eta_sq_partial_wave),
which normalises the flux of the J > 0 blocks.  The reference has J = 0 only
(common/EtaSq.m:4-18); the pins are:

* the Riccati-Hankel recursion against scipy's spherical Bessel functions;
* J = 0 is EtaSq.m's closed form;
* completeness: the packet's projections onto the regular free waves
  F_J(kR) = kR j_J(kR) (scipy) integrate to its norm over all energies, and,
  where the packet is incoming and outside the centrifugal barrier, they equal
  |eta_J|^2 / 2pi -- an independent construction of the same normalisation;
* the WKB limit: far outside the barrier |eta_J(E)|^2 is the plane-wave form
  at the local kinetic energy E - J(J+1)/(2 mu R0^2).
"""
import math

import numpy as np
import pytest
from scipy.special import spherical_jn, spherical_yn


@pytest.fixture(scope="module")
def pb(qd):
    return qd.problem


@pytest.fixture(scope="module")
def packet(pb):
    """The F+H2 packet of the batched workloads: R0 = 8, delta = 0.3, 0.2 eV."""
    mu = pb.F_H2.mu_R
    return dict(mu=mu, R0=8.0, delta=0.3, k0=math.sqrt(2 * mu * 0.2 / pb.H_EV))


@pytest.mark.parametrize("l", [0, 1, 2, 7, 20, 40, 61])
def test_riccati_hankel_matches_scipy(pb, l):
    x = np.linspace(0.3, 300.0, 5000)
    ref = x * (spherical_jn(l, x) + 1j * spherical_yn(l, x))
    got = pb.riccati_hankel_plus(l, x)
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 1e-12


def test_partial_wave_eta_reduces_to_etasq_at_J0(pb, packet):
    E = np.linspace(0.05, 0.6, 200) / pb.H_EV
    got = pb.eta_sq_partial_wave(packet["delta"], packet["R0"], packet["k0"], packet["mu"], E, 0)
    ref = pb.eta_sq(packet["delta"], packet["k0"], packet["mu"], E)
    assert np.max(np.abs(got - ref) / ref) < 1e-12


def _regular_projection(pk, E, J, n=2001):
    """|a_F(E)|^2 = (2 mu / (pi k)) |int F_J(kR) g(R) dR|^2 with F_J from scipy."""
    R = np.linspace(pk["R0"] - 10 * pk["delta"], pk["R0"] + 10 * pk["delta"], n)
    g = (np.pi * pk["delta"] ** 2) ** -0.25 * np.exp(-(R - pk["R0"]) ** 2 / (2 * pk["delta"] ** 2)
                                                       - 1j * pk["k0"] * R)
    k = np.sqrt(2 * pk["mu"] * E)
    x = k[:, None] * R[None, :]
    F = x * spherical_jn(J, x)
    I = np.trapezoid(F * g[None, :], R, axis=1)
    return 2 * pk["mu"] / (np.pi * k) * np.abs(I) ** 2


@pytest.mark.parametrize("J", [0, 10, 20, 40])
def test_completeness_and_incoming_projection(pb, packet, J):
    # the regular free waves are complete: their projections carry the packet's norm
    E = np.linspace(1e-7, 4.0, 6000) / pb.H_EV
    aF = _regular_projection(packet, E, J)
    assert abs(np.trapezoid(aF, E) - 1.0) < 2e-3
    # where the packet sits outside the barrier and its momenta point inward,
    # |eta_J|^2 / 2 pi is that same projection (up to the packet's small outgoing part)
    kmin = math.sqrt(J * (J + 1)) / (packet["R0"] - 3 * packet["delta"]) + 1.0
    Emin = kmin ** 2 / (2 * packet["mu"])
    Ew = np.linspace(max(Emin, 0.1 / pb.H_EV), 0.6 / pb.H_EV, 100)
    eta = pb.eta_sq_partial_wave(packet["delta"], packet["R0"], packet["k0"], packet["mu"], Ew, J)
    aw = _regular_projection(packet, Ew, J)
    assert np.max(np.abs(eta / (2 * np.pi) - aw) / aw) < 0.03


@pytest.mark.parametrize("J", [5, 10, 20])
def test_wkb_limit(pb, packet, J):
    E = np.linspace(0.15, 0.6, 100) / pb.H_EV
    eta = pb.eta_sq_partial_wave(packet["delta"], packet["R0"], packet["k0"], packet["mu"], E, J)
    shift = J * (J + 1) / (2 * packet["mu"] * packet["R0"] ** 2)
    wkb = pb.eta_sq(packet["delta"], packet["k0"], packet["mu"], E - shift)
    # the correction is second order in the barrier over the local kinetic energy
    assert np.max(np.abs(eta / wkb - 1)) < 5e-3


def test_jk_problem_uses_the_partial_wave_normalisation(pb):
    base = pb.config(0, nR=32, nr=24, ntheta=6, steps=2)
    jk = pb.make_jk_problem(base, pb.H_H2, [0, 6], jmax=5)
    assert np.allclose(jk.eta_sq_J(0), jk.base.eta_sq, rtol=1e-11, atol=0)
    pk = jk.packet
    want = pb.eta_sq_partial_wave(pk["delta"], pk["R0"], pk["k0"], pk["mu"], jk.base.energies - pk["e_v"], 6)
    assert np.array_equal(jk.eta_sq_J(6), want)
