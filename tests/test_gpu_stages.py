"""Each kernel of the step on its own, through the C-ABI test hooks, against a
numpy statement of the same operator on random data.

Stages: theta forward / backward (MFMA GEMMs), FFT along R and r (LDS
Stockham passes, forward and inverse), and the fused R and r passes with
their rotational and kinetic phase factors (and, for J > 0, the K <-> l
mixing).  FP64 must agree to ~1e-13 relative; the mixed path to ~1e-5.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rnd(shape, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal(shape) + 1j * r.standard_normal(shape)


def as_f64(z):
    return np.ascontiguousarray(z).view(np.float64).reshape(-1).copy()


def as_c(a, shape):
    return a.view(np.complex128).reshape(shape)


def momenta(n, L):
    return (2 * np.pi / L) * np.concatenate([np.arange(0, n // 2 + 1), np.arange(-n // 2 + 1, 0)])


TOL = {"fp64": 1e-12, "mixed": 3e-5}


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def small(qd):
    # 64 x 48 grid, 40 theta points: the H+H2 workload shape
    return qd.problem.config(0)


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_theta_transforms(qd, small, prec):
    p = small
    n1, n2, nt, M = p.n1, p.n2, p.ntheta, p.m + 1
    print("case: theta transforms", prec, n1, n2, nt, flush=True)
    L = p.legendre  # (M, nt)
    Wf = (np.arange(M)[None, :] + 0.5) * p.w[:, None] * L.T  # [k, l]
    with qd.Propagator(p, precision=prec) as prop:
        z = rnd((nt, n2 * n1), 1)
        prop.set_psi(as_f64(z))
        prop.debug_stage(0)
        got = as_c(prop.get_fbr(), (M, n2 * n1))
        ref = np.einsum("kx,kl->lx", z, Wf)
        assert relerr(got, ref) < TOL[prec]
        y = rnd((M, n2 * n1), 2)
        prop.set_fbr(as_f64(y))
        prop.debug_stage(3)
        got = as_c(prop.psi(), (nt, n2 * n1))
        ref = np.einsum("lx,lk->kx", y, L)
        assert relerr(got, ref) < TOL[prec]


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fft_along_each_axis(qd, small, prec):
    p = small
    n1, n2, M = p.n1, p.n2, p.m + 1
    print("case: FFT axes", prec, flush=True)
    with qd.Propagator(p, precision=prec) as prop:
        for stage, axis, fn in ((4, 2, np.fft.fft), (5, 1, np.fft.fft), (6, 2, np.fft.ifft), (7, 1, np.fft.ifft)):
            y = rnd((M, n2, n1), 10 + stage)
            prop.set_fbr(as_f64(y))
            prop.debug_stage(stage)
            got = as_c(prop.get_fbr(), (M, n2, n1))
            ref = fn(y, axis=axis)
            if fn is np.fft.ifft:
                ref = ref * y.shape[axis]
            assert relerr(got, ref) < TOL[prec], stage


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_fused_passes_with_phases(qd, small, prec):
    p = small
    n1, n2, M = p.n1, p.n2, p.m + 1
    dt = p.dt
    print("case: fused passes", prec, flush=True)
    I1 = 1 / (2 * p.mass1 * p.r1 ** 2)
    I2 = 1 / (2 * p.mass2 * p.r2 ** 2)
    ll = np.arange(M) * (np.arange(M) + 1.0)
    DR = np.exp(-0.5j * dt * ll[:, None] * I1[None, :])     # [l, i]
    Dr = np.exp(-0.5j * dt * ll[:, None] * I2[None, :])     # [l, j]
    T1 = momenta(n1, n1 * p.dr1) ** 2 / (2 * p.mass1)
    T2 = momenta(n2, n2 * p.dr2) ** 2 / (2 * p.mass2)
    with qd.Propagator(p, precision=prec) as prop:
        y = rnd((M, n2, n1), 21)
        prop.set_fbr(as_f64(y))
        prop.debug_stage(1)
        got = as_c(prop.get_fbr(), (M, n2, n1))
        x = y * DR[:, None, :]
        x = np.fft.ifft(np.fft.fft(x, axis=2) * np.exp(-1j * dt * T1)[None, None, :], axis=2)
        ref = x * DR[:, None, :]
        assert relerr(got, ref) < TOL[prec]
        y = rnd((M, n2, n1), 22)
        x = y * Dr[:, :, None]
        x = np.fft.ifft(np.fft.fft(x, axis=1) * np.exp(-1j * dt * T2)[None, :, None], axis=1)
        ref = x * Dr[:, :, None]
        # the r pass with batches of 256 / N1 and of 512 / N1 consecutive R
        # points: the same arithmetic per line, so the same bits
        out = {}
        for nth in (256, 512):
            prop.debug_select(4, nth)
            prop.set_fbr(as_f64(y))
            prop.debug_stage(2)
            out[nth] = as_c(prop.get_fbr(), (M, n2, n1)).copy()
            assert relerr(out[nth], ref) < TOL[prec], nth
        assert np.array_equal(out[256], out[512])


# (R points, r points, J values, j_max): LDS-staged mixing (32 points), the
# prefetching FP32 kernel with 32-line bundles (64 points, groups up to 10
# lines; "two": its variant for two work groups per CU, on 256 points with
# groups of up to 21 lines) and with 48-line bundles (groups of 34 lines), and
# the split pass (mixing in launches of its own around the FFT pass) on 64-,
# 128- (groups of 41 lines) and 256-point R grids
MIX_CASES = {"staged": (32, 24, [0, 2, 3], 5), "prefetch32": (64, 16, [0, 3, 9], 11),
             "prefetch32two": (256, 16, [0, 4, 20], 22), "prefetch48": (64, 16, [33], 35),
             "split64": (64, 32, [0, 3, 9], 11), "split128": (128, 16, [40], 42),
             "split256": (256, 16, [0, 4, 20], 22)}


@pytest.mark.parametrize("case", sorted(MIX_CASES))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_r_pass_mixing_jk(qd, prec, case):
    """J > 0: the R pass mixes the K blocks of each (J, j) into the l basis,
    applies D_R/2 with l(l+1), transforms, and mixes back."""
    nR, nr, Js, jmax = MIX_CASES[case]
    if prec == "fp64" and case == "prefetch48":
        pytest.skip("FP64 groups above 32 lines take the split pass, whose 32-line FFT batches "
                    "need a multiple of 32 r points on a 64-point R grid (this case has 16)")
    pb = qd.problem
    base = pb.config(0, nR=nR, nr=nr, ntheta=6, steps=2)
    jk = pb.make_jk_problem(base, pb.H_H2, Js, jmax=jmax)
    n1, n2, dt = base.n1, base.n2, base.dt
    print("case: R pass mixing", prec, case, flush=True)
    I1 = 1 / (2 * base.mass1 * base.r1 ** 2)
    T1 = momenta(n1, n1 * base.dr1) ** 2 / (2 * base.mass1)
    kin = np.exp(-1j * dt * T1)
    with qd.propagator.JKPropagator(jk, precision=prec) as prop:
        tot = jk.psi.size
        y = rnd(tot, 31)
        if prec == "mixed" and case.startswith("prefetch"):
            two = 1 if case.endswith("two") else 0
            prop.debug_select(0, two)
            assert prop.tuning()["rpass_prefetch_mix_two_groups"] == two
            if case == "prefetch32":
                prop.debug_select(1, 3)  # three consecutive bundles per turn of a work group
        if case.startswith("split"):
            prop.debug_select(2, 1)
        elif prop.tuning()["rpass_split_mix"]:
            prop.debug_select(2, 0)  # FP64 takes the split pass by default where it has no other register-FFT one
        prop.set_fbr(as_f64(y))
        prop.debug_stage(1)
        got = prop.get_fbr().view(np.complex128)
        uses_prefetch = prop.tuning()["rpass_prefetch_mix_bundle_lines"] > 0
        uses_split = prop.tuning()["rpass_split_mix"] == 1
    assert uses_prefetch == (prec == "mixed" and case.startswith("prefetch"))
    assert uses_split == case.startswith("split")
    # numpy statement
    ref = y.copy()
    cols = {}
    c0 = 0
    for b in jk.blocks:
        cols[(b.J, b.K)] = c0
        c0 += b.n
    n12 = n1 * n2
    Y = ref.reshape(-1, n2, n1)
    out = Y.copy()
    for J in jk.Js:
        for j in range(jk.jmax + 1):
            if J == 0:
                U, lv = np.eye(1), np.array([j])
            else:
                U, lv = pb.coriolis_eigen(J, j)
            g = U.shape[0]
            idx = [cols[(J, K)] + (j - K) for K in range(g)]
            c = Y[idx]                                   # [K, r, R]
            cm = np.einsum("Km,Kab->mab", U, c)
            DR = np.exp(-0.5j * dt * (lv * (lv + 1.0))[:, None] * I1[None, :])
            cm = cm * DR[:, None, :]
            cm = np.fft.ifft(np.fft.fft(cm, axis=2) * kin[None, None, :], axis=2)
            cm = cm * DR[:, None, :]
            out[idx] = np.einsum("Km,mab->Kab", U, cm)
    assert relerr(got.reshape(out.shape), out) < TOL[prec]
