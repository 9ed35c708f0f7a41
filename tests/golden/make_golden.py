"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in a container where /root/reference is mounted, after `make -C oracle ref`
has compiled the reference's own Fortran kernels into oracle/_ref/libqdref.so:

    python tests/golden/make_golden.py

Fixtures (inputs and outputs only, small):
  legtransform.npz   ForwardLegendreTransform / BackwardLegendreTransform
                     (common/LegTransform.F) on seeded random inputs
  crpaux.npz         PsiTimeToPsiEnergyOnSurface / CalculateCRP (common/CRPAux.F)
  gausslegendre.npz  GaussLegendre (common/GaussLegendre.F, table n <= 199) for
                     several n, and the tabulated 32-digit half-grids of
                     GaussLegendre/<n>-LegendreGauss.grids for n = 40, 80, 120, 300
  ksg_leps.npz       OHClKSGPES (OHCl/OHClKSG.F) at seeded geometries
  hcl_ksg_levels.npz the H-Cl curve of OHClKSG.F (O far away, as OHCl/HClPES.m
                     places it) on a 512-point grid over 1..40 a0, the H-Cl
                     reduced mass of OHCl/HClVibRotTest.m:7-20, and the 25
                     vibrational levels that script records (:33-57)
  stencils.json      central-difference coefficient table of
                     common-cuda/gradients2.cu:37-581 (half-widths 1..30)
"""
from __future__ import annotations

import ctypes as C
import json
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
_dp = C.POINTER(C.c_double)


def P(a):
    return a.ctypes.data_as(_dp)


def I(v):
    return C.byref(C.c_int(v))


def D(v):
    return C.byref(C.c_double(v))


def main():
    lib = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libqdref.so"))
    rng = np.random.default_rng(0x5EED0001)

    # --- LegTransform.F ---
    nr, nth, nleg = 37, 9, 7
    cpsi = rng.standard_normal(2 * nr * nth)
    wlegp = rng.standard_normal(nth * nleg)
    legp = rng.standard_normal(nleg * nth)
    lpsi_in = rng.standard_normal(2 * nr * nleg)
    lpsi = np.zeros(2 * nr * nleg)
    lib.forwardlegendretransform_(P(cpsi), P(lpsi), I(nr), I(nth), I(nleg), P(wlegp))
    cpsi_out = np.zeros(2 * nr * nth)
    lib.backwardlegendretransform_(P(cpsi_out), P(lpsi_in), I(nr), I(nth), I(nleg), P(legp))
    np.savez(os.path.join(HERE, "legtransform.npz"), nr=nr, nth=nth, nleg=nleg, cpsi=cpsi, wlegp=wlegp,
             lpsi_fwd=lpsi, legp=legp, lpsi_in=lpsi_in, cpsi_bwd=cpsi_out)

    # --- CRPAux.F ---
    n1, nt, ne = 11, 5, 6
    e = rng.standard_normal(2 * ne)
    psit = rng.standard_normal(2 * n1 * nt)
    dpsit = rng.standard_normal(2 * n1 * nt)
    fai0 = rng.standard_normal(2 * n1 * nt * ne)
    dfai0 = rng.standard_normal(2 * n1 * nt * ne)
    fai, dfai = fai0.copy(), dfai0.copy()
    lib.psitimetopsienergyonsurface_(I(n1), I(nt), I(ne), P(e), P(psit), P(dpsit), P(fai), P(dfai))
    w = rng.uniform(0.1, 1.0, nt)
    eta = rng.uniform(0.5, 2.0, ne)
    crp = np.zeros(ne)
    dr1, mu2 = 0.07, 918.575
    lib.calculatecrp_(I(n1), I(nt), I(ne), D(dr1), D(mu2), P(w), P(eta), P(fai), P(dfai), P(crp))
    np.savez(os.path.join(HERE, "crpaux.npz"), n1=n1, nt=nt, ne=ne, e=e, psit=psit, dpsit=dpsit, fai0=fai0,
             dfai0=dfai0, fai=fai, dfai=dfai, w=w, eta=eta, dr1=dr1, mu2=mu2, crp=crp)

    # --- GaussLegendre.F and the .grids tables ---
    out = {}
    for n in (1, 2, 3, 7, 40, 80, 120, 199):
        x = np.zeros(n)
        wt = np.zeros(n)
        lib.gausslegendre_(I(n), P(x), P(wt))
        out["f_x_%d" % n] = x
        out["f_w_%d" % n] = wt
    for n in (40, 80, 120, 300):
        rows = []
        with open(os.path.join(REF, "GaussLegendre", "%d-LegendreGauss.grids" % n)) as f:
            for line in f:
                if line.strip():
                    rows.append([float(v) for v in line.split()])
        a = np.array(rows)
        out["g_x_%d" % n] = a[:, 0]
        out["g_w_%d" % n] = a[:, 1]
    np.savez(os.path.join(HERE, "gausslegendre.npz"), **out)

    # --- OHClKSG.F: LEPS with angle-dependent r_e and beta ---
    geo = []
    vals = []
    for _ in range(64):
        # O-H, O-Cl, H-Cl satisfying the triangle inequality
        a = rng.uniform(1.2, 4.0)
        b = rng.uniform(2.0, 6.0)
        ang = rng.uniform(0.3, np.pi - 0.05)
        c = float(np.sqrt(a * a + b * b - 2 * a * b * np.cos(ang)))
        v = C.c_double()
        lib.ohclksgpes_(D(a), D(b), D(c), C.byref(v))
        geo.append((a, b, c))
        vals.append(v.value)
    np.savez(os.path.join(HERE, "ksg_leps.npz"), geo=np.array(geo), v=np.array(vals))

    # --- HCl levels: the KSG H-Cl curve and the levels OHCl/HClVibRotTest.m records ---
    mass_au = 1.822888484929367e+03
    m_h, m_cl = 1.0079 * mass_au, 35.453 * mass_au
    r = np.linspace(1.0, 40.0, 512)
    v_hcl = np.empty(r.size)
    for i, x in enumerate(r):
        v = C.c_double()
        lib.ohclksgpes_(D(float(np.sqrt(x * x + 100.0 * 100.0))), D(100.0), D(float(x)), C.byref(v))
        v_hcl[i] = v.value
    script = open(os.path.join(REF, "OHCl", "HClVibRotTest.m")).read()
    levels = np.array([float(t) for t in script.split("return", 1)[1].split()])
    np.savez(os.path.join(HERE, "hcl_ksg_levels.npz"), r=r, v=v_hcl, mu=m_h * m_cl / (m_h + m_cl), levels=levels)

    # --- gradients2.cu coefficient table (decimal literals, half-widths 1..30) ---
    src = open(os.path.join(REF, "common-cuda", "gradients2.cu")).read()
    body = src[src.index("const double a[] = {"):]
    body = body[:body.index("};")]
    nums = [float(v) for v in re.findall(r"[-+]?\d+\.\d+E[-+]\d+", body)]
    table = {}
    k = 0
    for h in range(1, 31):
        if k + h > len(nums):
            break
        table[str(h)] = nums[k:k + h]
        k += h
    with open(os.path.join(HERE, "stencils.json"), "w") as f:
        json.dump(table, f)
    print("wrote fixtures to", HERE, "; stencil half-widths:", len(table))


if __name__ == "__main__":
    main()
