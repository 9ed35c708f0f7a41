"""Host-side angular tables of the J > 0 path on the CPU (no GPU): the
Coriolis eigenvectors built in C++ (qd_tables.hpp, compiled here with g++)
and the FP32 rounding that keeps their column lengths, which the mixing R pass
depends on.  With plain rounding U_f = U + E, applying U_f^T then U_f every
step changes the norm by 4 sum_m (U^T E)_mm |z_m|^2 per step, coherently; the
norm-preserving rounding brings the column terms (U^T E)_mm down by orders of
magnitude while moving no entry more than 2 float spacings."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("tables") / "tables")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "cpu_harness", "tables_main.cpp")],
                   check=True)
    return exe


CASES = [(1, 5), (5, 30), (12, 12), (20, 20), (20, 40), (20, 60), (33, 35), (40, 41)]


def test_coriolis_eigenvectors_and_fp32_rounding(harness, qd):
    args = [str(v) for c in CASES for v in c]
    out = subprocess.run([harness] + args, check=True, capture_output=True, text=True).stdout.split("\n")
    rows = [list(map(float, line.split())) for line in out if line.strip()]
    assert len(rows) == len(CASES)
    for (J, j), r in zip(CASES, rows):
        g, orth, col_plain, col_fixed, ulp_plain, ulp_fixed = int(r[2]), r[3], r[4], r[5], r[6], r[7]
        assert g == min(J, j) + 1
        assert orth < 1e-13                       # orthonormal eigenvectors
        assert ulp_plain <= 0.5 + 1e-9            # nearest rounding
        assert ulp_fixed <= 2.0 + 1e-9            # at most two float spacings away
        assert col_fixed <= col_plain
        if g >= 6:  # enough entries per column to cancel the rounding
            assert col_fixed < 1e-9 and col_fixed < 0.1 * col_plain, (J, j, col_plain, col_fixed)
        # the C++ eigenvectors span the same space as numpy's (up to column signs)
        U, l = qd.problem.coriolis_eigen(J, j)
        assert U.shape == (g, g)


def test_centrifugal_eta_shift(qd):
    pb = qd.problem
    base = pb.config(0, nR=32, nr=24, ntheta=6, steps=2)
    jk = pb.make_jk_problem(base, pb.H_H2, [0, 6], jmax=5)
    # J = 0: no shift, the plain |eta|^2 of the packet
    assert np.allclose(jk.eta_sq_J(0, kind="wkb"), jk.base.eta_sq, rtol=1e-12, atol=0)
    pk = jk.packet
    J = 6
    shift = J * (J + 1) / (2 * pk["mu"] * pk["R0"] ** 2)
    Ek = jk.base.energies - pk["e_v"] - shift
    want = np.where(Ek > 0, pb.eta_sq(pk["delta"], pk["k0"], pk["mu"], np.maximum(Ek, 1e-300)), 0.0)
    assert np.allclose(jk.eta_sq_J(J, kind="wkb"), want, rtol=1e-12, atol=0)
    # the shifted |eta|^2 is the plain one evaluated shift lower in energy: its peak moves up by the shift
    assert (jk.base.energies[np.argmax(jk.eta_sq_J(J, kind="wkb"))]
            >= jk.base.energies[np.argmax(jk.base.eta_sq)])


@pytest.fixture(scope="module")
def pack_harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("pack") / "pack")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "cpu_harness", "pack_main.cpp")],
                   check=True)
    return exe


@pytest.mark.parametrize("Jmax,jmax,n2,nl,min_fill", [
    (20, 60, 192, 32, 0.99),   # the batched workload: 231 blocks, groups of up to 21 lines
    (40, 60, 192, 48, 0.90),   # every J up to 40: groups of up to 41 lines in 48-line bundles
    (9, 11, 16, 32, 0.0),      # a small batch (fewer units than bundles can pair)
])
def test_bundle_packing(pack_harness, Jmax, jmax, n2, nl, min_fill):
    """Every mixing unit lands in exactly one bundle of at most nl lines and two
    groups, in r order within its group; the largest-first packing with a
    complementary partner fills the batched workload's line slots completely."""
    out = subprocess.run([pack_harness, str(Jmax), str(jmax), str(n2), str(nl)], check=True, capture_output=True,
                         text=True).stdout.split()
    assert out[0] == "ok", " ".join(out)
    nb, lines, fill = int(out[1]), int(out[2]), float(out[3])
    groups = sum(min(J, j) + 1 for J in range(Jmax + 1) for j in range(jmax + 1))
    assert lines == groups * n2
    assert fill >= min_fill, fill


@pytest.fixture(scope="module")
def records_harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("records") / "records")
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-o", exe,
                    os.path.join(HERE, "cpu_harness", "records_main.cpp")], check=True)
    return exe


# (n1, n2, jmax, coriolis, Js): the batched workloads and the GPU tests' batches
RECORD_WORKLOADS = [
    (256, 192, 60, 1, list(range(21))),               # throughput workload: J = 0..20, 231 blocks
    (256, 192, 60, 1, list(range(41))),               # every J of the learned-interpolation system
    (256, 192, 60, 1, list(range(0, 41, 4))),         # its propagated J
    (256, 192, 60, 1, [1, 9, 17, 25, 33]),            # its held-out J
    (256, 192, 60, 1, [0, 10, 20, 40]),               # the full-length accuracy batch
    (256, 192, 60, 0, [0, 10, 20]),                   # centrifugal-sudden (no mixing matrices)
    (48, 32, 7, 1, [0, 1, 3]),                        # oracle-parity batch
    (64, 32, 11, 1, [0, 4, 9]),                       # prefetch-mixing accuracy batch
    (64, 16, 35, 1, [33]),                            # 34-line groups (48-line bundles)
    (64, 48, 8, 1, list(range(33))),                  # H+H2 learned-interpolation batch
]


@pytest.mark.parametrize("n1,n2,jmax,coriolis,Js", RECORD_WORKLOADS)
def test_work_records_stay_in_bounds(records_harness, n1, n2, jmax, coriolis, Js):
    """Every index the mixing R pass and the fused theta kernel form from their
    host-built records lies inside the arrays they read and write, for every
    batched workload (qd_records.hpp: the library's own checks, then an
    independent walk over lines, matrix reads and tiles)."""
    r = subprocess.run([records_harness, str(n1), str(n2), str(jmax), str(coriolis)] + [str(J) for J in Js],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5])
def test_work_record_checks_catch_planted_mistakes(records_harness, plant):
    """A line past the FBR array, a unit past its bundle, a D_R row past the
    table, a tile past the grid and an uncovered tile are each reported."""
    env = dict(os.environ, QD_PLANT=str(plant))
    r = subprocess.run([records_harness, "64", "32", "11", "1", "0", "4", "9"], capture_output=True, text=True,
                       timeout=120, env=env)
    assert r.returncode == 0 and r.stdout.startswith("fail"), r.stdout
