"""End-to-end parity of the C host program (sai-qmc_amd/dqmc_gpu_main)
against the reference's own output files for the same input: identical
header lines, and every observable, Szz(q) and Sperp(q) within the combined
jackknife error.  The engine replays the reference's RNG stream, so in FP64
the two runs follow the same Markov chain: on both cases every printed
digit of observables.dat agrees; the error-bar test is what stays robust
if a single marginal accept ever resolved differently.

C1: chain L=4, U=4, beta=4, 300+3000 sweeps, 30 bins, seed 1.
C2: 8x8, U=4, beta=4, stab 10, 64 replicas (replica_seed(1, 0, r)),
    200+2000 sweeps, 20 bins.
C3: 4x4, U=4, beta=3, stab 5, 8 replicas, sweep_order=alternating,
    green_rebuild=two_sided, 200+2000 sweeps, 20 bins."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "sai-qmc_amd", "dqmc_gpu_main")

CASES = {
    "c1": "lattice=chain\nLx=4\nU=4\ndtau=0.1\nbeta_list=4\nnwarm=300\nnmeas=3000\nnbin=30\n"
          "seed=1\nszz_q=all\nsperp_q=all\n",
    "c2": "lattice=square\nLx=8\nLy=8\nU=4\ndtau=0.1\nbeta_list=4\nstab=10\nnwarm=200\n"
          "nmeas=2000\nnbin=20\nseed=1\nparallel=omp\nnrep=64\nszz_q=all\nsperp_q=all\n",
    "c3": "lattice=square\nLx=4\nLy=4\nU=4\ndtau=0.1\nbeta_list=3\nstab=5\nnwarm=200\nnmeas=2000\nnbin=20\n"
          "seed=3\nparallel=omp\nnrep=8\nsweep_order=alternating\ngreen_rebuild=two_sided\nszz_q=all\nsperp_q=all\n",
}


def split(path):
    head, rows = [], []
    for line in open(path):
        (head if line.startswith("#") else rows).append(line.rstrip("\n"))
    return head, np.array([[float(x) for x in r.split()] for r in rows])


def within(a, ea, b, eb, what):
    tol = 3.0 * np.sqrt(ea ** 2 + eb ** 2) + 1e-12 * np.abs(b)
    bad = np.abs(a - b) > tol
    assert not bad.any(), (what, a[bad], b[bad], tol[bad])


@pytest.mark.parametrize("case", ["c1", "c2", "c3"])
def test_driver_matches_reference_outputs(case, tmp_path):
    if not os.path.exists(os.path.join(GOLD, f"{case}_observables.txt")):
        pytest.skip(f"no reference outputs for {case}")
    assert os.path.exists(EXE), "dqmc_gpu_main is not built"
    (tmp_path / "in.txt").write_text(CASES[case])
    print(f"running dqmc_gpu_main on {case}", flush=True)
    subprocess.run([EXE, "in.txt"], cwd=tmp_path, check=True, timeout=600)
    # observables.dat: T, (E_hub, dE), (E_gc, dE), (E_ph, dE), (N, dN), (D, dD), sign, (acc, dacc)
    h0, r0 = split(os.path.join(GOLD, f"{case}_observables.txt"))
    h1, r1 = split(tmp_path / "observables.dat")
    print("reference:", open(os.path.join(GOLD, f"{case}_observables.txt")).read().splitlines()[-1])
    print("engine:   ", open(tmp_path / "observables.dat").read().splitlines()[-1], flush=True)
    assert h1 == h0
    assert r1.shape == r0.shape
    assert r1[0, 0] == r0[0, 0]
    for c, name in ((1, "E_hub"), (3, "E_gc"), (5, "E_ph"), (7, "ntot"), (9, "doublon"), (12, "acceptance")):
        within(r1[:, c], r1[:, c + 1], r0[:, c], r0[:, c + 1], name)
    assert np.all(r1[:, 11] == r0[:, 11])  # average sign (1 on the PH path)
    for f in ("szz", "sperp"):
        g0, s0 = split(os.path.join(GOLD, f"{case}_{f}.txt"))
        g1, s1 = split(tmp_path / f"{f}.dat")
        assert g1 == g0
        assert s1.shape == s0.shape
        assert np.array_equal(s1[:, :10], s0[:, :10])  # beta, T, q index, mx, my, q/pi columns
        within(s1[:, 10], s1[:, 11], s0[:, 10], s0[:, 11], f)
