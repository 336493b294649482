"""Observables of the GPU engine against exact results: the U = 0 free
fermion closed forms (tests/test_integration.c:15-38, 114-153 of the
reference) and the exact diagonalisation of the 4-site chain at U = 4
(data/benchmark_L468_U4_full_diag_20260626/reference/L4_U4_FullDiag.dat,
tests/golden/ed_l4.npz), the latter after the Trotter extrapolation the
reference itself applies to its ED comparison
(data/L6_U4_mu2_dtau_ed_comparison/: at dtau = 0.1 the energy sits ~5
error bars below ED, the dtau^2 extrapolation meets it)."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def free_szz(Lx, Ly, mx, my, beta, t=-1.0):
    # src/../tests/test_integration.c:15-38
    s = 0.0
    for ky in range(Ly):
        for kx in range(Lx):
            kqx, kqy = (kx + mx) % Lx, (ky + my) % Ly
            e = 2 * t * math.cos(2 * math.pi * kx / Lx)
            eq = 2 * t * math.cos(2 * math.pi * kqx / Lx)
            if Ly > 1:
                e += 2 * t * math.cos(2 * math.pi * ky / Ly)
                eq += 2 * t * math.cos(2 * math.pi * kqy / Ly)
            f, fq = 1 / (1 + math.exp(beta * e)), 1 / (1 + math.exp(beta * eq))
            s += f * (1 - fq)
    return s / (2 * Lx * Ly)


@pytest.mark.parametrize("kind,Lx,Ly,L,stab", [("chain", 4, 1, 40, 8), ("square", 4, 4, 40, 10),
                                               ("square", 8, 8, 40, 10)])
def test_u0_observables_exact(eng_mod, kind, Lx, Ly, L, stab):
    # U = 0: no interaction, every B_l = expK, so the sampled G is the free
    # G for any field: E = 2 sum_k eps_k f(eps_k), n = 2 sum_k f, doublon =
    # (n/2N)^2 per site ... and Szz(q) = free_szz, Sperp = 2 Szz, exactly
    dtau = 0.1
    beta = L * dtau
    eng = eng_mod.Engine(eng_mod.FP64)
    mid = eng.add_model(kind, Lx, Ly, 0.0, dtau, L, stab, delay_k=8)
    eng.add_replicas(mid, [11, 12])
    mq = eng.set_measure(mid, nbin=1, q="all")
    eng.init()
    eng.sweep(2)
    eng.sweep(3, 0)
    th, _, _ = eng_mod.lattice(kind, Lx, Ly)
    w = np.linalg.eigvalsh(th)
    f = 1 / (1 + np.exp(beta * w))
    n = Lx * Ly
    E = 2 * np.sum(w * f)
    N = 2 * np.sum(f)
    for r in range(2):
        b8, cnt, szz, sperp = eng.bins(r)
        assert cnt[0] == 3 and b8[0, 5] == 3.0  # three samples with sign +1
        e_hub, e_gc, e_ph, ntot, dbl = b8[0, :5] / 3.0
        assert abs(e_gc - E) < 1e-10 and abs(ntot - N) < 1e-10
        # doublon = (1/n) sum_i n_up,i n_dn,i with n_up = n_dn = G-diagonal
        # complements; translation invariance makes every site N/(2n)
        assert abs(dbl - (N / (2 * n)) ** 2) < 1e-10
        for k, (mx, my) in enumerate(mq):
            fs = free_szz(Lx, Ly, mx, my, beta)
            assert abs(szz[0, k] / 3.0 - fs) < 1e-10, (mx, my)
            assert abs(sperp[0, k] / 3.0 - 2 * fs) < 1e-10, (mx, my)
    eng.close()


def qmc_l4(eng_mod, dtau, beta=4.0, nrep=128, nwarm=100, nbin=20, per=50):
    L = int(round(beta / dtau))
    eng = eng_mod.Engine(eng_mod.FP64)
    mid = eng.add_model("chain", 4, 1, 4.0, dtau, L, max(1, L // 5), delay_k=8)
    eng.add_replicas(mid, [eng_mod.replica_seed(7, 0, r) for r in range(nrep)])
    eng.set_measure(mid, nbin=nbin, q=None)
    eng.init()
    eng.run(nwarm, per, nbin)
    st = eng.status()
    assert (st["status"] == 0).all()
    egc, dbl = [], []
    for r in range(nrep):
        b8, cnt, _, _ = eng.bins(r)
        egc.extend(b8[:, 1] / b8[:, 5])
        dbl.extend(b8[:, 4] / b8[:, 5])
    eng.close()
    out = []
    for x in (np.array(egc), np.array(dbl)):
        out.append((x.mean(), x.std(ddof=1) / math.sqrt(len(x))))
    return out


def test_ed_l4_trotter_extrapolated(eng_mod):
    # chain L=4, U=4, mu=U/2, beta=4: 128 replicas x 1000 measured sweeps at
    # dtau = 0.1 and 0.05; E_gc and the double occupancy extrapolated as
    # a + b dtau^2 from the two (Richardson) meet the exact diagonalisation
    # at T = 0.25 within 4 combined error bars
    ed = np.load(os.path.join(GOLD, "ed_l4.npz"))
    row = ed["data"][np.argmin(np.abs(ed["data"][:, 0] - 0.25))]
    assert abs(row[0] - 0.25) < 1e-12
    e_ed, d_ed = row[1], row[6] / 4.0
    print("running chain L=4 at dtau 0.1 and 0.05", flush=True)
    (e1, se1), (d1, sd1) = qmc_l4(eng_mod, 0.1)
    (e2, se2), (d2, sd2) = qmc_l4(eng_mod, 0.05)
    e0, se0 = (4 * e2 - e1) / 3, math.sqrt(16 * se2 ** 2 + se1 ** 2) / 3
    d0, sd0 = (4 * d2 - d1) / 3, math.sqrt(16 * sd2 ** 2 + sd1 ** 2) / 3
    print(f"E_gc: dtau 0.1 {e1:.5f}+-{se1:.5f}, 0.05 {e2:.5f}+-{se2:.5f}, extrapolated {e0:.5f}+-{se0:.5f}, "
          f"ED {e_ed:.5f}")
    print(f"D: dtau 0.1 {d1:.6f}+-{sd1:.6f}, 0.05 {d2:.6f}+-{sd2:.6f}, extrapolated {d0:.6f}+-{sd0:.6f}, "
          f"ED {d_ed:.6f}")
    assert abs(e0 - e_ed) < 4 * se0
    assert abs(d0 - d_ed) < 4 * sd0
