"""Writes the golden trajectories of the two large configurations from the
compiled reference (oracle/_ref, built from /root/reference by
oracle/Makefile), in the build container:

    python tests/golden/make_large.py

  traj_c3.npz  16x16 periodic, U=4, dtau=0.1, beta=8 (80 slices), stab=8
               (the metric's workload), seeds 1 and 2;
  traj_c5.npz  20x20 periodic, U=8, dtau=0.05, beta=12 (240 slices), stab=4
               (the stabilisation stress case), seeds 1 and 2.

Per seed and sweep (2 sweeps): the field (packed bits, 1 = +1), the
accepted-flip count, trace and Frobenius norm of G_up, and G_up at a fixed
set of entries (the diagonal at every 4th site and 128 seeded random
off-diagonal positions): the full G is 0.5 / 1.3 MB per sample.  For seed
1 also the initial state: the field, the reference's initial G at the
sampled entries, the exact initial G (oracle/exact.c, 80-bit) at the same
entries on the engine host layer's propagators (dqmc_gpu.model), and the
reference's own distance from the exact G of its propagators (max|dG| /
max|G|).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "sai-qmc_amd"))
import dqmc_gpu  # noqa: E402
import pyoracle  # noqa: E402

CASES = {"c3": ("square", 16, 16, 4.0, 0.1, 80, 8), "c5": ("square", 20, 20, 8.0, 0.05, 240, 4)}


def sample_index(n):
    rng = np.random.default_rng(12345)
    d = np.arange(0, n, 4)
    ii = rng.integers(0, n, 128)
    jj = rng.integers(0, n, 128)
    return np.concatenate([d, ii]), np.concatenate([d, jj])


def main(which=("c3", "c5")):
    for name in which:
        kind, Lx, Ly, U, dtau, L, stab = CASES[name]
        n = Lx * Ly
        ii, jj = sample_index(n)
        out = {"params": np.array([Lx, Ly, U, dtau, L, stab]), "ii": ii, "jj": jj}
        for seed in (1, 2):
            r = pyoracle.RefReplica(kind, Lx, Ly, U, dtau, L, stab, seed)
            if seed == 1:
                f0 = r.field()
                g0 = r.green()[0]
                eKr, _, _, _, lamr = r.model()
                gxr = pyoracle.green_exact(eKr, lamr, U, dtau, f0, L, 0, stab)
                th, _, _ = dqmc_gpu.lattice(kind, Lx, Ly)
                eK, _, _, lam = dqmc_gpu.model(th, U, dtau)
                gx = pyoracle.green_exact(eK, lam, U, dtau, f0, L, 0, stab)
                out["field0"] = np.packbits(f0 > 0)
                out["ginit_ref"] = g0[ii, jj]
                out["ginit_exact"] = gx[ii, jj]
                out["ginit_ref_err"] = np.array(np.abs(g0 - gxr).max() / np.abs(gxr).max())
                print(name, "init: reference vs exact", float(out["ginit_ref_err"]), flush=True)
            fs, acc, tr, fro, gs = [], [], [], [], []
            for s in range(2):
                assert r.sweep() == 0
                g = r.green()[0]
                fs.append(np.packbits(r.field() > 0))
                acc.append(r.stats()[0])
                tr.append(np.trace(g))
                fro.append(np.linalg.norm(g))
                gs.append(g[ii, jj])
                print(name, seed, s, flush=True)
            out[f"fields_s{seed}"] = np.array(fs)
            out[f"accepted_s{seed}"] = np.array(acc, dtype=np.uint64)
            out[f"trace_s{seed}"] = np.array(tr)
            out[f"fro_s{seed}"] = np.array(fro)
            out[f"gsample_s{seed}"] = np.array(gs)
        np.savez_compressed(os.path.join(HERE, f"traj_{name}.npz"), **out)


if __name__ == "__main__":
    main(sys.argv[1:] or ("c3", "c5"))
