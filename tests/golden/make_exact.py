"""Writes tests/golden/exact_g.npz: Green's functions of fixed field
configurations evaluated with mpmath at 50 significant digits, the check of
the oracle's extended-precision G (oracle/exact.c).

    python tests/golden/make_exact.py

For each case (lattice, U, dtau, L, tau) the field is the oracle replica's
after one forward sweep from the given seed; expK and lambda are the
oracle's model (the same FP64 numbers the engine is given).  The mpmath
evaluation is the direct formula

    G(tau) = (1 + B_{tau-1} ... B_0 B_{L-1} ... B_tau)^-1,
    B_l = expK diag(d_l),  d_{l,j} = exp(+-lambda - dtau U / 2),

with no stabilisation: at 50 digits the product's range (up to ~1e17 here)
costs no accuracy that matters at FP64 output.
"""
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

CASES = [
    # kind, Lx, Ly, U, dtau, L, stab, seed, tau
    ("chain", 4, 1, 4.0, 0.1, 40, 8, 1, 0),
    ("square", 4, 4, 4.0, 0.1, 20, 5, 7, 0),
    ("square", 4, 4, 4.0, 0.1, 20, 5, 7, 10),
    ("square", 6, 6, 4.0, 0.1, 24, 5, 3, 0),
    ("square", 8, 8, 4.0, 0.1, 40, 10, 2, 0),
]


def exact_mp(expK, lam, U, dtau, field, L, tau):
    mp.mp.dps = 50
    n = expK.shape[0]
    K = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            K[i, j] = mp.mpf(float(expK[i, j]))
    c = -mp.mpf(dtau) * mp.mpf(U) / 2
    ev = {-1: mp.exp(-mp.mpf(lam) + c), 1: mp.exp(mp.mpf(lam) + c)}
    f = field.reshape(L, n)
    P = mp.eye(n)
    for k in range(L):
        l = (tau + k) % L
        D = mp.diag([ev[int(s)] for s in f[l]])
        P = K * (D * P)
    G = (mp.eye(n) + P) ** -1
    return np.array([[float(G[i, j]) for j in range(n)] for i in range(n)])


def main():
    out = {}
    for k, (kind, Lx, Ly, U, dtau, L, stab, seed, tau) in enumerate(CASES):
        rep = pyoracle.Replica(kind, Lx, Ly, U, dtau, L, stab, seed)
        rep.sweep()
        eK, _, lam = rep.model()
        f = rep.field()
        g = exact_mp(eK, lam, U, dtau, f, L, tau)
        out[f"case{k}_params"] = np.array([U, dtau, L, tau, lam])
        out[f"case{k}_expK"] = eK
        out[f"case{k}_field"] = f
        out[f"case{k}_G"] = g
        print(kind, Lx, Ly, L, tau, "done", flush=True)
    out["ncases"] = np.array(len(CASES))
    np.savez_compressed(os.path.join(HERE, "exact_g.npz"), **out)


if __name__ == "__main__":
    main()
