"""GPU parity of the batched UDV linear algebra against the oracle
(udv_lmul / udv_rmul / udv_combine / udv_inv_one_plus, src/linalg.c) and
against the reference's own golden vectors and mpmath fixture."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)


def bmats(orc, Lx, Ly, count, rng, U=4.0, dtau=0.1):
    """`count` single-slice B = expK diag(d) of a square lattice with random
    fields (flat column-major)."""
    th, _, _ = orc.lattice("square", Lx, Ly)
    eK, _, lam = orc.model(th, U, dtau)
    ev = np.exp(np.array([-lam, lam]) - dtau * U / 2)
    n = th.shape[0]
    out = []
    for _ in range(count):
        s = rng.integers(0, 2, n)
        out.append(np.asfortranarray(eK * ev[s][None, :]).ravel(order="F").copy())
    return out


def udv_chain(orc, Bs, left):
    """UDV of a product of B matrices accumulated by the oracle."""
    n = int(round(np.sqrt(Bs[0].size)))
    u, d, t = np.eye(n).ravel().copy(), np.ones(n), np.eye(n).ravel().copy()
    fn = orc.orc().orc_udv_lmul if left else orc.orc().orc_udv_rmul
    for B in Bs:
        assert fn(n, P(u), P(d), P(t), P(B)) == 0
    return u, d, t


def batch_udv(orc, Lx, Ly, b, nsl, left, seed):
    rng = np.random.default_rng(seed)
    U, D, T = [], [], []
    for _ in range(b):
        u, d, t = udv_chain(orc, bmats(orc, Lx, Ly, nsl, rng), left)
        U.append(u); D.append(d); T.append(t)
    return np.array(U), np.array(D), np.array(T)


def relerr(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("Lx,Ly", [(2, 2), (4, 4), (6, 6), (8, 8), (16, 16), (20, 20)])
def test_udv_lmul_rmul_vs_oracle(orc, eng_mod, Lx, Ly):
    n = Lx * Ly
    b = 3 if n <= 64 else 1
    U, D, T = batch_udv(orc, Lx, Ly, b, 10 if n <= 64 else 4, True, n)
    B = np.array(bmats(orc, Lx, Ly, b, np.random.default_rng(7 + n)))
    for left in (True, False):
        Ug, Dg, Tg = (eng_mod.udv_lmul if left else eng_mod.udv_rmul)(U, D, T, B)
        ofn = orc.orc().orc_udv_lmul if left else orc.orc().orc_udv_rmul
        for k in range(b):
            Uo, Do, To = U[k].copy(), D[k].copy(), T[k].copy()
            assert ofn(n, P(Uo), P(Do), P(To), P(B[k])) == 0
            assert np.allclose(Dg[k], Do, rtol=1e-11, atol=0), (left, k)
            assert np.abs(Ug[k] - Uo).max() < 1e-11
            assert relerr(Tg[k], To) < 1e-11


@pytest.mark.parametrize("n", [40, 96])
def test_udv_lmul_rmul_identity_reflectors(orc, eng_mod, n):
    """Columns already zero below the diagonal give tau = 0 (dlarfg's
    identity reflector, R(j,j) = alpha): an upper-triangular B gives Q = I
    exactly, and a B with only some columns filled below the diagonal mixes
    identity and proper reflectors within one panel."""
    rng = np.random.default_rng(11 + n)
    tri = np.triu(rng.standard_normal((n, n))) + np.diag(np.where(rng.random(n) < 0.5, -1.0, 1.0))
    mixed = tri.copy()
    for j in rng.choice(n, n // 3, replace=False):
        mixed[j + 1:, j] = rng.standard_normal(n - j - 1)
    B = np.array([tri.ravel(order="F"), mixed.ravel(order="F")])
    U = np.array([np.eye(n).ravel()] * 2)
    D = np.ones((2, n))
    T = np.array([np.eye(n).ravel()] * 2)
    for left in (True, False):
        Ug, Dg, Tg = (eng_mod.udv_lmul if left else eng_mod.udv_rmul)(U, D, T, B)
        ofn = orc.orc().orc_udv_lmul if left else orc.orc().orc_udv_rmul
        for k in range(2):
            Uo, Do, To = U[k].copy(), D[k].copy(), T[k].copy()
            assert ofn(n, P(Uo), P(Do), P(To), P(B[k])) == 0
            assert np.allclose(Dg[k], Do, rtol=1e-12, atol=0), (left, k)
            assert np.abs(Ug[k] - Uo).max() < 1e-12, (left, k)
            assert relerr(Tg[k], To) < 1e-12, (left, k)
        if left:
            assert np.array_equal(Ug[0], np.eye(n).ravel())  # all reflectors are the identity


@pytest.mark.parametrize("Lx,Ly,nl,nr", [(4, 4, 16, 24), (6, 6, 16, 24), (8, 8, 16, 24), (12, 12, 6, 6)])
def test_udv_combine_inv_one_plus_vs_oracle(orc, eng_mod, Lx, Ly, nl, nr):
    n = Lx * Ly
    b = 2
    Ul, Dl, Tl = batch_udv(orc, Lx, Ly, b, nl, True, 100 + n)
    Ur, Dr, Tr = batch_udv(orc, Lx, Ly, b, nr, False, 200 + n)
    Uo, Do, To = eng_mod.udv_combine(Ul, Dl, Tl, Ur, Dr, Tr)
    g, sg = eng_mod.udv_inv_one_plus(Uo, Do, To)
    for k in range(b):
        u, d, t = np.zeros(n * n), np.zeros(n), np.zeros(n * n)
        assert orc.orc().orc_udv_combine(n, P(Ul[k]), P(Dl[k]), P(Tl[k]), P(Ur[k]), P(Dr[k]), P(Tr[k]),
                                         P(u), P(d), P(t)) == 0
        assert np.allclose(Do[k], d, rtol=1e-10, atol=0)
        assert np.abs(Uo[k] - u).max() < 1e-10
        go = np.zeros(n * n)
        s = C.c_int()
        assert orc.orc().orc_udv_inv_one_plus(n, P(u), P(d), P(t), P(go), C.byref(s)) == 0
        assert sg[k] == s.value
        assert relerr(g[k], go) < 1e-10


def test_udv_ops_reference_golden(eng_mod):
    """lmul / rmul / combine / inv_one_plus against the reference's outputs."""
    g = np.load(os.path.join(GOLD, "udv_ops_n16.npz"))
    one = lambda a: a[None, :]
    for name, fn in (("lmul", eng_mod.udv_lmul), ("rmul", eng_mod.udv_rmul)):
        u, d, t = fn(one(g["U"]), one(g["D"]), one(g["T"]), one(g["B"]))
        assert np.allclose(d[0], g[f"{name}_D"], rtol=1e-12, atol=0)
        assert np.abs(u[0] - g[f"{name}_U"]).max() < 1e-12
        assert relerr(t[0], g[f"{name}_T"]) < 1e-12
    u, d, t = eng_mod.udv_combine(one(g["U"]), one(g["D"]), one(g["T"]), one(g["U2"]), one(g["D2"]), one(g["T2"]))
    assert np.allclose(d[0], g["comb_D"], rtol=1e-11, atol=0)
    G, s = eng_mod.udv_inv_one_plus(one(g["comb_U"]), one(g["comb_D"]), one(g["comb_T"]))
    assert s[0] == int(g["inv_sign"])
    assert relerr(G[0], g["inv_G"]) < 1e-10


def test_inv_one_plus_extreme_scales(eng_mod):
    """The reference's mpmath fixture: logD in [-645, 706]
    (tests/test_udv_phase6a_fixture.c:97-130): G to 1e-9, exact sign."""
    g = np.load(os.path.join(GOLD, "ref_fixtures.npz"))
    G, s = eng_mod.udv_inv_one_plus(g["U"][None, :], g["D"][None, :], g["T"][None, :])
    assert s[0] == int(g["det_sign"])
    assert relerr(G[0], g["G"]) <= 1e-9


def test_measure_vs_oracle(orc, eng_mod):
    gs = np.load(os.path.join(GOLD, "traj_chain4.npz"))
    th, bp, _ = orc.lattice("chain", 4)
    G = gs["green"][-5:]
    ph, mq = eng_mod.phase_cos(4, 1, "all")
    s5, z, p = eng_mod.measure(4, 1, G, th, bp, 4.0, ph, len(mq))
    for k in range(G.shape[0]):
        assert np.abs(s5[k] - gs["meas"][15 + k]).max() < 1e-12
        assert np.abs(z[k] - gs["szz"][15 + k]).max() < 1e-12
        assert np.abs(p[k] - gs["sperp"][15 + k]).max() < 1e-12
    # 6x6 random-ish G (asymmetric, dense) against the oracle's O(n^2) sums
    rng = np.random.default_rng(5)
    th6, bp6, _ = orc.lattice("square", 6, 6)
    G6 = rng.standard_normal((3, 36 * 36)) * 0.2
    ph6, mq6 = eng_mod.phase_cos(6, 6, "all")
    s5, z, p = eng_mod.measure(6, 6, G6, th6, bp6, 4.0, ph6, len(mq6))
    for k in range(3):
        m, zo, po = orc.measure(G6[k].reshape(36, 36, order="F"), th6, bp6, 4.0, 6, 6)
        assert np.abs(s5[k] - m).max() < 1e-12
        assert np.abs(z[k] - zo).max() < 1e-13
        assert np.abs(p[k] - po).max() < 1e-13
