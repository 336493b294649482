"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something it cannot simply restate:
golden vectors written by the compiled reference (tests/golden/*.npz), the
reference's own test fixtures, published test vectors, LAPACK through numpy,
closed forms and invariants.  A deliberately wrong oracle (dropped term, wrong
sign or index, transposed operand) fails at least one of these.
"""
import ctypes as C
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
P = lambda a: a.ctypes.data_as(C.c_void_p)


def gold(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


# ------------------------------------------------------------------ RNG

def test_splitmix_published_vector(orc):
    # Vigna's splitmix64 reference outputs for seed 1234567
    s = np.zeros(4, dtype=np.uint64)
    orc.orc().orc_rng_seed(P(s), 1234567)
    assert [int(x) for x in s] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                   4593380528125082431]


def test_field_init_stream_matches_reference(orc):
    g = gold("traj_chain4.npz")
    r = orc.Replica("chain", 4, 1, 4.0, 0.1, 40, 8, 1)
    # a fresh replica holds the reference's initial field and RNG state
    r2 = orc.Replica("chain", 4, 1, 4.0, 0.1, 40, 8, 1)
    assert np.array_equal(r2.field(), g["field0"])
    assert np.array_equal(r2.rng_state(), g["rng0"])
    del r


def test_replica_seed(orc):
    g = gold("seeds.npz")
    for (b, i, r), v in zip(g["cases"], g["values"]):
        assert orc.orc().orc_replica_seed(int(b), int(i), int(r)) == int(v)


# --------------------------------------------------------------- model

def test_lattice_and_model_match_reference(orc):
    g = gold("model.npz")
    for name, kind, Lx, Ly in (("chain4", "chain", 4, 1), ("sq4", "square", 4, 4)):
        th, bp, bip = orc.lattice(kind, Lx, Ly)
        assert bip
        assert np.array_equal(th, g[f"{name}_thop"])
        assert np.array_equal(bp, g[f"{name}_bipart"])
        eK, eKi, lam = orc.model(th, 4.0, 0.1)
        assert np.abs(eK - g[f"{name}_expK"]).max() < 1e-14
        assert np.abs(eKi - g[f"{name}_expKinv"]).max() < 1e-14
        assert lam == pytest.approx(float(g[f"{name}_lambda"]), rel=1e-15)
        assert np.abs(eK @ eKi - np.eye(th.shape[0])).max() < 1e-13


def test_expm_closed_form_2x2(orc):
    # exp(s [[a, b], [b, a]]) = e^{sa} [[cosh sb, sinh sb], [sinh sb, cosh sb]]
    a, b, s = 0.7, -1.3, 0.4
    A = np.array([[a, b], [b, a]])
    out = np.zeros(4)
    orc.orc().orc_expm_sym(2, P(np.asfortranarray(A)), s, P(out))
    e = np.exp(s * a) * np.array([[np.cosh(s * b), np.sinh(s * b)], [np.sinh(s * b), np.cosh(s * b)]])
    assert np.abs(out.reshape(2, 2, order="F") - e).max() < 1e-15


def test_lambda_identity(orc):
    # cosh(lambda) = exp(dtau U / 2) (tests/test_field.c:13)
    _, _, lam = orc.model(np.zeros((2, 2)), 6.0, 0.05)
    assert np.cosh(lam) == pytest.approx(np.exp(0.05 * 6.0 / 2), rel=1e-15)


# ------------------------------------------------------- linear algebra

@pytest.mark.parametrize("n", [1, 5, 16, 33])
def test_qr_matches_lapack_convention(orc, n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    a = np.asfortranarray(A).ravel(order="F").copy()
    tau = np.zeros(n)
    orc.orc().orc_geqrf(n, P(a), P(tau))
    R = np.triu(a.reshape(n, n, order="F"))
    Rl = np.linalg.qr(A, mode="r")  # LAPACK dgeqrf: same reflector signs
    assert np.abs(R - Rl).max() < 1e-12 * max(1.0, np.abs(Rl).max())
    orc.orc().orc_orgqr(n, P(a), P(tau))
    Q = a.reshape(n, n, order="F")
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    assert np.abs(Q @ R - A).max() < 1e-12


def test_inverse_and_trtri(orc):
    rng = np.random.default_rng(3)
    n = 20
    A = rng.standard_normal((n, n))
    Ai = np.zeros(n * n)
    assert orc.orc().orc_inverse(n, P(np.asfortranarray(A)), P(Ai)) == 0
    assert np.abs(Ai.reshape(n, n, order="F") - np.linalg.inv(A)).max() < 1e-11
    T = np.triu(rng.standard_normal((n, n)) * 0.4, 1) + np.diag(rng.standard_normal(n))
    t = np.asfortranarray(T).ravel(order="F").copy()
    orc.orc().orc_trtri_unit_upper(n, P(t))
    Tu = np.triu(T, 1) + np.eye(n)  # the diagonal is taken as 1
    assert np.abs(t.reshape(n, n, order="F") - np.linalg.inv(Tu)).max() < 1e-12


def test_udv_ops_match_reference(orc):
    g = gold("udv_ops_n16.npz")
    n = 16
    for name, fn in (("lmul", orc.orc().orc_udv_lmul), ("rmul", orc.orc().orc_udv_rmul)):
        U, D, T = g["U"].copy(), g["D"].copy(), g["T"].copy()
        assert fn(n, P(U), P(D), P(T), P(g["B"].copy())) == 0
        assert np.allclose(D, g[f"{name}_D"], rtol=1e-12, atol=0)
        assert np.abs(U - g[f"{name}_U"]).max() < 1e-12
        assert np.abs(T - g[f"{name}_T"]).max() < 1e-12 * np.abs(g[f"{name}_T"]).max()
    uo, do, to = np.zeros(n * n), np.zeros(n), np.zeros(n * n)
    assert orc.orc().orc_udv_combine(n, P(g["U"]), P(g["D"]), P(g["T"]), P(g["U2"]), P(g["D2"]), P(g["T2"]),
                                     P(uo), P(do), P(to)) == 0
    assert np.allclose(do, g["comb_D"], rtol=1e-11, atol=0)
    assert np.abs(uo - g["comb_U"]).max() < 1e-11
    gg = np.zeros(n * n)
    s = C.c_int()
    assert orc.orc().orc_udv_inv_one_plus(n, P(uo), P(do), P(to), P(gg), C.byref(s)) == 0
    assert s.value == int(g["inv_sign"])
    assert np.abs(gg - g["inv_G"]).max() / np.abs(g["inv_G"]).max() < 1e-11


def test_inv_one_plus_reference_fixture(orc):
    """The reference's mpmath fixture: logD in [-645, 706], G to 1e-9
    relative, exact sign (tests/test_udv_phase6a_fixture.c:97-130)."""
    g = gold("ref_fixtures.npz")
    n = 16
    assert float(g["max_logD"]) - float(g["min_logD"]) == pytest.approx(1350.7731737311086, abs=1e-9)
    G = np.zeros(n * n)
    s = C.c_int()
    assert orc.orc().orc_udv_inv_one_plus(n, P(g["U"].copy()), P(g["D"].copy()), P(g["T"].copy()), P(G),
                                          C.byref(s)) == 0
    assert s.value == int(g["det_sign"])
    assert np.abs(G - g["G"]).max() / np.abs(g["G"]).max() <= 1e-9


def test_inv_one_plus_dense_formula(orc):
    rng = np.random.default_rng(11)
    n = 12
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    T = np.triu(rng.standard_normal((n, n)) * 0.3, 1) + np.eye(n)
    D = np.sign(rng.standard_normal(n)) * np.exp(rng.uniform(-2, 2, n))
    G = np.zeros(n * n)
    s = C.c_int()
    assert orc.orc().orc_udv_inv_one_plus(n, P(np.asfortranarray(q)), P(D), P(np.asfortranarray(T)), P(G),
                                          C.byref(s)) == 0
    Mx = np.eye(n) + q @ np.diag(D) @ T
    assert np.abs(G.reshape(n, n, order="F") - np.linalg.inv(Mx)).max() < 1e-12
    assert s.value == int(np.sign(np.linalg.det(Mx)))


# ------------------------------------------------------------- sweeps

def test_trajectory_chain4_matches_reference(orc):
    g = gold("traj_chain4.npz")
    r = orc.Replica("chain", 4, 1, 4.0, 0.1, 40, 8, 1)
    th, bp = r.thop, r.bipart
    for s in range(20):
        assert r.sweep() == 0
        assert np.array_equal(r.field(), g["fields"][s]), s
        assert r.stats()[0] == int(g["accepted"][s])
        gu = r.green()[0]
        assert np.abs(gu - g["green"][s]).max() < 1e-12
        m, z, p = orc.measure(gu, th, bp, 4.0, 4, 1)
        assert np.abs(m - g["meas"][s]).max() < 1e-11
        assert np.abs(z - g["szz"][s]).max() < 1e-12
        assert np.abs(p - g["sperp"][s]).max() < 1e-12


def test_trajectory_8x8_matches_reference(orc):
    g = gold("traj_sq8.npz")
    for seed in (1, 2):
        r = orc.Replica("square", 8, 8, 4.0, 0.1, 40, 10, seed)
        assert np.array_equal(r.field(), g[f"field0_s{seed}"])
        for s in range(3):
            assert r.sweep() == 0
            assert np.array_equal(r.field(), g[f"fields_s{seed}"][s])
            assert r.stats()[0] == int(g[f"accepted_s{seed}"][s])
            gu = r.green()[0]
            assert np.trace(gu) == pytest.approx(float(g[f"trace_s{seed}"][s]), rel=1e-11)
            assert np.linalg.norm(gu) == pytest.approx(float(g[f"fro_s{seed}"][s]), rel=1e-11)
        ref = g[f"green_s{seed}"]
        assert np.abs(r.green()[0] - ref).max() / np.abs(ref).max() < 1e-10


def test_alternating_trajectory_matches_reference(orc):
    # sweep_order=alternating (src/dqmc.c:815-963): forward sweeps that
    # store the prefix stack, backward sweeps that store the suffix stack,
    # against the compiled reference's own trajectory
    g = gold("traj_alt.npz")
    cases = (("chain", 6, 1, 20, 7, 5), ("square", 4, 4, 24, 5, 3))
    for c, (kind, Lx, Ly, L, stab, seed) in enumerate(cases):
        r = orc.Replica(kind, Lx, Ly, 4.0, 0.1, L, stab, seed, alternating=True)
        for s in range(6):
            assert r.sweep() == 0
            assert np.array_equal(r.field(), g[f"fields_{c}"][s]), (c, s)
            acc, _, sign, _ = r.stats()
            assert acc == int(g[f"accepted_{c}"][s]) and sign == g[f"sign_{c}"][s]
            ref = g[f"green_{c}"][s]
            assert np.abs(r.green()[0] - ref).max() / np.abs(ref).max() < 1e-11


def test_backward_wrap_inverts_forward_wrap(orc):
    # green_wrap_backward at slice l undoes green_wrap at slice l:
    # B_l^-1 (B_l G B_l^-1) B_l = G, and it is not the forward wrap again
    rep = orc.Replica("square", 4, 4, 4.0, 0.1, 12, 4, 9)
    eK, eKi, lam = rep.model()
    n = 16
    ev = np.zeros(2)
    orc.orc().orc_expv(lam, 4.0, 0.1, 1.0, orc._ptr(ev))
    s = rep.field()[3 * n:4 * n].copy()
    g0 = np.asfortranarray(np.random.default_rng(4).standard_normal((n, n)))
    g = g0.copy(order="F")
    args = (n, orc._ptr(np.asfortranarray(eK)), orc._ptr(np.asfortranarray(eKi)), orc._ptr(ev), orc._ptr(s))
    orc.orc().orc_wrap(*args, orc._ptr(g))
    fwd = g.copy(order="F")
    orc.orc().orc_wrap_backward(*args, orc._ptr(g))
    assert np.abs(g - g0).max() < 1e-12
    orc.orc().orc_wrap_backward(*args, orc._ptr(g))
    B = eK * ev[(s > 0).astype(int)][None, :]
    assert np.abs(g - np.linalg.solve(B, g0) @ B).max() < 1e-11
    assert np.abs(g - fwd).max() > 1e-3


def test_alternating_green_is_the_rebuilt_green(orc):
    # after every sweep, either direction, G = (1 + B_{L-1} ... B_0)^-1 of
    # the current field (the end-of-sweep rebuild, src/green.c:352-370)
    r = orc.Replica("square", 4, 4, 4.0, 0.1, 20, 6, 21, alternating=True)
    q = orc.Replica("square", 4, 4, 4.0, 0.1, 20, 6, 21)
    for s in range(4):
        assert r.sweep() == 0
        q.set_field(r.field())
        assert q.rebuild() == 0
        a, b = r.green()[0], q.green()[0]
        assert np.abs(a - b).max() < 1e-12
    # the backward legs visit the slices in the other order: the trajectory
    # differs from the forward sweep's
    f = orc.Replica("square", 4, 4, 4.0, 0.1, 20, 6, 21)
    for _ in range(2):
        f.sweep()
    assert not np.array_equal(f.field(), r.field())


MODE_CASES = (("two_sided", False, 8.0, 80, 8, 3, 4), ("centered", False, 8.0, 80, 8, 3, 4),
              ("centered", True, 8.0, 80, 8, 4, 4), ("centered", False, 16.0, 400, 10, 3, 2))


def test_rebuild_modes_match_reference(orc):
    # green_rebuild=two_sided and centered (src/green.c:308-345,
    # src/linalg.c:453-794, 1117-1389) against the compiled reference's
    # trajectories; the last case (U=16, beta=40) carries D beyond 1e300
    g = gold("traj_modes.npz")
    for c, (mode, alt, U, L, stab, seed, nsw) in enumerate(MODE_CASES):
        r = orc.Replica("square", 4, 4, U, 0.1, L, stab, seed, alternating=alt, rebuild=mode)
        for s in range(nsw):
            assert r.sweep() == 0
            assert np.array_equal(r.field(), g[f"fields_{c}"][s]), (c, s)
            acc, _, sign, _ = r.stats()
            assert acc == int(g[f"accepted_{c}"][s]) and sign == g[f"sign_{c}"][s]
            ref = g[f"green_{c}"][s]
            assert np.abs(r.green()[0] - ref).max() / np.abs(ref).max() < 1e-9, (c, s)


def _udv_args(orc, U, D, T):
    return orc._ptr(U), orc._ptr(D), orc._ptr(T)


def test_log_offset_is_a_scale_of_d(orc):
    # tests/test_udv.c:173-196: a UDV whose D is scaled by e^-2 with
    # log_offset 2 is the same matrix: same G and sign
    O = orc.orc()
    rng = np.random.default_rng(11)
    n = 3
    U, _ = np.linalg.qr(rng.standard_normal((n, n)))
    U = np.asfortranarray(U)
    T = np.asfortranarray(np.triu(rng.standard_normal((n, n)), 1) + np.eye(n))
    D = np.array([-3.0, 0.5, 2.0])
    g0, g1 = np.zeros((n, n), order="F"), np.zeros((n, n), order="F")
    s0, s1 = orc.C.c_int(), orc.C.c_int()
    assert O.orc_udv_inv_one_plus_off(n, *_udv_args(orc, U, D, T), 0.0, orc._ptr(g0), orc.C.byref(s0)) == 0
    Ds = D * np.exp(-2.0)
    assert abs(Ds[0]) < 1.0
    assert O.orc_udv_inv_one_plus_off(n, *_udv_args(orc, U, Ds, T), 2.0, orc._ptr(g1), orc.C.byref(s1)) == 0
    assert s0.value == s1.value and np.abs(g0 - g1).max() < 1e-12
    assert np.abs(g0 - np.linalg.inv(np.eye(n) + U @ np.diag(D) @ T)).max() < 1e-12
    # the two-sided inverse with offsets on both sides is the same matrix too
    g2, s2 = np.zeros((n, n), order="F"), orc.C.c_int()
    I = np.asfortranarray(np.eye(n))
    assert O.orc_udv_inv_two_sided(n, *_udv_args(orc, U, Ds, T), 2.0, *_udv_args(orc, I, np.ones(n), I), 0.0,
                                   orc._ptr(g2), orc.C.byref(s2)) == 0
    assert s2.value == s0.value and np.abs(g2 - g0).max() < 1e-12
    # log_offset = 800 leaves no representable split (tests/test_udv.c:198-208)
    eye = np.asfortranarray(np.eye(n))
    assert O.orc_udv_inv_one_plus_off(n, *_udv_args(orc, eye, np.ones(n), eye), 800.0, orc._ptr(g0),
                                      orc.C.byref(s0)) == 3
    assert s0.value == 0


def test_recenter_and_update_margin(orc):
    O = orc.orc()
    n = 4
    # recentring keeps D e^offset and centres log|D| (tests/test_udv_stack.c:196-212)
    D = np.array([np.exp(30.0), -np.exp(-10.0), 2.0, 1e-3])
    off = orc.C.c_double(1.5)
    Dc = D.copy()
    assert O.orc_udv_recenter(n, orc._ptr(Dc), orc.C.byref(off)) == 0
    ld = np.log(np.abs(Dc))
    assert abs(ld.min() + ld.max()) < 1e-9 and off.value == pytest.approx(1.5 + 10.0, abs=1e-12)
    assert np.allclose(Dc * np.exp(off.value - 1.5), D, rtol=1e-12) and np.array_equal(np.sign(Dc), np.sign(D))
    # an update that would leave the double range fails with the inputs
    # untouched (tests/test_udv_stack.c:214-252)
    for fn in (O.orc_udv_lmul_centered, O.orc_udv_rmul_centered):
        U = np.asfortranarray(np.eye(n))
        T = np.asfortranarray(np.eye(n))
        Dw = np.array([np.exp(700.0), np.exp(-700.0), 1.0, 1.0])
        amp = np.asfortranarray(np.exp(2.0) * np.eye(n))
        o = orc.C.c_double(0.0)
        assert fn(n, *_udv_args(orc, U, Dw, T), orc.C.byref(o), orc._ptr(amp)) == 2
        assert np.array_equal(U, np.eye(n)) and np.array_equal(T, np.eye(n)) and o.value == 0.0
        assert np.array_equal(Dw, [np.exp(700.0), np.exp(-700.0), 1.0, 1.0])


def test_u0_green_closed_form(orc):
    # U = 0: B_l = expK for every slice, G = (1 + exp(-beta K))^-1
    # (tests/test_green_init.c:29-46)
    L, dtau = 30, 0.1
    r = orc.Replica("square", 4, 4, 0.0, dtau, L, 6, 5)
    th = r.thop
    w, V = np.linalg.eigh(th)
    Gx = (V * (1.0 / (1.0 + np.exp(-L * dtau * w)))) @ V.T
    assert np.abs(r.green()[0] - Gx).max() < 1e-12


@pytest.mark.parametrize("kind,Lx,Ly,L", [("chain", 4, 1, 10), ("square", 4, 4, 10), ("square", 4, 4, 20)])
def test_ratio_log_is_the_determinant_ratio(orc, kind, Lx, Ly, L):
    # The logged ratio of each proposal of slice 0 is the ratio of the
    # determinants det(1 + B_{L-1} .. B_0) after and before the flip, for
    # both spin species, with B_l = expK diag(expv_sigma[s_l]) built
    # directly from the field (src/green.c:63-67, 126-137): the definition,
    # independent of the R = (1 + (1 - G_ii) N_up)(1 + G_ii N_dn) form and of
    # the PH map it uses.  Slice 0 is final after one forward sweep, so the
    # field before proposal i is the initial one with sites < i taken from
    # the swept one.  beta <= 2 keeps cond(1 + B..B) small enough for FP64
    # determinants to 1e-9.
    U, dtau, stab = 4.0, 0.1, 5
    r = orc.Replica(kind, Lx, Ly, U, dtau, L, stab, 11, ratio_log=True)
    n = r.n
    s0 = r.field().reshape(L, n).copy()
    assert r.sweep() == 0
    s1 = r.field().reshape(L, n)
    R = r.ratios()[0]
    eK, _, lam = r.model()
    ev = {}
    for sg in (1.0, -1.0):
        e = np.zeros(2)
        orc.orc().orc_expv(lam, U, dtau, sg, e.ctypes.data_as(C.c_void_p))
        ev[sg] = e

    def det_ratio(field, i):
        out = 1.0
        for sg in (1.0, -1.0):
            vals = []
            for f in (field, None):
                g = field.copy()
                if f is None:
                    g[0, i] = -g[0, i]
                P = np.eye(n)
                for l in range(L):
                    P = (eK * ev[sg][(g[l] > 0).astype(int)][None, :]) @ P
                sign, logdet = np.linalg.slogdet(np.eye(n) + P)
                vals.append((sign, logdet))
            out *= vals[1][0] * vals[0][0] * np.exp(vals[1][1] - vals[0][1])
        return out

    field = s0.copy()
    for i in range(n):
        x = det_ratio(field, i)
        assert abs(R[i] - x) < 1e-9 * max(1.0, abs(x)), (i, R[i], x)
        field[0, i] = s1[0, i]
    assert np.array_equal(field, np.vstack([s1[:1], s0[1:]]))


def test_drift_bound(orc):
    # tests/test_dqmc_stab_drift.c:53-58 (PH case): Ltr=8, stab=4, U=4,
    # 2 sweeps, 2 samples per sweep, max drift < 1e-8
    r = orc.Replica("chain", 4, 1, 4.0, 0.1, 8, 4, 777, drift=True)
    for _ in range(2):
        assert r.sweep() == 0
    mx, mn, cnt = r.drift()
    assert cnt == 4
    assert 0.0 <= mn <= mx < 1e-8


# ----------------------------------------------------------- measurement

def test_sum_rules_and_free_fermion(orc):
    # sum_q Szz(q) = 1/4 (ntot - 2 n doublon), sum_q Sperp = 1/2 (...)
    # (src/structure_factor.c:472-474); at U=0 Sperp = 2 Szz
    # (tests/test_integration.c:75, 152)
    g = gold("traj_chain4.npz")
    r = orc.Replica("chain", 4, 1, 4.0, 0.1, 40, 8, 1)
    gu = g["green"][-1]
    m, z, p = orc.measure(gu, r.thop, r.bipart, 4.0, 4, 1)
    rhs = m[3] - 2 * 4 * m[4]
    assert z.sum() == pytest.approx(0.25 * rhs, abs=1e-12)
    assert p.sum() == pytest.approx(0.5 * rhs, abs=1e-12)
    r0 = orc.Replica("square", 4, 4, 0.0, 0.1, 20, 5, 3)
    m0, z0, p0 = orc.measure(r0.green()[0], r0.thop, r0.bipart, 0.0, 4, 4)
    assert np.abs(p0 - 2 * z0).max() < 1e-12


def test_jackknife_is_standard_error(orc):
    x = np.random.default_rng(1).standard_normal(37)
    m, e = orc.jackknife(x)
    assert m == pytest.approx(x.mean(), rel=1e-14)
    assert e == pytest.approx(np.sqrt(((x - x.mean()) ** 2).sum() / (37 * 36)), rel=1e-12)


@pytest.mark.skipif(not os.path.exists("/root/reference/src"), reason="compiled reference only in the build container")
def test_oracle_vs_live_reference_4x4(orc):
    # more seeds than the fixtures hold, straight against the compiled reference
    for seed in (11, 12):
        a = orc.Replica("square", 4, 4, 6.0, 0.125, 32, 6, seed)
        b = orc.RefReplica("square", 4, 4, 6.0, 0.125, 32, 6, seed)
        for _ in range(4):
            a.sweep()
            b.sweep()
            assert np.array_equal(a.field(), b.field())
            ga, gb = a.green()[0], b.green()[0]
            assert np.abs(ga - gb).max() / np.abs(gb).max() < 1e-11


@pytest.mark.skipif(not os.path.exists("/root/reference/src"), reason="compiled reference only in the build container")
def test_oracle_vs_live_reference_alternating(orc):
    for kind, Lx, Ly, L, stab, seed in (("square", 4, 4, 32, 6, 13), ("chain", 8, 1, 40, 8, 14)):
        a = orc.Replica(kind, Lx, Ly, 6.0, 0.125, L, stab, seed, alternating=True)
        b = orc.RefReplica(kind, Lx, Ly, 6.0, 0.125, L, stab, seed, alternating=True)
        for _ in range(5):
            a.sweep()
            b.sweep()
            assert np.array_equal(a.field(), b.field())
            ga, gb = a.green()[0], b.green()[0]
            assert np.abs(ga - gb).max() / np.abs(gb).max() < 1e-11


# ------------------------------------------------ extended-precision G

def test_exact_green_matches_mpmath(orc):
    # oracle/exact.c against the direct formula (1 + B..B)^-1 in mpmath at
    # 50 digits (tests/golden/make_exact.py): chain 4 (beta=4), 4x4 at tau=0
    # and tau=10, 6x6, 8x8 (beta=4); the long double's own rounding, grown to
    # ~2e-15 over 40 single-slice QR steps at 8x8, and the FP64 rounding of
    # the stored values are the only differences
    d = gold("exact_g.npz")
    for k in range(int(d["ncases"])):
        U, dtau, L, tau, lam = d[f"case{k}_params"]
        G = d[f"case{k}_G"]
        for blk in (1, 5, 8):
            g = orc.green_exact(d[f"case{k}_expK"], lam, U, dtau, d[f"case{k}_field"], int(L), int(tau), blk)
            assert np.abs(g - G).max() / np.abs(G).max() < 1e-14, (k, blk)


def test_exact_green_u0_and_shifted_tau(orc):
    # U = 0: every B_l = expK, G(tau) = (1 + exp(-beta K))^-1 for all tau
    # (tests/test_green_init.c:29-46), here on 8x8 at beta = 4, where the
    # closed form by eigh is good to ~1e-15
    L, dtau = 40, 0.1
    r = orc.Replica("square", 8, 8, 0.0, dtau, L, 10, 3)
    eK, _, lam = r.model()
    K = r.thop - 0.0 * np.eye(64)
    w, V = np.linalg.eigh(K)
    Gx = (V * (1.0 / (1.0 + np.exp(-L * dtau * w)))) @ V.T
    for tau in (0, 17):
        g = orc.green_exact(eK, lam, 0.0, dtau, r.field(), L, tau, 10)
        assert np.abs(g - Gx).max() < 1e-13, tau


def test_exact_green_wraps_like_g(orc):
    # G(tau + 1) = B_tau G(tau) B_tau^-1 exactly (src/green.c:509-527): the
    # extended-precision G at two neighbouring slices related by the FP64
    # wrap of the oracle, on an 8x8 field at beta = 2
    n, L = 64, 20
    r = orc.Replica("square", 8, 8, 4.0, 0.1, L, 5, 11)
    r.sweep()
    eK, eKi, lam = r.model()
    f = r.field()
    g3 = orc.green_exact(eK, lam, 4.0, 0.1, f, L, 3, 5)
    g4 = orc.green_exact(eK, lam, 4.0, 0.1, f, L, 4, 5)
    ev = np.zeros(2)
    orc.orc().orc_expv(lam, 4.0, 0.1, 1.0, P(ev))
    g = np.asfortranarray(g3.copy())
    orc.orc().orc_wrap(n, P(np.asfortranarray(eK)), P(np.asfortranarray(eKi)), P(ev),
                       P(np.ascontiguousarray(f[3 * n:4 * n])), P(g))
    assert np.abs(g - g4).max() / np.abs(g4).max() < 1e-12
