"""Failure paths, the drift guard and the drift definitions on the GPU, each
forced from the host so that every kernel exits normally.

Reference behaviour: a replica whose numerics fail gets a sticky status
and reason (LinalgWork.failed / failure_reason, src/linalg.h:48-53;
D->status, src/dqmc.h:79-81) and its later sweeps are no-ops
(src/dqmc.c:611-614; tests/test_dqmc.c:317-334); udv_inv_one_plus_work
refuses an offset whose effective log range underflows
(LINALG_FAILURE_EFFECTIVE_LOG_RANGE, tests/test_udv.c:199-209); the
stabilisation drift is max|G_wrap - G_scratch(tau)| against a from-scratch
G before each rebuild (dqmc_record_stab_drift, src/dqmc.c:64-98)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def run_traj(eng_mod, seeds, nsw, drift=False, guard=0.0, order="forward", nan_at=None, precision=None,
             kind="square", Lx=4, Ly=4, L=20, stab=5, U=4.0):
    eng = eng_mod.Engine(eng_mod.FP64 if precision is None else precision, drift=drift, drift_guard=guard)
    mid = eng.add_model(kind, Lx, Ly, U, 0.1, L, stab, delay_k=8)
    eng.set_sweep_order(mid, order)
    eng.add_replicas(mid, seeds)
    eng.init()
    out = []
    for s in range(nsw):
        if nan_at is not None and s == nan_at[0]:
            g = eng.green(nan_at[1])
            g[0, 0] = np.nan
            eng.set_green(nan_at[1], g)
        eng.sweep(1)
        st = eng.status()
        out.append(dict(st=st, fields=[eng.field(r).copy() for r in range(len(seeds))],
                        green=[eng.green(r).copy() for r in range(len(seeds))]))
    return eng, out


@pytest.mark.parametrize("persist", ["1", "0"])
def test_nonfinite_green_fails_sticky(eng_mod, monkeypatch, persist):
    # a NaN written into one replica's G: the rebuild that ends its block
    # finds the wrapped G non-finite and fails the replica with
    # DQMC_GPU_FAIL_NONFINITE; from then on its sweeps change nothing
    # (fields, accept and attempt counters frozen), and the other replicas
    # follow exactly the trajectory of an undisturbed run.  On the
    # persistent sweep kernel (4 x 4 runs there by default) and on the
    # batched path.
    monkeypatch.setenv("DQMC_GPU_PERSIST", persist)
    seeds = [3, 4, 5]
    print("NaN injection into replica 1 before sweep 1", flush=True)
    eng, bad = run_traj(eng_mod, seeds, 4, nan_at=(1, 1))
    eng.close()
    eng, good = run_traj(eng_mod, seeds, 4)
    eng.close()
    assert (bad[0]["st"]["status"] == 0).all()
    for s in (1, 2, 3):
        st = bad[s]["st"]
        assert st["status"][1] == 1 and st["reason"][1] == eng_mod.FAIL_NONFINITE, (s, st["reason"])
        assert st["status"][0] == 0 and st["status"][2] == 0
    # in the failing sweep the NaN reaches only row and column 0 during
    # slice 0 (site 0's ratio is NaN and rejected; other sites may flip),
    # the flush and the wrap spread it over all of G, every later proposal
    # of the block is rejected, and the first rebuild fails the replica:
    # the field can differ from before the sweep in slice 0 only
    n = 16
    assert np.array_equal(bad[1]["fields"][1][n:], bad[0]["fields"][1][n:])
    for s in (2, 3):
        assert np.array_equal(bad[s]["fields"][1], bad[1]["fields"][1])
        assert bad[s]["st"]["accepted"][1] == bad[1]["st"]["accepted"][1]
        assert bad[s]["st"]["attempts"][1] == bad[1]["st"]["attempts"][1]
    for s in range(4):
        for r in (0, 2):
            assert np.array_equal(bad[s]["fields"][r], good[s]["fields"][r]), (s, r)
            assert np.array_equal(bad[s]["green"][r], good[s]["green"][r]), (s, r)


def test_mixed_precision_needs_a_guard(eng_mod):
    for prec in (eng_mod.MIXED, eng_mod.MIXED_FP32_CHAIN):
        with pytest.raises(RuntimeError, match="drift guard"):
            eng_mod.Engine(prec)


def test_drift_guard_fails_drifting_replicas(eng_mod):
    # 16x16, U=4, beta=8, stab=8 (the metric's workload): the FP32 wrap and
    # delayed updates leave G O(1e-2 .. 10) away from the rebuilt one, so
    # with the guard at 1e-6 every replica fails with DQMC_GPU_FAIL_DRIFT
    # at its first rebuild and stays failed; the FP64 engine on 8x8,
    # beta=4 (drift ~1e-8) never trips the same guard
    print("mixed 16x16 beta=8 with drift_guard=1e-6", flush=True)
    eng, tr = run_traj(eng_mod, list(range(1, 9)), 2, guard=1e-6, precision=eng_mod.MIXED, Lx=16, Ly=16, L=80,
                       stab=8)
    eng.close()
    for s in range(2):
        st = tr[s]["st"]
        assert (st["status"] == 1).all() and (st["reason"] == eng_mod.FAIL_DRIFT).all(), st["reason"]
    for r in range(8):
        assert np.array_equal(tr[1]["fields"][r], tr[0]["fields"][r])
    print("fp64 8x8 beta=4 with drift_guard=1e-6", flush=True)
    eng, tr = run_traj(eng_mod, [1, 2, 3], 3, guard=1e-6, Lx=8, Ly=8, L=40, stab=10)
    eng.close()
    assert all((t["st"]["status"] == 0).all() for t in tr)


def test_boundary_rebuild_matches_scratch_16x16_beta8(eng_mod):
    # at every boundary of a sweep on the metric's workload the drift of
    # the two definitions differs by at most |G_scratch - G_rebuilt|: the
    # two-sided boundary rebuild (prefix and transposed suffix, pre-pivoted
    # chains) and the from-scratch chain agree to 1e-9 while the drift
    # itself is ~1e-4
    seeds = [1, 2, 3, 4]
    out = {}
    for mode in ("scratch", "rebuild"):
        print(f"16x16 beta=8 drift mode {mode}", flush=True)
        eng, tr = run_traj(eng_mod, seeds, 2, drift=mode, Lx=16, Ly=16, L=80, stab=8)
        out[mode] = eng.drift()
        eng.close()
    (a, am, ac), (b, bm, bc) = out["scratch"], out["rebuild"]
    print("scratch max/mean", a, am, "rebuild max/mean", b, bm)
    assert (ac == 20).all() and (bc == 20).all()
    assert (np.abs(a - b) < 1e-9).all() and (np.abs(am - bm) < 1e-9).all()


@pytest.mark.parametrize("order,nsw", [("forward", 3), ("alternating", 4)])
def test_drift_from_scratch(eng_mod, orc, monkeypatch, order, nsw):
    # DQMC_GPU_DRIFT_SCRATCH (the reference's definition) and
    # DQMC_GPU_DRIFT_REBUILD record the same trajectory (G itself is never
    # touched by the drift samples), one sample per rebuild, and values
    # that differ only by the two reference G's rounding (both FP64 G's
    # are within ~1e-12 of the exact G here, the drift ~1e-10 .. 1e-9).
    # Against the oracle's drift (the reference's definition on its own FP64
    # chain) the engine's may only be smaller: the oracle's wraps restart
    # from G rebuilt through the unpivoted chain and the LU-based inverse,
    # 1e-11 .. 1e-10 from the exact G here, and that error is most of its
    # drift (mean samples 5e-9 .. 1e-7 against the engine's 2e-10 .. 3e-10)
    # Both modes on the batched path (the persistent kernel takes the
    # rebuild mode at this size but not the from-scratch one).
    monkeypatch.setenv("DQMC_GPU_PERSIST", "0")
    seeds = [1, 2, 3]
    kind, Lx, Ly, L, stab = "square", 8, 8, 40, 10
    runs = {}
    for mode in ("scratch", "rebuild"):
        print(f"drift mode {mode}, {order}", flush=True)
        eng, tr = run_traj(eng_mod, seeds, nsw, drift=mode, order=order, Lx=Lx, Ly=Ly, L=L, stab=stab)
        runs[mode] = (tr, eng.drift(), eng.drift_at())
        eng.close()
    (ta, da, aa), (tb, db, ab) = runs["scratch"], runs["rebuild"]
    for s in range(nsw):
        for r in range(len(seeds)):
            assert np.array_equal(ta[s]["fields"][r], tb[s]["fields"][r])
            assert np.array_equal(ta[s]["green"][r], tb[s]["green"][r])
    M = (L + stab - 1) // stab
    assert (da[2] == nsw * M).all() and (db[2] == nsw * M).all()
    print("scratch max/mean", da[0], da[1], "rebuild max/mean", db[0], db[1])
    assert (np.abs(da[0] - db[0]) < 1e-10 + 1e-3 * db[0]).all()
    assert (np.abs(da[1] - db[1]) < 1e-10 + 1e-3 * db[1]).all()
    assert ((aa[0] % stab) == 0).all()
    refs = [orc.Replica(kind, Lx, Ly, 4.0, 0.1, L, stab, sd, drift=True, alternating=order == "alternating",
                        delay_cap=8) for sd in seeds]
    for ref in refs:
        for _ in range(nsw):
            assert ref.sweep() == 0
    om = np.array([ref.drift()[1] for ref in refs])
    print("oracle mean", om)
    assert (da[1] < 2.0 * om).all()


def test_batched_inverse_effective_log_range(eng_mod):
    # tests/test_udv.c:173-209: an offset is a scale of D (D e^-2 with
    # offset 2 is the same matrix: same G and sign), and the identity UDV
    # with offset 800 fails with LINALG_FAILURE_EFFECTIVE_LOG_RANGE
    n = 3
    U = np.eye(n)[None].copy()
    T = np.eye(n)[None].copy()
    D = np.array([[-3.0, 0.5, 2.0]])
    g0, s0 = eng_mod.udv_inv_one_plus(U, D, T)
    g1, s1, rs, rc = eng_mod.udv_inv_one_plus_off(U, D * np.exp(-2.0), T, np.array([2.0]))
    assert rc == 0 and rs[0] == 0
    assert s1[0] == s0[0]
    assert np.abs(g1 - g0).max() < 1e-12
    g2, s2, rs, rc = eng_mod.udv_inv_one_plus_off(U, np.ones((1, n)), T, np.array([800.0]))
    assert rc != 0 and rs[0] == 3
