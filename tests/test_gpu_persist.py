"""The persistent sweep kernel (k_sweep_persist: one workgroup per replica
runs a whole forward sweep with G in LDS, every model of at most 64 sites in
one launch) against the batched path and the exact G.

Both paths follow the reference's accept/reject rule on the same RNG
streams (src/dqmc.c:656-679), so fields, accept counts, signs and RNG
states must be identical after every sweep; G differs only by rounding
order and both must be within 1e-11 of the exact Green's function of the
field (oracle/exact.c, 80-bit extended precision).  The batched path is
itself held to the oracle and the compiled reference's trajectories
(tests/test_gpu_sweep.py), and those tests now run the small lattices on
this kernel by default."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U, DTAU = 4.0, 0.1


def run(eng_mod, monkeypatch, persist, models, nsw, drift=False, guard=0.0, per_sweep=None):
    """models: list of (kind, Lx, Ly, L, stab, seeds); one context, nsw
    sweeps; returns per-sweep snapshots of every replica."""
    monkeypatch.setenv("DQMC_GPU_PERSIST", persist)
    eng = eng_mod.Engine(eng_mod.FP64, drift=drift, drift_guard=guard)
    nrep = 0
    for kind, Lx, Ly, L, stab, seeds in models:
        mid = eng.add_model(kind, Lx, Ly, U, DTAU, L, stab, delay_k=16)
        eng.add_replicas(mid, seeds)
        nrep += len(seeds)
    eng.init()
    snaps = []
    for s in range(nsw):
        if per_sweep:
            per_sweep(eng, s)
        eng.sweep(1)
        st = eng.status()
        snaps.append(dict(st=st, fields=[eng.field(r).copy() for r in range(nrep)],
                          green=[eng.green(r).copy() for r in range(nrep)],
                          rng=[eng.rng(r).copy() for r in range(nrep)]))
    drift_out = eng.drift() if drift else None
    eng.close()
    return snaps, drift_out


def replica_models(models):
    out = []
    for kind, Lx, Ly, L, stab, seeds in models:
        out += [(kind, Lx, Ly, L)] * len(seeds)
    return out


CASES = [
    [("chain", 4, 1, 40, 8, [1, 2])],
    [("square", 4, 4, 20, 5, [7, 8, 9])],
    # n = 36 and 42: tiles past n; L not a multiple of stab (a short last block)
    [("square", 6, 6, 24, 5, [3, 4])],
    [("chain", 42, 1, 22, 6, [5])],
    [("square", 8, 8, 40, 10, [1, 2, 3, 4])],
]


@pytest.mark.parametrize("case", range(len(CASES)))
def test_persist_matches_batched(eng_mod, monkeypatch, exact, case):
    models = CASES[case]
    nsw = 3
    print(f"persistent vs batched: {models}", flush=True)
    a, _ = run(eng_mod, monkeypatch, "1", models, nsw)
    b, _ = run(eng_mod, monkeypatch, "0", models, nsw)
    reps = replica_models(models)
    for s in range(nsw):
        sa, sb = a[s]["st"], b[s]["st"]
        assert (sa["status"] == 0).all() and (sb["status"] == 0).all(), (sa["reason"], sb["reason"])
        assert np.array_equal(sa["accepted"], sb["accepted"]), s
        assert np.array_equal(sa["attempts"], sb["attempts"]), s
        assert np.array_equal(sa["sign"], sb["sign"]), s
        for r, (kind, Lx, Ly, L) in enumerate(reps):
            assert np.array_equal(a[s]["fields"][r], b[s]["fields"][r]), (s, r)
            assert np.array_equal(a[s]["rng"][r], b[s]["rng"][r]), (s, r)
            gx = exact(kind, Lx, Ly, U, DTAU, L, a[s]["fields"][r])
            sc = np.abs(gx).max()
            ea = np.abs(a[s]["green"][r] - gx).max() / sc
            eb = np.abs(b[s]["green"][r] - gx).max() / sc
            print(f"sweep {s} replica {r}: persistent vs exact {ea:.2e}, batched vs exact {eb:.2e}", flush=True)
            assert ea < 1e-11 and eb < 1e-11, (s, r, ea, eb)


def test_persist_uneven_batch_one_launch(eng_mod, monkeypatch, exact):
    # four models of different n, L and stab in one context: one persistent
    # launch per sweep holds all of their replicas, and each model follows
    # the trajectory it follows alone on the batched path
    models = [("chain", 4, 1, 40, 8, [1, 2]), ("square", 4, 4, 20, 10, [3, 4, 5]),
              ("square", 6, 6, 40, 10, [6, 7]), ("square", 8, 8, 40, 10, [8, 9])]
    nsw = 2
    a, _ = run(eng_mod, monkeypatch, "1", models, nsw)
    r0 = 0
    for mdl in models:
        b, _ = run(eng_mod, monkeypatch, "0", [mdl], nsw)
        k = len(mdl[5])
        for s in range(nsw):
            assert np.array_equal(a[s]["st"]["accepted"][r0:r0 + k], b[s]["st"]["accepted"]), (mdl, s)
            for q in range(k):
                assert np.array_equal(a[s]["fields"][r0 + q], b[s]["fields"][q]), (mdl, s, q)
                e = np.abs(a[s]["green"][r0 + q] - b[s]["green"][q]).max() / np.abs(b[s]["green"][q]).max()
                assert e < 1e-11, (mdl, s, q, e)
        r0 += k


def test_persist_drift_rebuild_matches_batched(eng_mod, monkeypatch):
    # drift against the boundary rebuild (DQMC_GPU_DRIFT_REBUILD): a sample
    # at every rebuild on both paths.  The drift is the rounding the wraps
    # accumulate over a block, so its value depends on the order of the
    # sums: the two paths must agree in size (within a factor of 5, measured
    # 0.4 .. 2.5 at 8 x 8, beta = 4, ~1e-9), not in digits.
    models = [("square", 8, 8, 40, 10, [1, 2, 3])]
    a, da = run(eng_mod, monkeypatch, "1", models, 3, drift="rebuild")
    b, db = run(eng_mod, monkeypatch, "0", models, 3, drift="rebuild")
    (ma, sa, ca), (mb, sb, cb) = da, db
    print(f"drift max {ma} vs {mb}, mean {sa} vs {sb}, samples {ca} vs {cb}", flush=True)
    assert np.array_equal(ca, cb) and (ca == 3 * 4).all()
    assert ma.max() < 1e-7 and mb.max() < 1e-7
    assert 0.2 < ma.mean() / mb.mean() < 5.0 and 0.2 < sa.mean() / sb.mean() < 5.0


def test_persist_guard_fails_sticky(eng_mod, monkeypatch):
    # a guard no drift can meet fails every replica at its first rebuild
    # with DQMC_GPU_FAIL_DRIFT; later sweeps leave fields and counters alone
    models = [("square", 6, 6, 24, 6, [1, 2])]
    a, _ = run(eng_mod, monkeypatch, "1", models, 3, drift="rebuild", guard=1e-300)
    for s in range(3):
        st = a[s]["st"]
        assert (st["status"] == 1).all() and (st["reason"] == eng_mod.FAIL_DRIFT).all(), st["reason"]
    for s in (1, 2):
        for r in range(2):
            assert np.array_equal(a[s]["fields"][r], a[0]["fields"][r])
            assert np.array_equal(a[s]["green"][r], a[0]["green"][r])
        assert np.array_equal(a[s]["st"]["attempts"], a[0]["st"]["attempts"])
