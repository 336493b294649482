"""Writes the golden fixtures under tests/golden/ from the compiled reference.

Run in the build container (where /root/reference and oracle/_ref exist):

    python tests/golden/make_golden.py

Fixtures (all numpy .npz, no pickles):
  traj_chain4.npz   chain L=4, U=4, dtau=0.1, Ltr=40, stab=8, seed=1:
                    initial field and RNG state, then 20 sweeps of field,
                    accept counts, G_up and the measurement sample.
  traj_sq8.npz      8x8, U=4, dtau=0.1, Ltr=40, stab=10, seeds 1 and 2:
                    3 sweeps of field and accept counts, final G_up and
                    per-sweep trace / Frobenius norm of G_up.
  traj_alt.npz      sweep_order=alternating (forward, backward, forward, ...):
                    chain L=6 Ltr=20 stab=7 seed=5 and 4x4 Ltr=24 stab=5
                    seed=3 (both with a shorter last block), U=4, dtau=0.1:
                    6 sweeps of field, accept counts, sign and G_up.
  traj_modes.npz    green_rebuild=two_sided / centered trajectories, 4x4: U=8
                    L=80 stab=8 (two_sided; centered; centered alternating)
                    and U=16 L=400 stab=10 (centered, D beyond 1e300).
  model.npz         expK, expKinv, lambda of the reference's model_init for
                    the 4-site chain and the 4x4 square lattice.
  udv_ops_n16.npz   seeded random UDV inputs and the reference's udv_lmul,
                    udv_rmul, udv_combine, udv_inv_one_plus outputs at n=16.
  seeds.npz         replica_seed(base, beta_index, replica) values.
  ref_fixtures.npz  the reference's own test data (tests/fixtures/udv_phase6a_*):
                    the n=16 UDV with logD in [-645, 706] and its mpmath
                    Green's function and sign.
  ed_l4.npz         exact-diagonalisation rows of the 4-site chain at U=4
                    (data/benchmark_L468_U4_full_diag_20260626/reference/).
  c1_observables.txt, c1_szz.txt, c1_sperp.txt
  c2_observables.txt, c2_szz.txt, c2_sperp.txt (8x8, 64 replicas)
                    the reference's output files for the 4-site chain run
                    (nwarm=300 nmeas=3000 nbin=30 seed=1, szz_q=all, sperp_q=all).
  c3_observables.txt, c3_szz.txt, c3_sperp.txt (4x4, beta=3, stab=5, 8
                    replicas, sweep_order=alternating, green_rebuild=two_sided)
                    and its replicas.dat / hopping_used.txt (c3_replicas.txt, c3_hopping.txt)
  c4_*.txt          chain L=6, beta_list=2,3, 2 serial replicas: obs, szz, sperp,
                    su2 (spin_consistency_file), drift (stab_drift_file), reps
                    (replica_log) and hopping_used.txt
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402


def traj_chain4():
    r = pyoracle.RefReplica("chain", 4, 1, 4.0, 0.1, 40, 8, 1)
    f0 = r.field()
    rng0 = r.rng_state()
    fields, accs, gs, meas, szz, sperp = [], [], [], [], [], []
    for _ in range(20):
        assert r.sweep() == 0
        fields.append(r.field())
        accs.append(r.stats()[0])
        gs.append(r.green()[0])
        m, z, p, rc = r.measure()
        assert rc == 0
        meas.append(m)
        szz.append(z)
        sperp.append(p)
    np.savez_compressed(os.path.join(HERE, "traj_chain4.npz"), field0=f0, rng0=rng0, fields=np.array(fields),
                        accepted=np.array(accs, dtype=np.uint64), green=np.array(gs), meas=np.array(meas),
                        szz=np.array(szz), sperp=np.array(sperp))


def traj_sq8():
    out = {}
    for seed in (1, 2):
        r = pyoracle.RefReplica("square", 8, 8, 4.0, 0.1, 40, 10, seed)
        out[f"field0_s{seed}"] = r.field()
        fs, acc, tr, fro = [], [], [], []
        for _ in range(3):
            assert r.sweep() == 0
            g = r.green()[0]
            fs.append(r.field())
            acc.append(r.stats()[0])
            tr.append(np.trace(g))
            fro.append(np.linalg.norm(g))
        out[f"fields_s{seed}"] = np.array(fs)
        out[f"accepted_s{seed}"] = np.array(acc, dtype=np.uint64)
        out[f"trace_s{seed}"] = np.array(tr)
        out[f"fro_s{seed}"] = np.array(fro)
        out[f"green_s{seed}"] = r.green()[0]
    np.savez_compressed(os.path.join(HERE, "traj_sq8.npz"), **out)


ALT_CASES = (("chain", 6, 1, 20, 7, 5), ("square", 4, 4, 24, 5, 3))


def traj_alt():
    out = {}
    for c, (kind, Lx, Ly, L, stab, seed) in enumerate(ALT_CASES):
        r = pyoracle.RefReplica(kind, Lx, Ly, 4.0, 0.1, L, stab, seed, alternating=True)
        fs, acc, sg, gs = [], [], [], []
        for _ in range(6):
            assert r.sweep() == 0
            fs.append(r.field())
            st = r.stats()
            acc.append(st[0])
            sg.append(st[2])
            gs.append(r.green()[0])
        out[f"fields_{c}"] = np.array(fs)
        out[f"accepted_{c}"] = np.array(acc, dtype=np.uint64)
        out[f"sign_{c}"] = np.array(sg)
        out[f"green_{c}"] = np.array(gs)
    np.savez_compressed(os.path.join(HERE, "traj_alt.npz"), **out)


MODE_CASES = (("two_sided", False, 8.0, 80, 8, 3, 4), ("centered", False, 8.0, 80, 8, 3, 4),
              ("centered", True, 8.0, 80, 8, 4, 4), ("centered", False, 16.0, 400, 10, 3, 2))


def traj_modes():
    """green_rebuild=two_sided and centered trajectories of the compiled
    reference, 4x4, dtau=0.1: fields, accept counts, sign and G per sweep."""
    out = {}
    for c, (mode, alt, U, L, stab, seed, nsw) in enumerate(MODE_CASES):
        r = pyoracle.RefReplica("square", 4, 4, U, 0.1, L, stab, seed, alternating=alt, rebuild=mode)
        fs, acc, sg, gs = [], [], [], []
        for _ in range(nsw):
            assert r.sweep() == 0
            fs.append(r.field())
            st = r.stats()
            acc.append(st[0])
            sg.append(st[2])
            gs.append(r.green()[0])
        out[f"fields_{c}"] = np.array(fs)
        out[f"accepted_{c}"] = np.array(acc, dtype=np.uint64)
        out[f"sign_{c}"] = np.array(sg)
        out[f"green_{c}"] = np.array(gs)
    np.savez_compressed(os.path.join(HERE, "traj_modes.npz"), **out)


def model():
    out = {}
    for name, kind, Lx, Ly in (("chain4", "chain", 4, 1), ("sq4", "square", 4, 4)):
        r = pyoracle.RefReplica(kind, Lx, Ly, 4.0, 0.1, 10, 5, 1)
        eK, eKi, th, bp, lam = r.model()
        out[f"{name}_expK"] = eK
        out[f"{name}_expKinv"] = eKi
        out[f"{name}_thop"] = th
        out[f"{name}_bipart"] = bp
        out[f"{name}_lambda"] = lam
    np.savez_compressed(os.path.join(HERE, "model.npz"), **out)


def udv_ops():
    import ctypes as C
    rng = np.random.default_rng(20260603)
    n = 16
    P = lambda a: a.ctypes.data_as(C.c_void_p)
    # realistic factors: products of B_l = expK diag(d_l) of the 4x4 model at
    # U=4, dtau=0.1 with random fields, accumulated by the reference itself
    r = pyoracle.RefReplica("square", 4, 4, 4.0, 0.1, 10, 5, 1)
    eK, _, _, _, lam = r.model()
    ev = np.exp(np.array([-lam, lam]) - 0.1 * 4.0 / 2)

    def bmat():
        s = rng.integers(0, 2, n)
        return np.asfortranarray(eK * ev[s][None, :]).ravel(order="F").copy()

    def chain(nslices, left):
        u, d, t = np.eye(n).ravel(), np.ones(n), np.eye(n).ravel()
        fn = pyoracle.ref().ref_udv_lmul if left else pyoracle.ref().ref_udv_rmul
        for _ in range(nslices):
            assert fn(n, P(u), P(d), P(t), P(bmat())) == 0
        return u, d, t

    U, D, T = chain(12, True)
    B = bmat()
    out = dict(U=U, D=D, T=T, B=B)
    for name, fn in (("lmul", pyoracle.ref().ref_udv_lmul), ("rmul", pyoracle.ref().ref_udv_rmul)):
        u, d, t = U.copy(), D.copy(), T.copy()
        assert fn(n, P(u), P(d), P(t), P(B)) == 0
        out[f"{name}_U"], out[f"{name}_D"], out[f"{name}_T"] = u, d, t
    U2, D2, T2 = chain(20, False)
    out.update(U2=U2, D2=D2, T2=T2)
    uo, do, to = np.zeros(n * n), np.zeros(n), np.zeros(n * n)
    assert pyoracle.ref().ref_udv_combine(n, P(U), P(D), P(T), P(U2), P(D2), P(T2), P(uo), P(do), P(to)) == 0
    out.update(comb_U=uo, comb_D=do, comb_T=to)
    g = np.zeros(n * n)
    s = C.c_int()
    assert pyoracle.ref().ref_udv_inv_one_plus(n, P(uo), P(do), P(to), P(g), C.byref(s)) == 0
    out.update(inv_G=g, inv_sign=np.int32(s.value))
    np.savez_compressed(os.path.join(HERE, "udv_ops_n16.npz"), **out)


def seeds():
    cases = [(1, 0, 0), (1, 0, 1), (12345, 0, 7), (12345, 3, 0), (12345, 3, 127), (2**63 + 5, 1, 64)]
    vals = [pyoracle.ref().ref_replica_seed(b, i, r) for b, i, r in cases]
    np.savez_compressed(os.path.join(HERE, "seeds.npz"), cases=np.array(cases, dtype=np.uint64),
                        values=np.array(vals, dtype=np.uint64))


def ref_fixtures():
    def load_udv(path):
        # "key value" lines are metadata, a lone non-numeric token opens a
        # section, lone numbers are that section's values
        sec, vals, meta = None, {}, {}
        for line in open(path):
            parts = line.split()
            if not parts:
                continue
            if len(parts) == 2:
                meta[parts[0]] = float(parts[1])
                continue
            try:
                v = float(parts[0])
            except ValueError:
                sec = parts[0]
                vals[sec] = []
                continue
            vals[sec].append(v)
        return meta, {k: np.array(v) for k, v in vals.items()}

    m1, v1 = load_udv(os.path.join(REF, "tests/fixtures/udv_phase6a_actual_n16.txt"))
    m2, v2 = load_udv(os.path.join(REF, "tests/fixtures/udv_phase6a_actual_n16_green.txt"))
    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), U=v1["U"], D=v1["D"], T=v1["T"],
                        offset=m1.get("offset", 0.0), min_logD=m1["min_logD"], max_logD=m1["max_logD"],
                        G=v2["G"], det_sign=np.int32(m2["det_sign"]))


def ed_l4():
    a = np.loadtxt(os.path.join(REF, "data/benchmark_L468_U4_full_diag_20260626/reference/L4_U4_FullDiag.dat"))
    np.savez_compressed(os.path.join(HERE, "ed_l4.npz"), columns=np.array(["T", "E_gc", "C", "N", "Sz", "S2",
                                                                           "D_total", "Z"]), data=a)


def c1_outputs():
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "c1.in")
        with open(inp, "w") as f:
            f.write("lattice=chain\nLx=4\nU=4\ndtau=0.1\nbeta_list=4\nnwarm=300\nnmeas=3000\nnbin=30\n"
                    "seed=1\nszz_q=all\nsperp_q=all\n")
        subprocess.run([pyoracle.REF_OMP, inp], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
        for src, dst in (("observables.dat", "c1_observables.txt"), ("szz.dat", "c1_szz.txt"),
                         ("sperp.dat", "c1_sperp.txt")):
            with open(os.path.join(d, src)) as fi, open(os.path.join(HERE, dst), "w") as fo:
                fo.write(fi.read())


def c2_outputs():
    """8x8, beta=4, stab=10, 64 replicas with seeds replica_seed(1, 0, r):
    the reference's OpenMP driver on the host cores."""
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "c2.in")
        with open(inp, "w") as f:
            f.write("lattice=square\nLx=8\nLy=8\nU=4\ndtau=0.1\nbeta_list=4\nstab=10\nnwarm=200\n"
                    "nmeas=2000\nnbin=20\nseed=1\nparallel=omp\nnrep=64\nszz_q=all\nsperp_q=all\n")
        subprocess.run([pyoracle.REF_OMP, inp], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1), OPENBLAS_NUM_THREADS="1"))
        for src, dst in (("observables.dat", "c2_observables.txt"), ("szz.dat", "c2_szz.txt"),
                         ("sperp.dat", "c2_sperp.txt")):
            with open(os.path.join(d, src)) as fi, open(os.path.join(HERE, dst), "w") as fo:
                fo.write(fi.read())


C3_INPUT = ("lattice=square\nLx=4\nLy=4\nU=4\ndtau=0.1\nbeta_list=3\nstab=5\nnwarm=200\nnmeas=2000\nnbin=20\n"
            "seed=3\nparallel=omp\nnrep=8\nsweep_order=alternating\ngreen_rebuild=two_sided\nszz_q=all\nsperp_q=all\n")


def c3_outputs():
    """4x4, beta=3, stab=5, 8 replicas, sweep_order=alternating with the
    two-sided boundary rebuild: the reference's OpenMP driver."""
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "c3.in")
        with open(inp, "w") as f:
            f.write(C3_INPUT)
        subprocess.run([pyoracle.REF_OMP, inp], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 1), OPENBLAS_NUM_THREADS="1"))
        for src, dst in (("observables.dat", "c3_observables.txt"), ("szz.dat", "c3_szz.txt"),
                         ("sperp.dat", "c3_sperp.txt"), ("replicas.dat", "c3_replicas.txt"),
                         ("hopping_used.txt", "c3_hopping.txt")):
            with open(os.path.join(d, src)) as fi, open(os.path.join(HERE, dst), "w") as fo:
                fo.write(fi.read())


C4_INPUT = ("lattice=chain\nLx=6\nU=4\ndtau=0.1\nbeta_list=2,3\nstab=5\nnwarm=100\nnmeas=1000\nnbin=10\nseed=5\n"
            "parallel=serial\nnrep=2\nszz_q=all\nsperp_q=all\nspin_consistency_file=su2.dat\n"
            "stab_drift_file=drift.dat\nreplica_log=reps.dat\noutput_file=obs.dat\n")


def c4_outputs():
    """chain L=6, beta_list=2,3, 2 serial replicas with every optional output
    of the driver: spin_consistency_file, stab_drift_file, replica_log."""
    with tempfile.TemporaryDirectory() as d:
        inp = os.path.join(d, "c4.in")
        with open(inp, "w") as f:
            f.write(C4_INPUT)
        subprocess.run([pyoracle.REF_OMP, inp], cwd=d, check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1"))
        for src in ("obs.dat", "szz.dat", "sperp.dat", "su2.dat", "drift.dat", "reps.dat", "hopping_used.txt"):
            with open(os.path.join(d, src)) as fi, open(os.path.join(HERE, "c4_" + src.replace(".dat", ".txt")),
                                                        "w") as fo:
                fo.write(fi.read())


if __name__ == "__main__":
    if not pyoracle.ref_available():
        sys.exit("oracle/_ref not built: make -C oracle ref")
    gens = [traj_chain4, traj_sq8, traj_alt, traj_modes, model, udv_ops, seeds, ref_fixtures, ed_l4, c1_outputs, c2_outputs,
            c3_outputs, c4_outputs]
    # optional names on the command line select generators (default: all)
    pick = set(sys.argv[1:])
    for g in gens:
        if not pick or g.__name__ in pick:
            g()
    print("golden fixtures written to", HERE)
