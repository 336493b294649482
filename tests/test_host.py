"""CPU tests of the C host layer (sai-qmc_amd/libdqmc_host.so: parser,
lattice, exp(scale*K), replica seeds, jackknife) against the Python mirror,
numpy and the golden seeds of the reference."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "sai-qmc_amd", "libdqmc_host.so")
GOLD = os.path.join(ROOT, "tests", "golden")
P = lambda a: np.ascontiguousarray(a).ctypes.data_as(C.c_void_p)


class Params(C.Structure):
    _fields_ = [("lattice", C.c_char * 16), ("Lx", C.c_int), ("Ly", C.c_int), ("pbc", C.c_int),
                ("thop", C.c_double), ("U", C.c_double), ("dtau", C.c_double),
                ("beta_list", C.c_double * 64), ("nbeta", C.c_int),
                ("nwarm", C.c_int), ("nmeas", C.c_int), ("nbin", C.c_int), ("stab", C.c_int),
                ("output_file", C.c_char * 256), ("parallel", C.c_char * 16), ("nrep", C.c_int),
                ("szz_q", C.c_char * 256), ("szz_file", C.c_char * 256), ("sperp_q", C.c_char * 256),
                ("sperp_file", C.c_char * 256), ("stab_drift_file", C.c_char * 256),
                ("seed", C.c_ulonglong), ("replica_log", C.c_char * 256), ("spin_consistency_file", C.c_char * 256),
                ("sweep_order", C.c_char * 16), ("green_rebuild", C.c_char * 16),
                ("precision", C.c_char * 24), ("delay", C.c_int), ("device", C.c_int),
                ("drift_guard", C.c_double), ("autotune", C.c_int)]


class Lattice(C.Structure):
    _fields_ = [("n", C.c_int), ("Lx", C.c_int), ("Ly", C.c_int), ("is_bipartite", C.c_int),
                ("is_chain", C.c_int), ("t", C.POINTER(C.c_double)), ("bipart", C.POINTER(C.c_int))]


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(LIB):
        pytest.fail("libdqmc_host.so is not built (run __graft_entry__.build())")
    h = C.CDLL(LIB)
    h.replica_seed.restype = C.c_ulonglong
    h.replica_seed.argtypes = [C.c_ulonglong, C.c_int, C.c_int]
    h.expm_sym.argtypes = [C.c_int, C.c_void_p, C.c_double, C.c_void_p]
    h.jackknife.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    h.params_read.argtypes = [C.POINTER(Params), C.c_char_p]
    h.lattice_build.argtypes = [C.POINTER(Lattice), C.POINTER(Params)]
    h.lattice_release.argtypes = [C.POINTER(Lattice)]
    return h


def read(host, tmp_path, text):
    f = tmp_path / "in.txt"
    f.write_text(text)
    p = Params()
    return host.params_read(C.byref(p), str(f).encode()), p


def test_replica_seed_matches_reference(host, eng_mod):
    g = np.load(os.path.join(GOLD, "seeds.npz"))
    for (base, b, r), s in zip(g["cases"], g["values"]):
        assert host.replica_seed(int(base), int(b), int(r)) == int(s)
    for base in (1, 12345, 2**63 + 5):
        for b in range(3):
            for r in range(5):
                assert host.replica_seed(base, b, r) == eng_mod.replica_seed(base, b, r)


def test_expm_sym_against_eigh(host, eng_mod, orc):
    for kind, Lx, Ly in (("chain", 4, 1), ("square", 4, 4), ("square", 6, 6), ("square", 16, 16)):
        th, _, _ = eng_mod.lattice(kind, Lx, Ly)
        n = th.shape[0]
        K = th - 2.0 * np.eye(n)
        for s in (-0.1, 0.1, -0.05):
            out = np.zeros((n, n))
            assert host.expm_sym(n, P(np.asfortranarray(K)), s, P(out)) == 0
            w, V = np.linalg.eigh(K)
            ref = (V * np.exp(s * w)) @ V.T
            assert np.abs(out - ref).max() < 1e-14 * max(1.0, np.abs(ref).max())
        # the oracle's long-double exponential (model_init's expK)
        eK, eKi, _ = orc.model(th, 4.0, 0.1)
        out = np.zeros((n, n))
        assert host.expm_sym(n, P(np.asfortranarray(K)), -0.1, P(out)) == 0
        assert np.abs(out - eK).max() < 1e-14
    # closed form for a 2x2: exp(s [[a, b], [b, a]]) = e^{sa} [[cosh sb, sinh sb], [sinh sb, cosh sb]]
    A = np.array([[0.3, -1.2], [-1.2, 0.3]])
    out = np.zeros((2, 2))
    assert host.expm_sym(2, P(A), 0.7, P(out)) == 0
    ea = np.exp(0.7 * 0.3)
    exact = ea * np.array([[np.cosh(0.7 * -1.2), np.sinh(0.7 * -1.2)], [np.sinh(0.7 * -1.2), np.cosh(0.7 * -1.2)]])
    assert np.abs(out - exact).max() < 5e-15


def test_jackknife(host, orc):
    x = np.random.default_rng(4).standard_normal(41)
    m, e = C.c_double(), C.c_double()
    host.jackknife(P(x), 41, C.byref(m), C.byref(e))
    mo, eo = orc.jackknife(x)
    assert m.value == pytest.approx(mo, rel=1e-15)
    assert e.value == pytest.approx(eo, rel=1e-13)
    assert e.value == pytest.approx(np.sqrt(((x - x.mean()) ** 2).sum() / (41 * 40)), rel=1e-12)


@pytest.mark.parametrize("kind,Lx,Ly,pbc", [("chain", 4, 1, 1), ("chain", 5, 1, 1), ("square", 4, 4, 1),
                                            ("square", 3, 4, 0), ("square", 6, 2, 1)])
def test_lattice_matches_mirror_and_reference(host, eng_mod, orc, tmp_path, kind, Lx, Ly, pbc):
    rc, p = read(host, tmp_path, f"lattice={kind}\nLx={Lx}\nLy={Ly}\npbc={pbc}\n")
    assert rc == 0
    L = Lattice()
    assert host.lattice_build(C.byref(L), C.byref(p)) == 0
    n = L.n
    t = np.ctypeslib.as_array(L.t, shape=(n * n,)).reshape(n, n, order="F").copy()
    bp = np.ctypeslib.as_array(L.bipart, shape=(n,)).copy()
    th, bpm, bip = eng_mod.lattice(kind, Lx, Ly, -1.0, pbc)
    assert np.array_equal(t, th) and np.array_equal(bp, bpm) and bool(L.is_bipartite) == bip
    if pbc:
        tho, bpo, _ = orc.lattice(kind, Lx, Ly)
        assert np.array_equal(t, tho) and np.array_equal(bp, bpo)
    host.lattice_release(C.byref(L))


def test_params_defaults_and_keys(host, tmp_path):
    rc, p = read(host, tmp_path, "lattice=square\nLx=8\nLy=8\nU=4\nbeta_list=2,4.5\nseed=9\nszz_q=all\n"
                                 "precision=mixed\nnrep=3\n")
    assert rc == 0
    assert p.lattice == b"square" and p.Lx == 8 and p.Ly == 8 and p.nbeta == 2
    assert list(p.beta_list[:2]) == [2.0, 4.5]
    assert p.seed == 9 and p.nrep == 3 and p.precision == b"mixed" and p.szz_q == b"all"
    assert p.dtau == 0.1 and p.thop == -1.0 and p.pbc == 1
    assert p.output_file == b"observables.dat" and p.sperp_q == b"none"
    assert p.sweep_order == b"forward" and p.green_rebuild == b"combine"
    assert p.replica_log == b"" and p.spin_consistency_file == b"none"


def test_params_sweep_order_and_rebuild(host, tmp_path):
    rc, p = read(host, tmp_path, "sweep_order=alternating\ngreen_rebuild=two_sided\n")
    assert rc == 0 and p.sweep_order == b"alternating" and p.green_rebuild == b"two_sided"
    rc, p = read(host, tmp_path, "green_rebuild=centered\nreplica_log=none\nspin_consistency_file=d.dat\n")
    assert rc == 0 and p.green_rebuild == b"centered" and p.replica_log == b"none"
    assert p.spin_consistency_file == b"d.dat"


@pytest.mark.parametrize("text", ["lattice=hex\n", "Lx=abc\n", "nosuchkey=1\n", "precision=fp16\n",
                                  "sweep_order=backward\n", "green_rebuild=centered\nprecision=mixed_fp32_chain\n",
                                  "green_rebuild=lu\n",
                                  "dtau=-1\n", "beta_list=\n", "Lx\n"])
def test_params_rejects(host, tmp_path, text):
    rc, _ = read(host, tmp_path, text)
    assert rc != 0


def test_square_propagators_split(host, eng_mod):
    # On a 16 x 16 square lattice the engine replaces the dense B-matrix
    # products by separable ones when expK = E_y (x) E_x (site x + 16 y) to
    # 1e-12 of its largest entry.  The propagators of both host layers (the
    # C program's Taylor expm and the Python layer's eigh) split far inside
    # that; open boundaries split as well, and a lattice with an extra
    # diagonal bond does not (the engine then keeps the dense GEMMs).
    L = 16
    n = L * L

    def split_err(M):
        fx = M[:L, :L]
        fy = np.array([[M[L * r, L * c] / M[0, 0] for c in range(L)] for r in range(L)])
        return np.abs(np.kron(fy, fx) - M).max() / np.abs(M).max()

    for pbc in (1, 0):
        th, _, _ = eng_mod.lattice("square", L, L, pbc=pbc)
        K = th - 2.0 * np.eye(n)
        for s in (-0.1, 0.1):
            out = np.zeros((n, n))
            assert host.expm_sym(n, P(np.asfortranarray(K)), s, P(out)) == 0
            assert split_err(out) < 1e-14
        eK, eKi, _, _ = eng_mod.model(th, 4.0, 0.1)
        assert split_err(eK) < 1e-14 and split_err(eKi) < 1e-14
    th, _, _ = eng_mod.lattice("square", L, L)
    th = th.copy()
    th[0, L + 1] = th[L + 1, 0] = -1.0  # a next-nearest-neighbour bond breaks the split
    out = np.zeros((n, n))
    assert host.expm_sym(n, P(np.asfortranarray(th - 2.0 * np.eye(n))), -0.1, P(out)) == 0
    assert split_err(out) > 1e-6
