"""ctypes binding of libdqmc_gpu.so (include/dqmc_gpu.h) plus the host-side
model set-up that the reference does before it reaches the engine.

The host side mirrors the reference's per-replica order
(dqmc_run_replica, src/replica_run.c:89-477): lattice (src/lattice.c),
model_init with mu = U/2 (src/model.c:8-31), rng_seed + field_init
(src/rng.c, src/field.c), the initial rebuild, warm-up and measured sweeps,
bins and acceptance deltas (src/replica_run.c:284-451).  Everything from the
initial rebuild on runs in the HIP library; this module never computes a
Green's function itself, and it raises if the library is missing.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdqmc_gpu.so")

FP64 = 0
MIXED = 1
MIXED_FP32_CHAIN = 2
STEP_PROPOSE_SLICE = 0
STEP_WRAP = 1
STEP_REBUILD = 2
SWEEP_FORWARD = 0
SWEEP_ALTERNATING = 1
REBUILD = {"combine": 0, "two_sided": 1, "centered": 2}

_lib = None

# Every entry point of include/dqmc_gpu.h with its ctypes signature.
_P = C.c_void_p
_I = C.c_int
_D = C.c_double
SIGNATURES = {
    "dqmc_gpu_create": (_I, [C.POINTER(_P), _I, _P]),
    "dqmc_gpu_add_model": (_I, [_P, _I, _I, _I, _I, _P, _P, _P, _P, _D, _D, _D, _D, _I, _I, C.POINTER(_I)]),
    "dqmc_gpu_add_replicas": (_I, [_P, _I, _I, _P, _P, C.POINTER(_I)]),
    "dqmc_gpu_set_measure": (_I, [_P, _I, _I, _P, _I]),
    "dqmc_gpu_set_sweep_order": (_I, [_P, _I, _I]),
    "dqmc_gpu_set_rebuild": (_I, [_P, _I, _I]),
    "dqmc_gpu_init": (_I, [_P]),
    "dqmc_gpu_sweep": (_I, [_P, _I, _I]),
    "dqmc_gpu_run": (_I, [_P, _I, _I, _I]),
    "dqmc_gpu_read_bins": (_I, [_P, _I, _P, _P, _P, _P]),
    "dqmc_gpu_read_status": (_I, [_P, _P, _P, _P, _P, _P]),
    "dqmc_gpu_read_drift": (_I, [_P, _P, _P, _P]),
    "dqmc_gpu_read_drift_at": (_I, [_P, _P, _P]),
    "dqmc_gpu_read_tuning": (_I, [_P, _I, _P, _P, _P]),
    "dqmc_gpu_read_ratios": (_I, [_P, _I, _P]),
    "dqmc_gpu_read_green": (_I, [_P, _I, _P]),
    "dqmc_gpu_write_green": (_I, [_P, _I, _P]),
    "dqmc_gpu_read_field": (_I, [_P, _I, _P]),
    "dqmc_gpu_write_field": (_I, [_P, _I, _P]),
    "dqmc_gpu_read_rng": (_I, [_P, _I, _P]),
    "dqmc_gpu_step": (_I, [_P, _I, _I]),
    "dqmc_gpu_last_elapsed_ms": (_I, [_P, C.POINTER(_D)]),
    "dqmc_gpu_profile_sweep": (_I, [_P, _I, _P, _P]),
    "dqmc_gpu_destroy": (None, [_P]),
    "dqmc_gpu_udv_lmul": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "dqmc_gpu_udv_rmul": (_I, [_I, _I, _I, _P, _P, _P, _P]),
    "dqmc_gpu_udv_combine": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "dqmc_gpu_udv_inv_one_plus": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "dqmc_gpu_udv_inv_one_plus_off": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "dqmc_gpu_measure": (_I, [_I, _I, _I, _I, _P, _P, _P, _D, _I, _P, _P, _P, _P]),
    "dqmc_gpu_udv_inv_one_plus_noalloc": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "dqmc_gpu_last_error": (C.c_char_p, []),
}


DRIFT = {False: 0, None: 0, "off": 0, True: 1, "scratch": 1, "rebuild": 2}
FAIL_NONFINITE = 16
FAIL_SIGN = 18
FAIL_DRIFT = 19


class _Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("use_graph", C.c_int), ("drift", C.c_int), ("autotune", C.c_int),
                ("drift_guard", C.c_double)]


def lib():
    """Load the HIP engine; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine not built: {LIB_PATH} missing (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {lib().dqmc_gpu_last_error().decode()}")


# ------------------------------------------------------------ host set-up

def lattice(kind: str, Lx: int, Ly: int = 1, t: float = -1.0, pbc: int = 1):
    """Hopping matrix and sublattice signs (lattice_chain / lattice_square,
    src/lattice.c:26-82).  Returns (thop[n,n] column-major-equal symmetric,
    bipart[n], is_bipartite)."""
    if kind == "chain":
        n = Lx
        th = np.zeros((n, n))
        bp = np.array([1 if x % 2 == 0 else -1 for x in range(Lx)], dtype=np.int32)
        for x in range(Lx):
            if x + 1 < Lx:
                j = x + 1
            elif pbc and Lx > 1:
                j = 0
            else:
                continue
            th[x, j] += t
            th[j, x] += t
        bip = not (pbc and Lx > 1 and Lx % 2 != 0)
        return th, bp, bip
    if kind == "square":
        n = Lx * Ly
        th = np.zeros((n, n))
        bp = np.zeros(n, dtype=np.int32)
        for y in range(Ly):
            for x in range(Lx):
                i = x + y * Lx
                for (xx, yy, ok) in ((x + 1, y, True), (x, y + 1, True)):
                    if xx >= Lx:
                        if not (pbc and Lx > 1):
                            continue
                        xx = 0
                    if yy >= Ly:
                        if not (pbc and Ly > 1):
                            continue
                        yy = 0
                    j = xx + yy * Lx
                    th[i, j] += t
                    th[j, i] += t
                bp[i] = 1 if (x + y) % 2 == 0 else -1
        odd = (pbc and Lx > 1 and Lx % 2 != 0) or (pbc and Ly > 1 and Ly % 2 != 0)
        return th, bp, not odd
    raise ValueError(f"unknown lattice {kind}")


def model(thop, U: float, dtau: float):
    """expK = exp(-dtau K), expKinv = exp(+dtau K), K = t - mu I, mu = U/2
    (model_init, src/model.c:8-31, via a symmetric eigendecomposition as
    la_expm_sym does, src/linalg.c:319-346) and lambda = acosh(exp(dtau U/2))
    (src/field.c:9)."""
    n = thop.shape[0]
    mu = U / 2.0
    K = thop - mu * np.eye(n)
    w, V = np.linalg.eigh(K)
    expK = (V * np.exp(-dtau * w)) @ V.T
    expKinv = (V * np.exp(dtau * w)) @ V.T
    lam = math.acosh(math.exp(dtau * U / 2.0))
    return expK, expKinv, mu, lam


def phase_cos(Lx: int, Ly: int, q="all"):
    """phase_cos[disp + q*n] for the momenta (structure_factor_plan_init,
    src/structure_factor.c:107-206).  q = 'all', 'af' or a list of (mx, my)."""
    n = Lx * Ly
    if q == "all":
        mq = [(mx, my) for my in range(Ly) for mx in range(Lx)]
    elif q == "af":
        mq = [(Lx // 2 if Lx > 1 else 0, Ly // 2 if Ly > 1 else 0)]
    else:
        mq = list(q)
    disp = np.arange(n)
    dx, dy = disp % Lx, disp // Lx
    ph = np.zeros((len(mq), n))
    for k, (mx, my) in enumerate(mq):
        ph[k] = np.cos(2.0 * math.pi * (mx * dx / Lx + my * dy / Ly))
    return np.ascontiguousarray(ph.reshape(-1)), mq


def replica_seed(base: int, beta_index: int, replica_id: int) -> int:
    """Per-replica seed (replica_seed, src/replica.c:17-34)."""
    M = (1 << 64) - 1
    legacy = (base + 1000 * beta_index) & M
    if replica_id == 0:
        return legacy

    def sm(x):
        z = (x + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    x = base & M
    x ^= (0xD1B54A32D192ED03 * ((beta_index + 1) & 0xFFFFFFFF)) & M
    x ^= (0xABC98388FB8FAC03 * ((replica_id + 1) & 0xFFFFFFFF)) & M
    m = sm(x)
    if m == legacy:
        m = sm(m)
    return m


@dataclass
class ModelInfo:
    n: int
    L: int
    stab: int
    Lx: int
    Ly: int
    U: float
    dtau: float
    mu: float
    thop: np.ndarray
    bipart: np.ndarray
    nq: int = 0
    nbin: int = 0
    first: int = 0
    nrep: int = 0


class Engine:
    """One engine context on one device (dqmc_gpu_create ... destroy)."""

    def __init__(self, precision: int = FP64, device: int = 0, use_graph: bool = True, drift=False,
                 drift_guard: float = 0.0, autotune: bool = False):
        """drift: False / "off", True / "scratch" (the reference's from-scratch
        G, dqmc_record_stab_drift) or "rebuild" (the rebuilt G itself);
        drift_guard > 0 fails a replica whose G drifts further (required by
        the mixed modes)."""
        L = lib()
        o = _Opts(precision, 1 if use_graph else 0, DRIFT[drift], 1 if autotune else 0, float(drift_guard))
        h = C.c_void_p()
        _check(L.dqmc_gpu_create(C.byref(h), device, C.byref(o)), "dqmc_gpu_create")
        self.h = h
        self.models: list[ModelInfo] = []
        self.nrep = 0

    def close(self):
        if self.h:
            lib().dqmc_gpu_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_model(self, kind, Lx, Ly, U, dtau, L, stab, delay_k=0, t=-1.0, pbc=1):
        th, bp, bip = lattice(kind, Lx, Ly, t, pbc)
        if not bip:
            raise ValueError("the engine implements the particle-hole symmetric path: lattice must be bipartite")
        eK, eKi, mu, lam = model(th, U, dtau)
        n = th.shape[0]
        mid = C.c_int()
        _check(lib().dqmc_gpu_add_model(self.h, n, L, stab, delay_k, _ptr(np.asfortranarray(eK)),
                                        _ptr(np.asfortranarray(eKi)), _ptr(np.asfortranarray(th)), _ptr(bp),
                                        U, dtau, mu, lam, Lx, Ly if kind == "square" else 1, C.byref(mid)),
               "dqmc_gpu_add_model")
        self.models.append(ModelInfo(n, L, stab, Lx, Ly if kind == "square" else 1, U, dtau, mu, th, bp))
        return mid.value

    def add_replicas(self, model_id, seeds, fields=None):
        seeds = np.asarray(seeds, dtype=np.uint64)
        f = None if fields is None else np.ascontiguousarray(fields, dtype=np.int8)
        first = C.c_int()
        _check(lib().dqmc_gpu_add_replicas(self.h, model_id, len(seeds), _ptr(seeds), _ptr(f), C.byref(first)),
               "dqmc_gpu_add_replicas")
        m = self.models[model_id]
        if m.nrep == 0:
            m.first = first.value
        m.nrep += len(seeds)
        self.nrep += len(seeds)
        return first.value

    def set_measure(self, model_id, nbin, q="all"):
        m = self.models[model_id]
        ph, mq = phase_cos(m.Lx, m.Ly, q) if q is not None else (np.zeros(1), [])
        nq = len(mq)
        _check(lib().dqmc_gpu_set_measure(self.h, model_id, nq, _ptr(ph), nbin), "dqmc_gpu_set_measure")
        m.nq, m.nbin = nq, nbin
        return mq

    def set_sweep_order(self, model_id, order="forward"):
        """sweep_order = "forward" or "alternating" (src/io.c:284-286)."""
        code = {"forward": SWEEP_FORWARD, "alternating": SWEEP_ALTERNATING}[order]
        _check(lib().dqmc_gpu_set_sweep_order(self.h, model_id, code), "dqmc_gpu_set_sweep_order")

    def set_rebuild(self, model_id, mode="combine"):
        """green_rebuild = "combine", "two_sided" or "centered" (src/io.c:287-289)."""
        _check(lib().dqmc_gpu_set_rebuild(self.h, model_id, REBUILD[mode]), "dqmc_gpu_set_rebuild")

    def init(self):
        _check(lib().dqmc_gpu_init(self.h), "dqmc_gpu_init")

    def sweep(self, nsweep=1, bin=-1):
        _check(lib().dqmc_gpu_sweep(self.h, nsweep, bin), "dqmc_gpu_sweep")

    def run(self, nwarm, per, nbin):
        _check(lib().dqmc_gpu_run(self.h, nwarm, per, nbin), "dqmc_gpu_run")

    def step(self, step, arg):
        _check(lib().dqmc_gpu_step(self.h, step, arg), "dqmc_gpu_step")

    def elapsed_ms(self):
        v = C.c_double()
        _check(lib().dqmc_gpu_last_elapsed_ms(self.h, C.byref(v)), "dqmc_gpu_last_elapsed_ms")
        return v.value

    PROF_NAMES = ["metro", "wrap", "stack", "lmul", "combine", "inv_one_plus", "accept", "measure", "qr",
                  "trsm", "gemm", "flush", "propose", "persist"]

    def profile(self, nsweep=1):
        """Per-stage milliseconds and stage counts of nsweep sweeps
        (dqmc_gpu_profile_sweep)."""
        ms = np.zeros(len(self.PROF_NAMES))
        cnt = np.zeros(len(self.PROF_NAMES), dtype=np.int32)
        _check(lib().dqmc_gpu_profile_sweep(self.h, nsweep, _ptr(ms), _ptr(cnt)), "dqmc_gpu_profile_sweep")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(self.PROF_NAMES)}

    def _model_of(self, r):
        for m in self.models:
            if m.first <= r < m.first + m.nrep:
                return m
        raise IndexError(r)

    def tuning(self, model_id):
        """(window, segments, separable) of a model (dqmc_gpu_read_tuning)."""
        w, sg, sp = C.c_int(), C.c_int(), C.c_int()
        _check(lib().dqmc_gpu_read_tuning(self.h, model_id, C.byref(w), C.byref(sg), C.byref(sp)),
               "dqmc_gpu_read_tuning")
        return w.value, sg.value, sp.value

    def ratios(self, r):
        """R_up R_dn of each site at the latest Metropolis slice of replica r
        (dqmc_gpu_read_ratios; the model needs DQMC_GPU_RATIO_LOG=1)."""
        m = self._model_of(r)
        x = np.zeros(m.n)
        _check(lib().dqmc_gpu_read_ratios(self.h, r, _ptr(x)), "dqmc_gpu_read_ratios")
        return x

    def green(self, r):
        m = self._model_of(r)
        g = np.zeros(m.n * m.n)
        _check(lib().dqmc_gpu_read_green(self.h, r, _ptr(g)), "dqmc_gpu_read_green")
        return g.reshape(m.n, m.n, order="F")

    def set_green(self, r, g):
        g = np.asfortranarray(g, dtype=np.float64)
        _check(lib().dqmc_gpu_write_green(self.h, r, _ptr(g)), "dqmc_gpu_write_green")

    def field(self, r):
        m = self._model_of(r)
        s = np.zeros(m.L * m.n, dtype=np.int8)
        _check(lib().dqmc_gpu_read_field(self.h, r, _ptr(s)), "dqmc_gpu_read_field")
        return s

    def set_field(self, r, s):
        s = np.ascontiguousarray(s, dtype=np.int8)
        _check(lib().dqmc_gpu_write_field(self.h, r, _ptr(s)), "dqmc_gpu_write_field")

    def rng(self, r):
        s = np.zeros(4, dtype=np.uint64)
        _check(lib().dqmc_gpu_read_rng(self.h, r, _ptr(s)), "dqmc_gpu_read_rng")
        return s

    def status(self):
        n = self.nrep
        st = np.zeros(n, dtype=np.int32)
        rs = np.zeros(n, dtype=np.int32)
        acc = np.zeros(n, dtype=np.uint64)
        att = np.zeros(n, dtype=np.uint64)
        sg = np.zeros(n)
        _check(lib().dqmc_gpu_read_status(self.h, _ptr(st), _ptr(rs), _ptr(acc), _ptr(att), _ptr(sg)),
               "dqmc_gpu_read_status")
        return dict(status=st, reason=rs, accepted=acc, attempts=att, sign=sg)

    def drift(self):
        n = self.nrep
        mx, mn = np.zeros(n), np.zeros(n)
        cnt = np.zeros(n, dtype=np.uint64)
        _check(lib().dqmc_gpu_read_drift(self.h, _ptr(mx), _ptr(mn), _ptr(cnt)), "dqmc_gpu_read_drift")
        return mx, mn, cnt

    def drift_at(self):
        """(max_tau, max_sweep) of each replica's largest drift sample."""
        tau = np.zeros(self.nrep, dtype=np.int32)
        sw = np.zeros(self.nrep, dtype=np.uint64)
        _check(lib().dqmc_gpu_read_drift_at(self.h, _ptr(tau), _ptr(sw)), "dqmc_gpu_read_drift_at")
        return tau, sw

    def bins(self, r):
        m = self._model_of(r)
        nb = max(m.nbin, 1)
        b8 = np.zeros((nb, 8))
        cnt = np.zeros(nb, dtype=np.int32)
        nq = max(m.nq, 1)
        z = np.zeros((nb, nq))
        p = np.zeros((nb, nq))
        _check(lib().dqmc_gpu_read_bins(self.h, r, _ptr(b8), _ptr(cnt), _ptr(z), _ptr(p)), "dqmc_gpu_read_bins")
        return b8, cnt, z[:, : m.nq], p[:, : m.nq]


# ------------------------------------------------ batched linear algebra

def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def udv_lmul(U, D, T, B, device=0):
    """Batched udv_lmul (src/linalg.c:515-642). Arrays [batch, n, n] holding
    column-major matrices as numpy transposes (a[b].T is the matrix)."""
    U, D, T, B = _f(U).copy(), _f(D).copy(), _f(T).copy(), _f(B)
    b, n = D.shape
    _check(lib().dqmc_gpu_udv_lmul(device, n, b, _ptr(U), _ptr(D), _ptr(T), _ptr(B)), "dqmc_gpu_udv_lmul")
    return U, D, T


def udv_rmul(U, D, T, B, device=0):
    U, D, T, B = _f(U).copy(), _f(D).copy(), _f(T).copy(), _f(B)
    b, n = D.shape
    _check(lib().dqmc_gpu_udv_rmul(device, n, b, _ptr(U), _ptr(D), _ptr(T), _ptr(B)), "dqmc_gpu_udv_rmul")
    return U, D, T


def udv_combine(Ul, Dl, Tl, Ur, Dr, Tr, device=0):
    Ul, Dl, Tl, Ur, Dr, Tr = map(_f, (Ul, Dl, Tl, Ur, Dr, Tr))
    b, n = Dl.shape
    Uo, Do, To = np.zeros_like(Ul), np.zeros_like(Dl), np.zeros_like(Tl)
    _check(lib().dqmc_gpu_udv_combine(device, n, b, _ptr(Ul), _ptr(Dl), _ptr(Tl), _ptr(Ur), _ptr(Dr), _ptr(Tr),
                                      _ptr(Uo), _ptr(Do), _ptr(To)), "dqmc_gpu_udv_combine")
    return Uo, Do, To


def udv_inv_one_plus(U, D, T, device=0):
    U, D, T = map(_f, (U, D, T))
    b, n = D.shape
    g = np.zeros_like(U)
    sg = np.zeros(b, dtype=np.int32)
    _check(lib().dqmc_gpu_udv_inv_one_plus(device, n, b, _ptr(U), _ptr(D), _ptr(T), _ptr(g), _ptr(sg)),
           "dqmc_gpu_udv_inv_one_plus")
    return g, sg


def udv_inv_one_plus_off(U, D, T, off, device=0):
    """Batched udv_inv_one_plus_work on UDVs with log offsets.  Returns
    (g, sign, reason, rc): rc != 0 when an entry failed (reason[b] != 0),
    g and sign then unset."""
    U, D, T = map(_f, (U, D, T))
    off = _f(off)
    b, n = D.shape
    g = np.zeros_like(U)
    sg = np.zeros(b, dtype=np.int32)
    rs = np.zeros(b, dtype=np.int32)
    rc = lib().dqmc_gpu_udv_inv_one_plus_off(device, n, b, _ptr(U), _ptr(D), _ptr(T), _ptr(off), _ptr(g), _ptr(sg),
                                             _ptr(rs))
    return g, sg, rs, rc


def measure(Lx, Ly, gu, thop, bipart, U, phase=None, nq=0, device=0):
    gu = _f(gu)
    b = gu.shape[0]
    s5 = np.zeros((b, 5))
    z = np.zeros((b, max(nq, 1)))
    p = np.zeros((b, max(nq, 1)))
    _check(lib().dqmc_gpu_measure(device, Lx, Ly, b, _ptr(gu), _ptr(np.asfortranarray(thop)),
                                  _ptr(np.ascontiguousarray(bipart, dtype=np.int32)), U, nq,
                                  _ptr(phase if phase is not None else np.zeros(1)), _ptr(s5), _ptr(z), _ptr(p)),
           "dqmc_gpu_measure")
    return s5, z[:, :nq], p[:, :nq]
