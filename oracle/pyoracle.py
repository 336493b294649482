"""ctypes access to the CPU oracle (liboracle.so) and, when it has been
built, to the reference itself (oracle/_ref/libsaiqmc_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsaiqmc_ref.so")
REF_OMP = os.path.join(HERE, "_ref", "dqmc_omp")

_P, _I, _D, _U64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


_orc = None
REBUILD = {"combine": 0, "two_sided": 1, "centered": 2}


def orc():
    global _orc
    if _orc is None:
        L = C.CDLL(ORACLE_SO)
        sig = {
            "orc_rng_seed": (None, [_P, _U64]),
            "orc_rng_double": (_D, [_P]),
            "orc_replica_seed": (_U64, [_U64, _I, _I]),
            "orc_lattice": (_I, [_I, _I, _I, _I, _D, _P, _P]),
            "orc_expm_sym": (None, [_I, _P, _D, _P]),
            "orc_model": (None, [_I, _P, _D, _D, _P, _P, _P]),
            "orc_matmul": (None, [_I, _P, _P, _P]),
            "orc_geqrf": (None, [_I, _P, _P]),
            "orc_orgqr": (None, [_I, _P, _P]),
            "orc_getrf": (_I, [_I, _P, _P]),
            "orc_inverse": (_I, [_I, _P, _P]),
            "orc_trtri_unit_upper": (None, [_I, _P]),
            "orc_udv_lmul": (_I, [_I, _P, _P, _P, _P]),
            "orc_udv_rmul": (_I, [_I, _P, _P, _P, _P]),
            "orc_udv_combine": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
            "orc_udv_inv_one_plus": (_I, [_I, _P, _P, _P, _P, _P]),
            "orc_expv": (None, [_D, _D, _D, _D, _P]),
            "orc_build_B": (None, [_I, _P, _P, _P, _P]),
            "orc_build_Binv": (None, [_I, _P, _P, _P, _P]),
            "orc_build_Bblock": (None, [_I, _P, _P, _P, _I, _I, _P]),
            "orc_wrap": (None, [_I, _P, _P, _P, _P, _P]),
            "orc_wrap_backward": (None, [_I, _P, _P, _P, _P, _P]),
            "orc_ph_down": (None, [_I, _P, _P, _P]),
            "orc_measure_sample": (None, [_I, _P, _D, _P, _P, _P]),
            "orc_phase_cos": (None, [_I, _I, _I, _P, _P, _P]),
            "orc_szz_sperp": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
            "orc_jackknife": (None, [_P, _I, _P, _P]),
            "orc_replica_new": (_P, [_I, _I, _I, _P, _P, _D, _D, _U64, _I, _I]),
            "orc_replica_status": (_I, [_P]),
            "orc_replica_sweep": (_I, [_P]),
            "orc_replica_set_alternating": (None, [_P, _I]),
            "orc_replica_set_rebuild": (_I, [_P, _I]),
            "orc_replica_reason": (_I, [_P]),
            "orc_udv_inv_one_plus_off": (_I, [_I, _P, _P, _P, _D, _P, _P]),
            "orc_udv_inv_two_sided": (_I, [_I, _P, _P, _P, _D, _P, _P, _P, _D, _P, _P]),
            "orc_udv_recenter": (_I, [_I, _P, _P]),
            "orc_udv_lmul_centered": (_I, [_I, _P, _P, _P, _P, _P]),
            "orc_udv_rmul_centered": (_I, [_I, _P, _P, _P, _P, _P]),
            "orc_replica_green": (None, [_P, _P, _P]),
            "orc_replica_field": (None, [_P, _P]),
            "orc_replica_set_field": (None, [_P, _P]),
            "orc_replica_stats": (None, [_P, _P, _P, _P, _P]),
            "orc_replica_rng_state": (None, [_P, _P]),
            "orc_replica_model": (None, [_P, _P, _P, _P]),
            "orc_replica_drift": (None, [_P, _P, _P, _P]),
            "orc_replica_rebuild": (_I, [_P]),
            "orc_replica_log_ratios": (_I, [_P]),
            "orc_replica_ratios": (None, [_P, _P]),
            "orc_replica_free": (None, [_P]),
            "orc_green_exact": (_I, [_I, _I, _I, _I, _P, _P, _P, _P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _orc = L
    return _orc


def lattice(kind, Lx, Ly=1, t=-1.0, pbc=1):
    n = Lx * (Ly if kind == "square" else 1)
    th = np.zeros(n * n)
    bp = np.zeros(n, dtype=np.int32)
    bip = orc().orc_lattice(0 if kind == "chain" else 1, Lx, Ly, pbc, t, _ptr(th), _ptr(bp))
    return th.reshape(n, n, order="F"), bp, bool(bip)


def model(thop, U, dtau):
    n = thop.shape[0]
    eK = np.zeros(n * n)
    eKi = np.zeros(n * n)
    lam = C.c_double()
    orc().orc_model(n, _ptr(np.asfortranarray(thop)), U, dtau, _ptr(eK), _ptr(eKi), C.byref(lam))
    return eK.reshape(n, n, order="F"), eKi.reshape(n, n, order="F"), lam.value


class Replica:
    """The oracle's PH sweep engine for one replica (forward, or alternating
    forward/backward sweeps with carried stacks)."""

    def __init__(self, kind, Lx, Ly, U, dtau, L, stab, seed, delay_cap=16, drift=False, t=-1.0,
                 alternating=False, rebuild="combine", ratio_log=False):
        th, bp, _ = lattice(kind, Lx, Ly, t)
        self.n = th.shape[0]
        self.L = L
        self.thop, self.bipart = th, bp
        self.U = U
        self.h = C.c_void_p(orc().orc_replica_new(self.n, L, stab, _ptr(np.asfortranarray(th)), _ptr(bp), U, dtau,
                                                  seed, delay_cap, 1 if drift else 0))
        if rebuild != "combine":
            orc().orc_replica_set_rebuild(self.h, REBUILD[rebuild])
        if alternating:
            orc().orc_replica_set_alternating(self.h, 1)
        if ratio_log and orc().orc_replica_log_ratios(self.h):
            raise MemoryError("oracle ratio log")

    def __del__(self):
        if getattr(self, "h", None):
            orc().orc_replica_free(self.h)
            self.h = None

    def sweep(self):
        return orc().orc_replica_sweep(self.h)

    def green(self):
        gu = np.zeros(self.n * self.n)
        gd = np.zeros(self.n * self.n)
        orc().orc_replica_green(self.h, _ptr(gu), _ptr(gd))
        return gu.reshape(self.n, self.n, order="F"), gd.reshape(self.n, self.n, order="F")

    def field(self):
        s = np.zeros(self.L * self.n, dtype=np.int8)
        orc().orc_replica_field(self.h, _ptr(s))
        return s

    def set_field(self, s):
        orc().orc_replica_set_field(self.h, _ptr(np.ascontiguousarray(s, dtype=np.int8)))

    def rebuild(self):
        return orc().orc_replica_rebuild(self.h)

    def stats(self):
        a, t, s, d = C.c_ulonglong(), C.c_ulonglong(), C.c_double(), C.c_int()
        orc().orc_replica_stats(self.h, C.byref(a), C.byref(t), C.byref(s), C.byref(d))
        return a.value, t.value, s.value, d.value

    def rng_state(self):
        s = np.zeros(4, dtype=np.uint64)
        orc().orc_replica_rng_state(self.h, _ptr(s))
        return s

    def model(self):
        eK = np.zeros(self.n * self.n)
        eKi = np.zeros(self.n * self.n)
        lam = C.c_double()
        orc().orc_replica_model(self.h, _ptr(eK), _ptr(eKi), C.byref(lam))
        return eK.reshape(self.n, self.n, order="F"), eKi.reshape(self.n, self.n, order="F"), lam.value

    def ratios(self):
        """R_up R_dn of the latest proposal at each (slice, site), L x n
        (needs ratio_log=True)."""
        r = np.zeros(self.L * self.n)
        orc().orc_replica_ratios(self.h, _ptr(r))
        return r.reshape(self.L, self.n)

    def drift(self):
        mx, mn, c = C.c_double(), C.c_double(), C.c_ulonglong()
        orc().orc_replica_drift(self.h, C.byref(mx), C.byref(mn), C.byref(c))
        return mx.value, mn.value, c.value


def green_exact(expK, lam, U, dtau, field, L, tau=0, blk=8):
    """G(tau) = (1 + B_{tau-1}..B_0 B_{L-1}..B_tau)^-1 of a field (L*n values,
    slice-major) in 80-bit extended precision (oracle/exact.c): the true G
    that the FP64 paths approximate, for the given FP64 expK and expv."""
    n = expK.shape[0]
    ev = np.zeros(2)
    orc().orc_expv(lam, U, dtau, 1.0, _ptr(ev))
    g = np.zeros(n * n)
    f = np.ascontiguousarray(field, dtype=np.int8)
    rc = orc().orc_green_exact(n, L, tau, blk, _ptr(np.asfortranarray(expK)), _ptr(ev), _ptr(f), _ptr(g))
    if rc:
        raise RuntimeError("orc_green_exact failed")
    return g.reshape(n, n, order="F")


def measure(gu, thop, bipart, U, Lx, Ly, q="all"):
    """(E, ekin, eint, ntot, doublon), Szz(q), Sperp(q) of the oracle."""
    n = gu.shape[0]
    gd = np.zeros(n * n)
    orc().orc_ph_down(n, _ptr(bipart), _ptr(np.asfortranarray(gu)), _ptr(gd))
    out = np.zeros(5)
    orc().orc_measure_sample(n, _ptr(np.asfortranarray(thop)), U, _ptr(np.asfortranarray(gu)), _ptr(gd), _ptr(out))
    if q == "all":
        mq = [(mx, my) for my in range(Ly) for mx in range(Lx)]
    else:
        mq = list(q)
    mx = np.array([a for a, _ in mq], dtype=np.int32)
    my = np.array([b for _, b in mq], dtype=np.int32)
    ph = np.zeros(n * len(mq))
    orc().orc_phase_cos(Lx, Ly, len(mq), _ptr(mx), _ptr(my), _ptr(ph))
    z = np.zeros(len(mq))
    p = np.zeros(len(mq))
    orc().orc_szz_sperp(Lx, Ly, len(mq), _ptr(ph), _ptr(np.asfortranarray(gu)), _ptr(gd), _ptr(z), _ptr(p))
    return out, z, p


def jackknife(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    m, e = C.c_double(), C.c_double()
    orc().orc_jackknife(_ptr(x), len(x), C.byref(m), C.byref(e))
    return m.value, e.value


# ------------------------------------------------------------- reference

_ref = None


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        sig = {
            "ref_replica_new": (_P, [_I, _I, _I, _I, _D, _D, _D, _I, _I, C.c_ulonglong, _I]),
            "ref_replica_new_mode": (_P, [_I, _I, _I, _I, _D, _D, _D, _I, _I, C.c_ulonglong, _I, _I, _I]),
            "ref_replica_init_rc": (_I, [_P]),
            "ref_replica_sweep": (_I, [_P]),
            "ref_replica_green": (None, [_P, _P, _P]),
            "ref_replica_field": (None, [_P, _P]),
            "ref_replica_set_field": (None, [_P, _P]),
            "ref_replica_stats": (None, [_P, _P, _P, _P, _P, _P]),
            "ref_replica_drift": (None, [_P, _P, _P, _P]),
            "ref_replica_model": (None, [_P, _P, _P, _P, _P, _P]),
            "ref_replica_rng_state": (None, [_P, _P]),
            "ref_replica_measure": (None, [_P, _P]),
            "ref_replica_szz_sperp": (_I, [_P, _P, _P]),
            "ref_replica_free": (None, [_P]),
            "ref_udv_lmul": (_I, [_I, _P, _P, _P, _P]),
            "ref_udv_rmul": (_I, [_I, _P, _P, _P, _P]),
            "ref_udv_combine": (_I, [_I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
            "ref_udv_inv_one_plus": (_I, [_I, _P, _P, _P, _P, _P]),
            "ref_replica_seed": (C.c_ulonglong, [C.c_ulonglong, _I, _I]),
            "ref_expm_sym": (None, [_I, _P, _D, _P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _ref = L
    return _ref


class RefReplica:
    """The compiled reference engine for one replica (oracle/ref_trace.c)."""

    def __init__(self, kind, Lx, Ly, U, dtau, L, stab, seed, drift=False, t=-1.0, alternating=False,
                 rebuild="combine"):
        self.n = Lx * (Ly if kind == "square" else 1)
        self.L = L
        self.h = C.c_void_p(ref().ref_replica_new_mode(0 if kind == "chain" else 1, Lx, Ly, 1, t, U, dtau, L, stab,
                                                       seed, 1 if drift else 0, 1 if alternating else 0,
                                                       REBUILD[rebuild]))

    def __del__(self):
        if getattr(self, "h", None):
            ref().ref_replica_free(self.h)
            self.h = None

    def sweep(self):
        return ref().ref_replica_sweep(self.h)

    def green(self):
        gu = np.zeros(self.n * self.n)
        gd = np.zeros(self.n * self.n)
        ref().ref_replica_green(self.h, _ptr(gu), _ptr(gd))
        return gu.reshape(self.n, self.n, order="F"), gd.reshape(self.n, self.n, order="F")

    def field(self):
        s = np.zeros(self.L * self.n, dtype=np.int8)
        ref().ref_replica_field(self.h, _ptr(s))
        return s

    def stats(self):
        a, t, s, st, d = C.c_ulonglong(), C.c_ulonglong(), C.c_double(), C.c_int(), C.c_int()
        ref().ref_replica_stats(self.h, C.byref(a), C.byref(t), C.byref(s), C.byref(st), C.byref(d))
        return a.value, t.value, s.value, d.value

    def model(self):
        eK = np.zeros(self.n * self.n)
        eKi = np.zeros(self.n * self.n)
        th = np.zeros(self.n * self.n)
        bp = np.zeros(self.n, dtype=np.int32)
        lam = C.c_double()
        ref().ref_replica_model(self.h, _ptr(eK), _ptr(eKi), _ptr(th), _ptr(bp), C.byref(lam))
        f = lambda a: a.reshape(self.n, self.n, order="F")
        return f(eK), f(eKi), f(th), bp, lam.value

    def rng_state(self):
        s = np.zeros(4, dtype=np.uint64)
        ref().ref_replica_rng_state(self.h, _ptr(s))
        return s

    def measure(self):
        out = np.zeros(5)
        ref().ref_replica_measure(self.h, _ptr(out))
        z = np.zeros(self.n)
        p = np.zeros(self.n)
        rc = ref().ref_replica_szz_sperp(self.h, _ptr(z), _ptr(p))
        return out, z, p, rc
