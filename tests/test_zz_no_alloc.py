"""dqmc_gpu_udv_inv_one_plus_noalloc (the batched (1 + U D T)^-1 on
caller-owned device buffers with a static device workspace) against
dqmc_gpu_udv_inv_one_plus on the same inputs: identical G and sign, on two
calls in a row, the second smaller than the first; a call beyond the
workspace is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def udvs(rng, batch, n):
    U = np.zeros((batch, n, n))
    T = np.zeros((batch, n, n))
    D = np.zeros((batch, n))
    for b in range(batch):
        q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        U[b] = q.T  # a[b].T is the column-major matrix
        t = np.triu(rng.standard_normal((n, n)) * 0.3, 1) + np.eye(n)
        T[b] = t.T
        D[b] = np.sign(rng.standard_normal(n)) * np.exp(np.linspace(25.0, -25.0, n) + rng.standard_normal(n))
    return U, D, T


class Hip:
    """Caller-side device buffers through the HIP runtime the engine itself
    links (libamdhip64): the test, not the entry point, allocates."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0
        self.bufs.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def noalloc(eng_mod, U, D, T):
    b, n = D.shape
    hip = Hip()
    dU, dD, dT = hip.put(U), hip.put(D), hip.put(T)
    g0 = np.full_like(U, np.nan)
    s0 = np.full(b, 7, dtype=np.int32)
    dg, ds, dr = hip.put(g0), hip.put(s0), hip.put(s0)
    assert hip.h.hipDeviceSynchronize() == 0
    rc = eng_mod.lib().dqmc_gpu_udv_inv_one_plus_noalloc(n, b, dU, dD, dT, dg, ds, dr, None)
    out = (rc, hip.get(dg, g0), hip.get(ds, s0), hip.get(dr, s0))
    hip.free()
    return out


def test_noalloc_matches_allocating_entry(eng_mod):
    rng = np.random.default_rng(2024)
    for (b, n) in ((8, 64), (3, 24)):
        print(f"noalloc inverse: batch {b}, n {n}", flush=True)
        U, D, T = udvs(rng, b, n)
        g0, s0 = eng_mod.udv_inv_one_plus(U, D, T)
        rc, g1, s1, r1 = noalloc(eng_mod, U, D, T)
        assert rc == 0, eng_mod.lib().dqmc_gpu_last_error()
        assert (r1 == 0).all()
        assert np.array_equal(s1, s0)
        assert np.array_equal(g1, g0)


def test_noalloc_refuses_oversize(eng_mod):
    n, b = 256, 64  # 64 x ~2.2 MB of workspace > 64 MiB
    hip = Hip()
    x = hip.put(np.zeros(8))
    i = hip.put(np.zeros(8, dtype=np.int32))
    rc = eng_mod.lib().dqmc_gpu_udv_inv_one_plus_noalloc(n, b, x, x, x, x, i, None, None)
    hip.free()
    assert rc != 0
    assert b"workspace" in eng_mod.lib().dqmc_gpu_last_error()
