"""msvg_search_noalloc, the allocation-free batched search: its results
equal msvg_search's (batched mode) bit for bit on two calls in a row, the
second over fewer profiles and a smaller database (the shared scratch keeps
the first call's contents); sizes beyond its scratch and column-segmented
models are refused with EINVAL before anything is launched."""
import ctypes as C

import numpy as np
import pytest

import cudahmmsearch_amd as ca

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return _hip


class DevArray:
    """A caller-owned device array (the caller allocates its outputs)."""

    def __init__(self, nbytes):
        self.n = nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(nbytes, 1))) == 0
        assert hip().hipMemset(self.p, 0xA5, C.c_size_t(max(nbytes, 1))) == 0   # not zero: every byte is written

    def get(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes == self.n
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.n), 2) == 0   # D2H
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(ctx, pids, did, n, F1=0.02):
    ids = np.ascontiguousarray(pids, np.int32)
    np_ = ids.size
    u, s, p = DevArray(np_ * n * 4), DevArray(np_ * n), DevArray(np_ * n)
    try:
        rc = ctx.lib.msvg_search_noalloc(ctx._c, ids.ctypes.data, np_, did, F1, u.p, s.p, p.p)
        if rc != 0:
            return rc, None
        return 0, (u.get(np.float32, (np_, n)), s.get(np.uint8, (np_, n)), p.get(np.uint8, (np_, n)))
    finally:
        u.free(); s.free(); p.free()


def _profiles(ctx, specs):
    out = []
    for seed, M in specs:
        prof = ca.sample_profile(seed, M)
        pid = ctx.add_profile(prof)
        prof.mu = ctx.calibrate(pid, prof.lam)
        out.append((prof, pid))
    return out


def test_noalloc_equals_search_twice():
    ctx = ca.MSVContext(0)
    profs = _profiles(ctx, [(10, 50), (11, 180), (12, 420), (13, 777), (14, 1300), (15, 2500), (16, 3000),
                            (17, 96)])
    db1 = ca.gen_database(41, 30000, 300, 0.6, 20, 12000, tail_frac=0.02, tail_xm=1000, tail_alpha=1.5,
                          tail_max=12000)
    db1, _ = ca.plant(db1, ca.core_emit(profs[2][0].hmm, 3, 300), 5, 0.01)
    db2 = ca.gen_database(42, 5000, 250, 0.7, 10, 4000)
    db2, _ = ca.plant(db2, ca.core_emit(profs[0][0].hmm, 4, 100), 6, 0.02)
    runs = [([pid for _, pid in profs], ctx.add_database(db1), db1.n),
            ([profs[0][1], profs[3][1], profs[5][1]], ctx.add_database(db2), db2.n)]
    for pids, did, n in runs:
        print(f"noalloc: {len(pids)} profiles x {n} targets", flush=True)
        want = ctx.search(pids, did, mode=ca.MODE_BATCHED)
        rc, got = noalloc(ctx, pids, did, n)
        assert rc == 0, ctx.lib.msvg_last_error(ctx._c)
        assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32)), "usc"
        assert np.array_equal(got[1], want[1]), "status"
        assert np.array_equal(got[2], want[2]), "pass"
        assert got[2].sum() > 0
    ctx.close()


def test_noalloc_refuses_beyond_its_scratch():
    ctx = ca.MSVContext(0)
    profs = _profiles(ctx, [(20 + i, 300) for i in range(8)] + [(40, 8000)])
    db = ca.gen_database(43, 6_000_000, 100, 0.3, 20, 400, nthreads=8)
    did = ctx.add_database(db)
    rc, _ = noalloc(ctx, [pid for _, pid in profs[:8]], did, db.n)     # 8 x 6M pairs: beyond 128 MiB
    assert rc == ca.eslEINVAL
    assert b"limit is" in ctx.lib.msvg_last_error(ctx._c)
    small = ctx.add_database(ca.gen_database(44, 100, 200, 0.5, 20, 1000))
    rc, _ = noalloc(ctx, [profs[8][1]], small, 100)                    # M = 8000: column-segmented
    assert rc == ca.eslEINVAL
    assert b"column-segmented" in ctx.lib.msvg_last_error(ctx._c)
    rc, got = noalloc(ctx, [profs[0][1]], small, 100)                   # and a fitting call still works
    assert rc == 0
    want = ctx.search([profs[0][1]], small, mode=ca.MODE_BATCHED)
    assert np.array_equal(got[0].view(np.uint32), want[0].view(np.uint32))
    ctx.close()
