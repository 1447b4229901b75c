"""ctypes loaders for the checker libraries.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package.

* :class:`Oracle` wraps ``oracle/liboracle.so`` -- the scalar C restatements of
  ``p7_MSVFilter`` (impl_sse/msvfilter.c:78-213 with impl_sse/ssvfilter.c:1364-1413)
  and ``p7_ViterbiFilter`` (impl_sse/vitfilter.c:84-250).
* :class:`Reference` wraps ``oracle/_ref/libmsvref.so`` -- the HMMER 3.1b1
  reference compiled from its own sources plus ``oracle/ref_harness.c``.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libmsvref.so")

eslOK, eslEINVAL, eslERANGE, eslENORESULT = 0, 11, 16, 19


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class _OrcProfile(C.Structure):
    _fields_ = [("M", C.c_int), ("Kp", C.c_int), ("rbv", C.c_void_p),
                ("tbm_b", C.c_int), ("tec_b", C.c_int), ("base_b", C.c_int),
                ("bias_b", C.c_int), ("scale_b", C.c_float)]


class _OrcVProfile(C.Structure):
    _fields_ = [("M", C.c_int), ("Kp", C.c_int), ("rwv", C.c_void_p), ("twv", C.c_void_p),
                ("e_loop", C.c_int), ("e_move", C.c_int), ("base_w", C.c_int), ("ddbound_w", C.c_int),
                ("scale_w", C.c_float), ("nj", C.c_float)]


def vprofile_from_fields(f):
    """Viterbi profile dict in model order (k = 1..M) from ref_vit_fields'
    padded arrays: {rwv (Kp x M int16), twv (8 x M int16), e_loop, e_move,
    base_w, ddbound_w, scale_w, nj, vmu, vlambda}."""
    M = f["M"]
    return {"M": M, "rwv": np.ascontiguousarray(f["rwv_pad"][:, :M]),
            "twv": np.ascontiguousarray(f["twv_pad"][:, :M]), "e_loop": f["e_loop"], "e_move": f["e_move"],
            "base_w": f["base_w"], "ddbound_w": f["ddbound_w"], "scale_w": f["scale_w"], "nj": f["nj"],
            "vmu": f["vmu"], "vlambda": f["vlambda"]}


class Oracle:
    """The CPU restatement.  Profiles are dicts with keys rbv (Kp x M uint8),
    tbm, tec, base, bias, scale (see cudahmmsearch_amd.ByteProfile)."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run make -C oracle")
        lib = C.CDLL(path)
        lib.orc_tjb.restype = C.c_int
        lib.orc_tjb.argtypes = [C.c_int]
        for f in ("orc_ssv", "orc_msv_bytes", "orc_msv"):
            getattr(lib, f).restype = C.c_int
            getattr(lib, f).argtypes = [C.POINTER(_OrcProfile), C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.orc_score.restype = C.c_float
        lib.orc_score.argtypes = [C.c_int, C.c_int, C.c_int, C.c_float]
        lib.orc_nullsc.restype = C.c_float
        lib.orc_nullsc.argtypes = [C.c_int]
        lib.orc_pass.restype = C.c_int
        lib.orc_pass.argtypes = [C.c_float, C.c_int, C.c_float, C.c_float, C.c_double]
        lib.orc_jfree_max.restype = C.c_int
        lib.orc_jfree_max.argtypes = [C.POINTER(_OrcProfile), C.c_void_p, C.c_int]
        lib.orc_decide.restype = C.c_int
        lib.orc_decide.argtypes = [C.POINTER(_OrcProfile), C.c_int, C.c_int, C.POINTER(C.c_int)]
        lib.orc_msv_batch.restype = C.c_int
        lib.orc_msv_batch.argtypes = [C.POINTER(_OrcProfile), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        lib.orc_vit_move.restype = C.c_int
        lib.orc_vit_move.argtypes = [C.POINTER(_OrcVProfile), C.c_int]
        lib.orc_vit.restype = C.c_int
        lib.orc_vit.argtypes = [C.POINTER(_OrcVProfile), C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.orc_vit_batch.restype = C.c_int
        lib.orc_vit_batch.argtypes = [C.POINTER(_OrcVProfile), C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        self.lib = lib

    @staticmethod
    def _prof(p):
        rbv = np.ascontiguousarray(p["rbv"], dtype=np.uint8)
        s = _OrcProfile(int(rbv.shape[1]), int(rbv.shape[0]), rbv.ctypes.data,
                        int(p["tbm"]), int(p["tec"]), int(p["base"]), int(p["bias"]), float(p["scale"]))
        return s, rbv

    def tjb(self, L):
        return self.lib.orc_tjb(int(L))

    def _one(self, fn, p, x):
        s, keep = self._prof(p)
        x = np.ascontiguousarray(x, dtype=np.uint8)
        usc = C.c_float()
        st = fn(C.byref(s), _p(x), int(x.size), C.byref(usc))
        return st, usc.value

    def msv(self, p, x):
        return self._one(self.lib.orc_msv, p, x)

    def ssv(self, p, x):
        return self._one(self.lib.orc_ssv, p, x)

    def msv_bytes(self, p, x):
        return self._one(self.lib.orc_msv_bytes, p, x)

    def msv_batch(self, p, res, off, lens, nthreads=8):
        s, keep = self._prof(p)
        res = np.ascontiguousarray(res, dtype=np.uint8)
        off = np.ascontiguousarray(off, dtype=np.int64)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        n = lens.size
        usc = np.zeros(n, np.float32)
        st = np.zeros(n, np.int32)
        self.lib.orc_msv_batch(C.byref(s), _p(res), _p(off), _p(lens), n, _p(usc), _p(st), nthreads)
        return st, usc

    def jfree_max(self, p, x):
        s, keep = self._prof(p)
        x = np.ascontiguousarray(x, dtype=np.uint8)
        return self.lib.orc_jfree_max(C.byref(s), _p(x), int(x.size))

    def decide(self, p, Rp, L):
        s, keep = self._prof(p)
        xJ = C.c_int(-1)
        code = self.lib.orc_decide(C.byref(s), int(Rp), int(L), C.byref(xJ))
        return code, xJ.value

    def score(self, xJ, tjb, base, scale):
        return self.lib.orc_score(int(xJ), int(tjb), int(base), float(scale))

    def nullsc(self, L):
        return self.lib.orc_nullsc(int(L))

    def passes(self, usc, L, mu, lam, F1=0.02):
        return self.lib.orc_pass(float(usc), int(L), float(mu), float(lam), float(F1))

    # ---- p7_ViterbiFilter ----
    @staticmethod
    def _vprof(v):
        rwv = np.ascontiguousarray(v["rwv"], dtype=np.int16)
        twv = np.ascontiguousarray(v["twv"], dtype=np.int16)
        s = _OrcVProfile(int(rwv.shape[1]), int(rwv.shape[0]), rwv.ctypes.data, twv.ctypes.data,
                         int(v["e_loop"]), int(v["e_move"]), int(v["base_w"]), int(v["ddbound_w"]),
                         float(v["scale_w"]), float(v["nj"]))
        return s, (rwv, twv)

    def vit_move(self, v, L):
        s, keep = self._vprof(v)
        return self.lib.orc_vit_move(C.byref(s), int(L))

    def vit(self, v, x):
        s, keep = self._vprof(v)
        x = np.ascontiguousarray(x, dtype=np.uint8)
        sc = C.c_float()
        st = self.lib.orc_vit(C.byref(s), _p(x), int(x.size), C.byref(sc))
        return st, sc.value

    def vit_batch(self, v, res, off, lens, nthreads=8):
        s, keep = self._vprof(v)
        res = np.ascontiguousarray(res, dtype=np.uint8)
        off = np.ascontiguousarray(off, dtype=np.int64)
        lens = np.ascontiguousarray(lens, dtype=np.int32)
        n = lens.size
        sc = np.zeros(n, np.float32)
        st = np.zeros(n, np.int32)
        self.lib.orc_vit_batch(C.byref(s), _p(res), _p(off), _p(lens), n, _p(sc), _p(st), nthreads)
        return st, sc


class Reference:
    """The compiled HMMER reference (only where oracle/_ref was built)."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is not built; run make -C oracle ref")
        lib = C.CDLL(path)
        lib.ref_profile_sample.restype = C.c_void_p
        lib.ref_profile_sample.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_int]
        lib.ref_profile_from_probs.restype = C.c_void_p
        lib.ref_profile_from_probs.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        lib.ref_profile_free.argtypes = [C.c_void_p]
        lib.ref_profile_M.argtypes = [C.c_void_p]
        lib.ref_profile_rbv.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_profile_scalars.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_profile_hmm.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_profile_msc.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_tjb.argtypes = [C.c_void_p, C.c_int]
        for f in ("ref_msv", "ref_ssv"):
            getattr(lib, f).argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.ref_nullsc.restype = C.c_float
        lib.ref_nullsc.argtypes = [C.c_void_p, C.c_int]
        lib.ref_pass.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_double]
        lib.ref_msv_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int]
        lib.ref_iid.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_void_p]
        lib.ref_uniform.argtypes = [C.c_uint32, C.c_int, C.c_int, C.c_void_p]
        lib.ref_profile_from_file.restype = C.c_void_p
        lib.ref_profile_from_file.argtypes = [C.c_char_p, C.c_int]
        lib.ref_profile_Kp.argtypes = [C.c_void_p]
        lib.ref_digitize.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
        lib.ref_core_emit.restype = C.c_int64
        lib.ref_core_emit.argtypes = [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_int64, C.c_void_p]
        lib.ref_vit.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.ref_vit_Q.argtypes = [C.c_void_p]
        lib.ref_vit_fields.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.ref_vit_move.argtypes = [C.c_void_p, C.c_int]
        lib.ref_profile_tsc.argtypes = [C.c_void_p, C.c_void_p]
        lib.ref_vit_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                      C.c_void_p, C.c_void_p, C.c_int]
        self.lib = lib

    def sample(self, seed, M, config_L=100, calibrate=True):
        h = self.lib.ref_profile_sample(seed, M, config_L, 1 if calibrate else 0)
        if not h:
            raise RuntimeError("reference profile sampling failed")
        return RefProfile(self, h)

    def from_probs(self, mat, t, config_L=100, calibrate=False):
        M = mat.shape[0] - 1
        mat = np.ascontiguousarray(mat, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        h = self.lib.ref_profile_from_probs(M, _p(mat), _p(t), config_L, 1 if calibrate else 0)
        if not h:
            raise RuntimeError("reference profile construction failed")
        return RefProfile(self, h)

    def from_file(self, path, which=0):
        h = self.lib.ref_profile_from_file(path.encode(), which)
        if not h:
            raise RuntimeError(f"cannot read model {which} of {path}")
        return RefProfile(self, h)

    def iid(self, seed, n, L):
        out = np.zeros(n * L, np.uint8)
        self.lib.ref_iid(seed, n, L, _p(out))
        return out.reshape(n, L)

    def uniform(self, seed, fast, n):
        out = np.zeros(n, np.float64)
        self.lib.ref_uniform(seed, fast, n, _p(out))
        return out


class RefProfile:
    def __init__(self, ref, h):
        self.ref, self.h, self.lib = ref, h, ref.lib
        self.M = self.lib.ref_profile_M(h)

    def __del__(self):
        try:
            self.lib.ref_profile_free(self.h)
        except Exception:
            pass

    def byte_profile(self):
        kp = self.Kp
        rbv = np.zeros(kp * self.M, np.uint8)
        self.lib.ref_profile_rbv(self.h, _p(rbv))
        si = np.zeros(5, np.int32)
        sf = np.zeros(3, np.float32)
        self.lib.ref_profile_scalars(self.h, _p(si), _p(sf))
        return {"rbv": rbv.reshape(kp, self.M), "tbm": int(si[0]), "tec": int(si[1]),
                "base": int(si[3]), "bias": int(si[4]), "scale": float(sf[0]),
                "mu": float(sf[1]), "lambda": float(sf[2])}

    def hmm(self):
        mat = np.zeros((self.M + 1) * 20, np.float32)
        t = np.zeros((self.M + 1) * 7, np.float32)
        self.lib.ref_profile_hmm(self.h, _p(mat), _p(t))
        return mat.reshape(self.M + 1, 20), t.reshape(self.M + 1, 7)

    def msc(self):
        out = np.zeros(self.Kp * self.M, np.float32)
        self.lib.ref_profile_msc(self.h, _p(out))
        return out.reshape(self.Kp, self.M)

    def tjb(self, L):
        return self.lib.ref_tjb(self.h, L)

    @property
    def Kp(self):
        return self.lib.ref_profile_Kp(self.h)

    def digitize(self, text):
        out = np.zeros(len(text) + 1, np.uint8)
        n = self.lib.ref_digitize(self.h, text.encode(), _p(out))
        if n < 0:
            raise ValueError("text has characters outside the alphabet")
        return out[:n]

    @staticmethod
    def _dsq(x):
        return np.concatenate([[255], np.asarray(x, np.uint8), [255]]).astype(np.uint8)

    def msv(self, x):
        d = self._dsq(x)
        usc = C.c_float()
        st = self.lib.ref_msv(self.h, _p(d), len(x), C.byref(usc))
        return st, usc.value

    def ssv(self, x):
        d = self._dsq(x)
        usc = C.c_float()
        st = self.lib.ref_ssv(self.h, _p(d), len(x), C.byref(usc))
        return st, usc.value

    def nullsc(self, L):
        return self.lib.ref_nullsc(self.h, L)

    def passes(self, L, usc, F1=0.02):
        return self.lib.ref_pass(self.h, L, usc, F1)

    def msv_batch(self, res, off, lens, nthreads=8):
        res = np.ascontiguousarray(res, np.uint8)
        off = np.ascontiguousarray(off, np.int64)
        lens = np.ascontiguousarray(lens, np.int32)
        n = lens.size
        usc = np.zeros(n, np.float32)
        st = np.zeros(n, np.int32)
        self.lib.ref_msv_batch(self.h, _p(res), _p(off), _p(lens), n, _p(usc), _p(st), nthreads)
        return st, usc

    # ---- p7_ViterbiFilter ----
    def vit_fields(self):
        """The profile's ViterbiFilter fields, padded arrays over 8*Q positions
        plus the scalars (see ref_vit_fields)."""
        Q = self.lib.ref_vit_Q(self.h)
        W, kp = 8 * Q, self.Kp
        rwv = np.zeros(kp * W, np.int16)
        twv = np.zeros(8 * W, np.int16)
        ints = np.zeros(10, np.int32)
        flts = np.zeros(4, np.float32)
        self.lib.ref_vit_fields(self.h, _p(rwv), _p(twv), _p(ints), _p(flts))
        return {"M": self.M, "Q": Q, "rwv_pad": rwv.reshape(kp, W), "twv_pad": twv.reshape(8, W),
                "e_loop": int(ints[0]), "e_move": int(ints[1]), "n_move": int(ints[2]), "n_loop": int(ints[3]),
                "c_move": int(ints[4]), "c_loop": int(ints[5]), "j_move": int(ints[6]), "j_loop": int(ints[7]),
                "base_w": int(ints[8]), "ddbound_w": int(ints[9]), "scale_w": float(flts[0]),
                "nj": float(flts[1]), "vmu": float(flts[2]), "vlambda": float(flts[3])}

    def vit_move(self, L):
        return self.lib.ref_vit_move(self.h, int(L))

    def tsc(self):
        out = np.zeros((self.M + 1) * 8, np.float32)
        self.lib.ref_profile_tsc(self.h, _p(out))
        return out.reshape(self.M + 1, 8)

    def vit(self, x):
        d = self._dsq(x)
        sc = C.c_float()
        st = self.lib.ref_vit(self.h, _p(d), len(x), C.byref(sc))
        return st, sc.value

    def vit_batch(self, res, off, lens, nthreads=8):
        res = np.ascontiguousarray(res, np.uint8)
        off = np.ascontiguousarray(off, np.int64)
        lens = np.ascontiguousarray(lens, np.int32)
        n = lens.size
        sc = np.zeros(n, np.float32)
        st = np.zeros(n, np.int32)
        self.lib.ref_vit_batch(self.h, _p(res), _p(off), _p(lens), n, _p(sc), _p(st), nthreads)
        return st, sc

    def core_emit(self, seed, n, cap=None):
        cap = cap or n * (self.M * 40 + 2000)
        out = np.zeros(cap, np.uint8)
        lens = np.zeros(n, np.int32)
        tot = self.lib.ref_core_emit(self.h, seed, n, _p(out), cap, _p(lens))
        if tot < 0:
            raise RuntimeError("core_emit buffer too small")
        return out[:tot], lens
