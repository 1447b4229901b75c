"""cudahmmsearch_amd -- HMMER3's MSV filter (p7_MSVFilter) on AMD MI355X (gfx950).

The product is ``libmsvgpu.so`` (hand-written HIP kernels plus a C++ host
runtime behind the C ABI in ``include/msv_gpu.h`` and ``include/msv_prep.h``).
This module binds that ABI with ctypes and mirrors the reference's Python-less
C interface one to one:

=========================  ===================================================
this module                reference (forestcheng/cudahmmsearch, code/)
=========================  ===================================================
``sample_profile``         p7_hmm_Sample + p7_ProfileConfig + p7_oprofile_Convert
                           (src/p7_hmm.c:877, src/modelconfig.c:49,
                           src/impl_sse/p7_oprofile.c:840), lambda as p7_Lambda
``MSVContext.calibrate``   the MSV part of p7_Calibrate (src/evalues.c:60-141)
``MSVContext.msv_filter``  p7_MSVFilter (src/impl_sse/msvfilter.c:78)
``MSVContext.search``      cudaHmmsearch's block loop (src/GPUFunction.cu:332-492)
                           + the F1 rule of p7_Pipeline (src/p7_pipeline.c:667-677)
``core_emit`` / ``iid``    p7_CoreEmit (src/emit.c:71), esl_rsq_xfIID
=========================  ===================================================

No CPU fallback exists: the GPU entry points raise if the library or a GPU is
missing.
"""
import ctypes as C
import os

import numpy as np

__all__ = [
    "MSVError", "ByteProfile", "MSVContext", "Database", "load_library", "sample_profile",
    "profile_from_hmm", "core_emit", "iid", "tjb", "nullsc", "gen_database", "calibration_targets",
    "eslOK", "eslERANGE", "MODE_PLAIN", "MODE_BATCHED",
]

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmsvgpu.so")

eslOK, eslEMEM, eslEINVAL, eslERANGE, eslENORESULT = 0, 5, 11, 16, 19
MODE_PLAIN, MODE_BATCHED, FLAG_EAGER = 0, 1, 0x10
KP = 29
RNA, DNA, AMINO = 1, 2, 3                 # Easel's alphabet type codes (esl_alphabet.h:12-14)
_KP_OF = {RNA: 18, DNA: 18, AMINO: 29}


class MSVError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


class _Prof(C.Structure):
    _fields_ = [("M", C.c_int32), ("Kp", C.c_int32), ("rbv", C.c_void_p),
                ("tbm_b", C.c_uint8), ("tec_b", C.c_uint8), ("base_b", C.c_uint8), ("bias_b", C.c_uint8),
                ("scale_b", C.c_float), ("mmu", C.c_float), ("mlambda", C.c_float)]


class _Stats(C.Structure):
    _fields_ = [("ms_fast", C.c_double), ("ms_slow", C.c_double), ("ms_finalize", C.c_double),
                ("ms_total", C.c_double), ("cells", C.c_int64), ("n_slow", C.c_int64), ("n_jrow", C.c_int64),
                ("n_fast_launches", C.c_int32), ("dropin_path", C.c_int32), ("n_batch", C.c_int32),
                ("batch_lanes", C.c_int32 * 4), ("batch_cells", C.c_int64 * 4), ("batch_ms", C.c_double * 4),
                ("ms_jrow", C.c_double)]


class _Hmm(C.Structure):
    _fields_ = [("M", C.c_int), ("alphatype", C.c_int), ("K", C.c_int), ("Kp", C.c_int), ("mat", C.c_void_p),
                ("ins", C.c_void_p), ("t", C.c_void_p), ("compo", C.c_void_p), ("evparam", C.c_float * 6),
                ("has_stats", C.c_int), ("max_length", C.c_int), ("name", C.c_char_p), ("acc", C.c_char_p),
                ("desc", C.c_char_p)]


class _VProf(C.Structure):
    _fields_ = [("M", C.c_int32), ("Kp", C.c_int32), ("rwv", C.c_void_p), ("twv", C.c_void_p),
                ("e_loop", C.c_int16), ("e_move", C.c_int16), ("base_w", C.c_int16), ("ddbound_w", C.c_int16),
                ("scale_w", C.c_float), ("nj", C.c_float), ("vmu", C.c_float), ("vlambda", C.c_float)]


class _QRes(C.Structure):
    _fields_ = [("ticket", C.c_int64), ("status", C.c_int32), ("M", C.c_int32), ("n_pass", C.c_int64),
                ("cells", C.c_int64), ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_done", C.c_double),
                ("batch", C.c_int32)]


class _SInfo(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("admitted", C.c_int64), ("completed", C.c_int64),
                ("pending", C.c_int64), ("running", C.c_int64), ("max_pending", C.c_int64),
                ("busy_s", C.c_double), ("uptime_s", C.c_double), ("cached_profiles", C.c_int64),
                ("batches", C.c_int64)]


class _Rng(C.Structure):
    _fields_ = [("fast", C.c_int), ("mt", C.c_uint32 * 624), ("mti", C.c_int), ("x", C.c_uint32),
                ("seed", C.c_uint32)]


_lib = None


def load_library(path=LIB_PATH):
    """Load libmsvgpu.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise MSVError(-1, f"{path} is not built (run __graft_entry__.build() or make -C cudahmmsearch_amd/csrc)")
    lib = C.CDLL(path)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    sig = {
        "msvg_init": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "msvg_free": (None, [vp]),
        "msvg_last_error": (C.c_char_p, [vp]),
        "msvg_profile_add": (C.c_int, [vp, C.POINTER(_Prof), C.POINTER(i32)]),
        "msvg_profile_remove": (C.c_int, [vp, i32]),
        "msvg_profile_set_evparams": (C.c_int, [vp, i32, C.c_float, C.c_float]),
        "msvg_calibrate_mu": (C.c_int, [vp, i32, C.c_double, C.POINTER(C.c_float)]),
        "msvg_db_add": (C.c_int, [vp, vp, vp, vp, i64, C.POINTER(i32)]),
        "msvg_db_remove": (C.c_int, [vp, i32]),
        "msvg_db_size": (i64, [vp, i32]),
        "msvg_search": (C.c_int, [vp, vp, i32, i32, C.c_double, C.c_int, vp, vp, vp]),
        "msvg_search_device": (C.c_int, [vp, vp, i32, i32, C.c_double, C.c_int]),
        "msvg_results": (C.c_int, [vp, vp, vp, vp]),
        "msvg_search_noalloc": (C.c_int, [vp, vp, i32, i32, C.c_double, vp, vp, vp]),
        "msvg_last_stats": (C.c_int, [vp, C.POINTER(_Stats)]),
        "msvg_MSVFilter": (C.c_int, [vp, i32, vp, C.c_int, C.POINTER(C.c_float)]),
        "msvg_set_tile": (C.c_int, [vp, i32, i32, i32, i32]),
        "msvg_set_vit_tile": (C.c_int, [vp, i32, i32, i32, i32]),
        "msvg_set_vit_vote": (C.c_int, [vp, i32]),
        "msvg_set_vit_segments": (C.c_int, [vp, i32]),
        "msvg_set_tail_guard": (C.c_int, [vp, C.c_int]),
        "msvg_probe_lds": (C.c_int, [vp, C.POINTER(C.c_double)]),
        "msvg_probe_forms": (C.c_int, [vp, i32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                       C.POINTER(C.c_int32)]),
        "msvg_tile_of": (C.c_int, [vp, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
        "msvg_profile_set_viterbi": (C.c_int, [vp, i32, C.POINTER(_VProf)]),
        "msvg_vit_search": (C.c_int, [vp, i32, i32, vp, i64, vp, vp]),
        "msvg_vit_search_many": (C.c_int, [vp, i32, vp, i32, vp, vp, vp, vp]),
        "msvg_scan": (C.c_int, [vp, i32, vp, vp, C.c_int, C.c_double, vp, vp, vp]),
        "msvg_scan_many": (C.c_int, [vp, i32, vp, i32, vp, vp, vp, C.c_double, vp, vp, vp]),
        "msvg_ViterbiFilter": (C.c_int, [vp, i32, vp, C.c_int, C.POINTER(C.c_float)]),
        "msvg_calibrate_vmu": (C.c_int, [vp, i32, C.c_double, C.POINTER(C.c_float)]),
        "msvh_vit_move": (C.c_int, [C.c_float, C.c_float, C.c_int]),
        "msvh_pressed_open": (C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(vp)]),
        "msvh_pressed_next": (C.c_int, [vp, C.POINTER(_Prof), C.POINTER(_VProf), C.c_char_p, C.c_int]),
        "msvh_pressed_close": (None, [vp]),
        "msvh_vit_convert": (C.c_int, [C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "msvh_vit_convert_abc": (C.c_int, [C.c_int, C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_float),
                                           C.POINTER(C.c_float)]),
        "msvh_profile_convert_abc": (C.c_int, [C.c_int, C.c_int, vp, C.c_int, vp, vp, C.POINTER(C.c_float), vp]),
        "msvh_lambda_abc": (C.c_double, [C.c_int, C.c_int, vp]),
        "msvh_hmmfile_open": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
        "msvh_hmmfile_next": (C.c_int, [vp, C.POINTER(_Hmm)]),
        "msvh_hmmfile_error": (C.c_char_p, [vp]),
        "msvh_hmmfile_close": (None, [vp]),
        "msvh_vit_score": (C.c_float, [C.c_int, C.c_int, C.c_int, C.c_float]),
        "msvh_fit_vmu": (C.c_double, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_double]),
        "msvh_vit_calibration_targets": (None, [C.c_int, C.c_int, vp]),
        "msvg_test_alloc_budget": (None, [C.c_long]),
        "msvg_server_start": (C.c_int, [vp, i32, i32, C.c_double, i64, i32, C.POINTER(vp)]),
        "msvg_server_submit": (C.c_int, [vp, C.POINTER(_Prof), C.POINTER(i64)]),
        "msvg_server_wait": (C.c_int, [vp, i64, C.c_double, C.POINTER(_QRes), vp, vp, i64]),
        "msvg_server_stats": (C.c_int, [vp, C.POINTER(_SInfo)]),
        "msvg_server_stop": (C.c_int, [vp]),
        "msvh_rng_init": (None, [C.POINTER(_Rng), C.c_uint32, C.c_int]),
        "msvh_random": (C.c_double, [C.POINTER(_Rng)]),
        "msvh_hmm_sample": (C.c_int, [C.POINTER(_Rng), C.c_int, vp, vp, vp]),
        "msvh_profile_convert": (C.c_int, [C.c_int, vp, C.c_int, vp, vp, C.POINTER(C.c_float), vp]),
        "msvh_lambda": (C.c_double, [C.c_int, vp]),
        "msvh_core_emit": (C.c_int, [C.POINTER(_Rng), C.c_int, vp, vp, vp, vp, C.c_int]),
        "msvh_iid": (None, [C.POINTER(_Rng), C.c_int, vp]),
        "msvh_tjb": (C.c_int, [C.c_int]),
        "msvh_nullsc": (C.c_float, [C.c_int]),
        "msvh_fit_mu": (C.c_double, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_double]),
        "msvh_calibration_targets": (None, [C.c_int, C.c_int, vp]),
        "msvh_gen_lengths": (i64, [C.c_uint32, i64, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                   C.c_double, C.c_double, C.c_int, vp]),
        "msvh_gen_residues": (None, [C.c_uint32, i64, vp, vp, vp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------------------
# Host-side profile preparation (no GPU needed)
# ---------------------------------------------------------------------------
class Rng:
    """Easel's RNG: Mersenne twister (esl_randomness_Create) or, with
    fast=True, the LCG of esl_randomness_CreateFast."""

    def __init__(self, seed, fast=False):
        self._r = _Rng()
        load_library().msvh_rng_init(C.byref(self._r), seed, 1 if fast else 0)

    def random(self):
        return load_library().msvh_random(C.byref(self._r))


class ByteProfile:
    """The MSV fields of a P7_OPROFILE (impl_sse.h:79-86), de-striped.

    ``rbv`` is (29, M) uint8; scalars as in the reference; ``mu``/``lam`` are
    the calibrated MSV Gumbel parameters (``None`` until calibrated)."""

    def __init__(self, rbv, tbm, tec, base, bias, scale, mu=None, lam=None, hmm=None, tjb_config=None,
                 alphabet=AMINO):
        self.rbv = np.ascontiguousarray(rbv, dtype=np.uint8)
        self.alphabet = alphabet
        self.M = self.rbv.shape[1]
        self.tbm, self.tec, self.base, self.bias = int(tbm), int(tec), int(base), int(bias)
        self.scale = float(scale)
        self.mu, self.lam = mu, lam
        self.hmm = hmm
        self.tjb_config = tjb_config

    def as_dict(self):
        return {"rbv": self.rbv, "tbm": self.tbm, "tec": self.tec, "base": self.base, "bias": self.bias,
                "scale": self.scale, "mu": self.mu, "lambda": self.lam}

    def viterbi(self):
        """The ViterbiFilter fields of the same HMM (needs the HMM this
        profile was converted from); vlambda = the MSV lambda, as p7_Calibrate
        sets it (evalues.c:112-113)."""
        if self.hmm is None or self.hmm[2] is None:
            raise MSVError(eslEINVAL, "profile has no HMM transitions to convert")
        mat, ins, t = self.hmm
        return viterbi_from_hmm(mat, t, vlambda=self.lam, alphabet=self.alphabet)

    def _struct(self):
        return _Prof(self.M, self.rbv.shape[0], self.rbv.ctypes.data, self.tbm, self.tec, self.base, self.bias,
                     self.scale, float(self.mu if self.mu is not None else -99999.0),
                     float(self.lam if self.lam is not None else -99999.0))


class ViterbiProfile:
    """The ViterbiFilter fields of a P7_OPROFILE (impl_sse.h:88-101) in model
    order: ``rwv`` (Kp, M) int16 match emissions, ``twv`` (8, M) int16
    transitions as the reference's striped vectors hold them at node k (BM,
    MM, IM, DM into k; MD, MI, II, DD out of k), the E-state costs, base_w,
    ddbound_w, scale_w, nj and the calibrated Viterbi Gumbel parameters."""

    def __init__(self, rwv, twv, e_loop, e_move, base_w, ddbound_w, scale_w, nj=1.0, vmu=None, vlambda=None):
        self.rwv = np.ascontiguousarray(rwv, dtype=np.int16)
        self.twv = np.ascontiguousarray(twv, dtype=np.int16)
        self.M = self.rwv.shape[1]
        self.e_loop, self.e_move, self.base_w, self.ddbound_w = int(e_loop), int(e_move), int(base_w), int(ddbound_w)
        self.scale_w, self.nj = float(scale_w), float(nj)
        self.vmu, self.vlambda = vmu, vlambda

    @classmethod
    def from_dict(cls, d):
        return cls(d["rwv"], d["twv"], d["e_loop"], d["e_move"], d["base_w"], d["ddbound_w"], d["scale_w"],
                   d.get("nj", 1.0), d.get("vmu"), d.get("vlambda"))

    def as_dict(self):
        return {"rwv": self.rwv, "twv": self.twv, "e_loop": self.e_loop, "e_move": self.e_move,
                "base_w": self.base_w, "ddbound_w": self.ddbound_w, "scale_w": self.scale_w, "nj": self.nj,
                "vmu": self.vmu, "vlambda": self.vlambda}

    def _struct(self):
        return _VProf(self.M, self.rwv.shape[0], self.rwv.ctypes.data, self.twv.ctypes.data, self.e_loop,
                      self.e_move, self.base_w, self.ddbound_w, self.scale_w, self.nj,
                      float(self.vmu if self.vmu is not None else -99999.0),
                      float(self.vlambda if self.vlambda is not None else -99999.0))


def viterbi_from_hmm(mat, t, vlambda=None, alphabet=AMINO):
    """vf_conversion (p7_oprofile.c:653-740) of an HMM configured as hmmsearch
    configures it: the ViterbiProfile for match emissions ``mat`` ((M+1) x K)
    and transitions ``t`` ((M+1) x 7, p7H order)."""
    lib = load_library()
    mat = np.ascontiguousarray(mat, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    M = mat.shape[0] - 1
    rwv = np.zeros((_KP_OF[alphabet], M), np.int16)
    twv = np.zeros((8, M), np.int16)
    sc = np.zeros(4, np.int32)
    scale_w, nj = C.c_float(), C.c_float()
    rc = lib.msvh_vit_convert_abc(alphabet, M, _p(mat), _p(t), _p(rwv), _p(twv), _p(sc), C.byref(scale_w),
                                  C.byref(nj))
    if rc != 0:
        raise MSVError(rc, "Viterbi conversion failed")
    return ViterbiProfile(rwv, twv, sc[0], sc[1], sc[2], sc[3], scale_w.value, nj.value, vlambda=vlambda)


def read_pressed(h3f, h3p=None):
    """Models of an hmmpress'ed database (format 3/f): a list of (name,
    ByteProfile, ViterbiProfile or None), de-striped from the .h3f (MSV part)
    and, when given, the .h3p (ViterbiFilter part), with the calibrated
    E-value parameters the files carry."""
    lib = load_library()
    h = C.c_void_p()
    rc = lib.msvh_pressed_open(h3f.encode(), h3p.encode() if h3p else None, C.byref(h))
    if rc != 0:
        raise MSVError(rc, f"cannot open {h3f}" + (f" / {h3p}" if h3p else ""))
    out = []
    try:
        while True:
            m, v = _Prof(), _VProf()
            name = C.create_string_buffer(256)
            rc = lib.msvh_pressed_next(h, C.byref(m), C.byref(v), name, 256)
            if rc == 3:
                break
            if rc != 0:
                raise MSVError(rc, f"bad pressed file {h3f} (model {len(out)})")
            rbv = np.ctypeslib.as_array(C.cast(m.rbv, C.POINTER(C.c_uint8)), (m.Kp, m.M)).copy()
            prof = ByteProfile(rbv, m.tbm_b, m.tec_b, m.base_b, m.bias_b, m.scale_b, mu=m.mmu, lam=m.mlambda)
            vit = None
            if v.M:
                rwv = np.ctypeslib.as_array(C.cast(v.rwv, C.POINTER(C.c_int16)), (v.Kp, v.M)).copy()
                twv = np.ctypeslib.as_array(C.cast(v.twv, C.POINTER(C.c_int16)), (8, v.M)).copy()
                vit = ViterbiProfile(rwv, twv, v.e_loop, v.e_move, v.base_w, v.ddbound_w, v.scale_w, v.nj,
                                     vmu=v.vmu, vlambda=v.vlambda)
            out.append((name.value.decode(), prof, vit))
    finally:
        lib.msvh_pressed_close(h)
    return out


def vit_move(scale_w, nj, L):
    """xw[N|C|J][MOVE] for target length L (p7_oprofile.c:944-962)."""
    return load_library().msvh_vit_move(float(scale_w), float(nj), int(L))


def profile_from_hmm(mat, ins=None, t=None, config_L=100, alphabet=AMINO):
    """p7_ProfileConfig match scores + p7_oprofile_Convert's MSV part for an
    HMM's match emissions ``mat`` ((M+1) x K, row 0 unused; K = 20 amino,
    4 DNA/RNA); lambda as p7_Lambda."""
    lib = load_library()
    mat = np.ascontiguousarray(mat, dtype=np.float32)
    M = mat.shape[0] - 1
    rbv = np.zeros((_KP_OF[alphabet], M), np.uint8)
    sc = np.zeros(5, np.int32)
    scale = C.c_float()
    rc = lib.msvh_profile_convert_abc(alphabet, M, _p(mat), config_L, _p(rbv), _p(sc), C.byref(scale), None)
    if rc != 0:
        raise MSVError(rc, "profile conversion failed")
    lam = float(np.float32(lib.msvh_lambda_abc(alphabet, M, _p(mat))))
    return ByteProfile(rbv, sc[0], sc[1], sc[3], sc[4], scale.value, lam=lam,
                       hmm=(mat, ins, t), tjb_config=int(sc[2]), alphabet=alphabet)


def read_hmm(path, config_L=100):
    """Models of an HMMER3 ASCII file (.hmm, formats 3/a to 3/f; the parser
    of p7_hmmfile_Read, p7_hmmfile.c:1295-1630): a list of (name,
    ByteProfile, ViterbiProfile), converted as hmmsearch configures a model
    (p7_ProfileConfig local multihit at L = config_L, p7_oprofile_Convert),
    with the file's STATS lines as the E-value parameters (mu None and lambda
    from p7_Lambda when the file has none).  Each ByteProfile keeps the HMM
    (mat, ins, t) it came from."""
    lib = load_library()
    h = C.c_void_p()
    rc = lib.msvh_hmmfile_open(path.encode(), C.byref(h))
    if rc != 0:
        raise MSVError(rc, f"cannot open {path} as an HMMER3 ASCII file")
    out = []
    try:
        while True:
            m = _Hmm()
            rc = lib.msvh_hmmfile_next(h, C.byref(m))
            err = lib.msvh_hmmfile_error(h).decode()
            if rc == 3 and not err:
                break
            if rc != 0:
                raise MSVError(rc, f"{path}, model {len(out) + 1}: {err}")
            M, K = m.M, m.K
            mat = np.ctypeslib.as_array(C.cast(m.mat, C.POINTER(C.c_float)), ((M + 1) * K,)).reshape(M + 1, K).copy()
            ins = np.ctypeslib.as_array(C.cast(m.ins, C.POINTER(C.c_float)), ((M + 1) * K,)).reshape(M + 1, K).copy()
            t = np.ctypeslib.as_array(C.cast(m.t, C.POINTER(C.c_float)), ((M + 1) * 7,)).reshape(M + 1, 7).copy()
            prof = profile_from_hmm(mat, ins, t, config_L=config_L, alphabet=m.alphatype)
            ev = [float(x) for x in m.evparam]
            vit = viterbi_from_hmm(mat, t, alphabet=m.alphatype)
            if m.has_stats:
                prof.mu, prof.lam = ev[0], ev[1]
                vit.vmu, vit.vlambda = ev[2], ev[3]
            else:
                prof.mu, vit.vlambda = None, prof.lam
            out.append((m.name.decode(), prof, vit))
    finally:
        lib.msvh_hmmfile_close(h)
    return out


def match_scores(mat, config_L=100):
    """Float match scores p7P_MSC (29 x M) as p7_ProfileConfig sets them."""
    lib = load_library()
    mat = np.ascontiguousarray(mat, dtype=np.float32)
    M = mat.shape[0] - 1
    rbv = np.zeros((KP, M), np.uint8)
    msc = np.zeros((KP, M), np.float32)
    sc = np.zeros(5, np.int32)
    scale = C.c_float()
    lib.msvh_profile_convert(M, _p(mat), config_L, _p(rbv), _p(sc), C.byref(scale), _p(msc))
    return msc


def sample_hmm(seed, M):
    """p7_hmm_Sample with esl_randomness_Create(seed): (mat, ins, t)."""
    lib = load_library()
    r = Rng(seed)
    mat = np.zeros((M + 1, 20), np.float32)
    ins = np.zeros((M + 1, 20), np.float32)
    t = np.zeros((M + 1, 7), np.float32)
    rc = lib.msvh_hmm_sample(C.byref(r._r), M, _p(mat), _p(ins), _p(t))
    if rc != 0:
        raise MSVError(rc, "hmm sampling failed")
    return mat, ins, t


def sample_profile(seed, M, config_L=100):
    """A sampled amino HMM turned into a byte profile (uncalibrated mu)."""
    mat, ins, t = sample_hmm(seed, M)
    return profile_from_hmm(mat, ins, t, config_L)


def core_emit(hmm, seed, n, rng=None):
    """n p7_CoreEmit samples (list of uint8 arrays) from esl_randomness_Create(seed)."""
    lib = load_library()
    mat, ins, t = (np.ascontiguousarray(a, dtype=np.float32) for a in hmm)
    M = mat.shape[0] - 1
    r = rng or Rng(seed)
    out = []
    cap = 64 * M + 4096
    buf = np.zeros(cap, np.uint8)
    saved = _Rng()
    for _ in range(n):
        C.memmove(C.byref(saved), C.byref(r._r), C.sizeof(_Rng))
        while True:
            L = lib.msvh_core_emit(C.byref(r._r), M, _p(mat), _p(ins), _p(t), _p(buf), cap)
            if L >= 0:
                break
            if L != -1:
                raise MSVError(L, "core emission reached an impossible state")
            # buffer too small: restore the generator and retry with room to spare
            C.memmove(C.byref(r._r), C.byref(saved), C.sizeof(_Rng))
            cap *= 4
            buf = np.zeros(cap, np.uint8)
        out.append(buf[:L].copy())
    return out


def iid(seed, n, L):
    """n i.i.d. background targets of length L from esl_randomness_Create(seed)."""
    lib = load_library()
    r = Rng(seed)
    out = np.zeros((n, L), np.uint8)
    for i in range(n):
        lib.msvh_iid(C.byref(r._r), L, _p(out[i]))
    return out


def tjb(L):
    return load_library().msvh_tjb(int(L))


def nullsc(L):
    return load_library().msvh_nullsc(int(L))


def calibration_targets(N=200, L=200):
    out = np.zeros(N * L, np.uint8)
    load_library().msvh_calibration_targets(N, L, _p(out))
    return out.reshape(N, L)


class Database:
    """Concatenated digital targets: ``res`` (uint8), ``off`` (int64), ``lens`` (int32)."""

    def __init__(self, res, off, lens):
        self.res = np.ascontiguousarray(res, dtype=np.uint8)
        self.off = np.ascontiguousarray(off, dtype=np.int64)
        self.lens = np.ascontiguousarray(lens, dtype=np.int32)

    @property
    def n(self):
        return int(self.lens.size)

    @classmethod
    def from_list(cls, seqs):
        lens = np.array([len(s) for s in seqs], np.int32)
        off = np.zeros(len(seqs), np.int64)
        if len(seqs):
            off[1:] = np.cumsum(lens[:-1])
        res = np.concatenate([np.asarray(s, np.uint8) for s in seqs]) if len(seqs) else np.zeros(0, np.uint8)
        return cls(res, off, lens)

    def seq(self, i):
        return self.res[self.off[i]:self.off[i] + self.lens[i]]

    @property
    def residues(self):
        return int(self.lens.sum(dtype=np.int64))


def gen_database(seed, n, median, sigma, lo, hi, tail_frac=0.0, tail_xm=1000.0, tail_alpha=1.5, tail_max=35000,
                 nthreads=8):
    """Synthetic i.i.d. background database (see msvh_gen_lengths in msv_prep.h)."""
    lib = load_library()
    lens = np.zeros(n, np.int32)
    total = lib.msvh_gen_lengths(seed, n, median, sigma, lo, hi, tail_frac, tail_xm, tail_alpha, tail_max,
                                 _p(lens))
    off = np.zeros(n, np.int64)
    if n:
        off[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    res = np.zeros(max(int(total), 1), np.uint8)
    lib.msvh_gen_residues(seed, n, _p(lens), _p(off), _p(res), nthreads)
    return Database(res[:total], off, lens)


def plant(db, seqs, seed, frac):
    """Replace a fraction of targets (chosen by numpy seed) with ``seqs``;
    returns a new Database and the replaced indices."""
    rs = np.random.default_rng(seed)
    k = min(len(seqs), int(round(frac * db.n)))
    idx = np.sort(rs.choice(db.n, size=k, replace=False))
    seqs_all = None
    lens = db.lens.copy()
    lens[idx] = [len(s) for s in seqs[:k]]
    off = np.zeros(db.n, np.int64)
    off[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    res = np.empty(int(lens.sum(dtype=np.int64)), np.uint8)
    mask = np.ones(db.n, bool)
    mask[idx] = False
    # copy unchanged targets in runs
    keep = np.flatnonzero(mask)
    for i in keep:
        res[off[i]:off[i] + lens[i]] = db.res[db.off[i]:db.off[i] + db.lens[i]]
    for j, i in enumerate(idx):
        res[off[i]:off[i] + lens[i]] = seqs[j]
    del seqs_all
    return Database(res, off, lens), idx


# ---------------------------------------------------------------------------
# GPU context
# ---------------------------------------------------------------------------
class MSVContext:
    """One GPU, one context (not shared between host threads)."""

    def __init__(self, device=0, tiles="default"):
        """tiles: "default" loads the persisted autotuner tables if present
        (the batched MSV pass's and the ViterbiFilter shapes), None uses the
        built-in rules, or a list of MSV tile dicts."""
        self.lib = load_library()
        self._c = C.c_void_p()
        rc = self.lib.msvg_init(device, C.byref(self._c))
        if rc != 0:
            raise MSVError(rc, f"msvg_init failed on device {device} (is a GPU visible?)")
        self._keep = {}
        vit_tiles = None
        if tiles == "default":
            from . import autotune
            tiles = autotune.load()
            vit_tiles = autotune.load(key="vit_tiles")
        if tiles:
            for t in tiles:
                self.set_tile(t["lo"], t["hi"], t["lanes"], t["waves"])
        if vit_tiles:
            for t in vit_tiles:
                self.set_vit_tile(t["lo"], t["hi"], t["lanes"], t["waves"])

    def close(self):
        if self._c:
            self.lib.msvg_free(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != 0:
            raise MSVError(rc, self.lib.msvg_last_error(self._c).decode())

    def add_profile(self, prof):
        pid = C.c_int32()
        s = prof._struct()
        self._chk(self.lib.msvg_profile_add(self._c, C.byref(s), C.byref(pid)))
        return pid.value

    def remove_profile(self, pid):
        self._chk(self.lib.msvg_profile_remove(self._c, pid))

    def set_evparams(self, pid, mu, lam):
        self._chk(self.lib.msvg_profile_set_evparams(self._c, pid, float(mu), float(lam)))

    def calibrate(self, pid, lam):
        """GPU p7_MSVMu; returns mu (float32) and stores (mu, lam) on the profile."""
        mu = C.c_float()
        self._chk(self.lib.msvg_calibrate_mu(self._c, pid, float(lam), C.byref(mu)))
        self.set_evparams(pid, mu.value, lam)
        return mu.value

    def add_database(self, db):
        did = C.c_int32()
        # a database of empty targets has no residues; the ABI still takes a
        # residue pointer when n > 0
        res = db.res if db.res.size else np.zeros(1, np.uint8)
        self._chk(self.lib.msvg_db_add(self._c, _p(res) if db.n else None, _p(db.off), _p(db.lens),
                                       db.n, C.byref(did)))
        return did.value

    def remove_database(self, did):
        self._chk(self.lib.msvg_db_remove(self._c, did))

    def search(self, pids, did, F1=0.02, mode=MODE_BATCHED):
        """Scores of every profile against every target: (usc, status, pass),
        each shaped [len(pids), n] in the caller's target order."""
        ids = np.ascontiguousarray(pids, dtype=np.int32)
        n = self.lib.msvg_db_size(self._c, did)
        if n < 0:
            raise MSVError(eslEINVAL, f"no database {did}")
        usc = np.zeros((ids.size, n), np.float32)
        st = np.zeros((ids.size, n), np.uint8)
        ps = np.zeros((ids.size, n), np.uint8)
        self._chk(self.lib.msvg_search(self._c, _p(ids), ids.size, did, F1, mode, _p(usc), _p(st), _p(ps)))
        return usc, st, ps

    def search_device(self, pids, did, F1=0.02, mode=MODE_BATCHED):
        ids = np.ascontiguousarray(pids, dtype=np.int32)
        self._chk(self.lib.msvg_search_device(self._c, _p(ids), ids.size, did, F1, mode))

    def stats(self):
        s = _Stats()
        self._chk(self.lib.msvg_last_stats(self._c, C.byref(s)))
        out = {}
        for k, _ in _Stats._fields_:
            v = getattr(s, k)
            out[k] = list(v) if isinstance(v, C.Array) else v
        return out

    def msv_filter(self, pid, x):
        """p7_MSVFilter on one target (codes, no sentinels): (status, usc)."""
        dsq = np.concatenate([[255], np.asarray(x, np.uint8), [255]]).astype(np.uint8)
        sc = C.c_float()
        rc = self.lib.msvg_MSVFilter(self._c, pid, _p(dsq), len(x), C.byref(sc))
        if rc not in (eslOK, eslERANGE):
            self._chk(rc)
        return rc, sc.value

    # ---- ViterbiFilter stage ----
    def set_viterbi(self, pid, vprof):
        """Attach a ViterbiProfile to profile pid (same M)."""
        s = vprof._struct()
        self._chk(self.lib.msvg_profile_set_viterbi(self._c, pid, C.byref(s)))

    def vit_search(self, pid, did, targets):
        """p7_ViterbiFilter on the listed targets (caller indices): (sc, status)."""
        t = np.ascontiguousarray(targets, dtype=np.int64)
        sc = np.zeros(t.size, np.float32)
        st = np.zeros(t.size, np.uint8)
        self._chk(self.lib.msvg_vit_search(self._c, pid, did, _p(t), t.size, _p(sc), _p(st)))
        return sc, st

    def scan(self, pids, x, F1=0.02):
        """hmmscan direction: the profiles pids against one query x (codes, no
        sentinels): (usc, status, pass), one entry per profile."""
        pids = np.ascontiguousarray(pids, dtype=np.int32)
        dsq = np.concatenate([[255], np.asarray(x, np.uint8), [255]]).astype(np.uint8)
        n = max(pids.size, 1)
        usc = np.zeros(n, np.float32)
        st = np.zeros(n, np.uint8)
        ps = np.zeros(n, np.uint8)
        self._chk(self.lib.msvg_scan(self._c, pids.size, _p(pids), _p(dsq), len(x), float(F1), _p(usc), _p(st), _p(ps)))
        return usc[:pids.size], st[:pids.size], ps[:pids.size]

    def scan_many(self, pids, queries, F1=0.02):
        """scan() for a batch of queries (a Database or a list of code
        arrays): (usc, status, pass), each [query][profile]."""
        pids = np.ascontiguousarray(pids, dtype=np.int32)
        db = queries if isinstance(queries, Database) else Database.from_list(list(queries))
        res = db.res if db.res.size else np.zeros(1, np.uint8)
        off = np.ascontiguousarray(db.off, np.int64)
        lens = np.ascontiguousarray(db.lens, np.int32)
        n = max(db.n * pids.size, 1)
        usc = np.zeros(n, np.float32)
        st = np.zeros(n, np.uint8)
        ps = np.zeros(n, np.uint8)
        self._chk(self.lib.msvg_scan_many(self._c, pids.size, _p(pids), db.n, _p(res), _p(off), _p(lens), float(F1),
                                          _p(usc), _p(st), _p(ps)))
        shape = (db.n, pids.size)
        k = db.n * pids.size
        return usc[:k].reshape(shape), st[:k].reshape(shape), ps[:k].reshape(shape)

    def vit_search_many(self, pids, did, target_lists):
        """vit_search for several profiles in one call (launches overlap on
        the context's streams): a list of (sc, status), one per profile."""
        pids = np.ascontiguousarray(pids, dtype=np.int32)
        lists = [np.ascontiguousarray(t, dtype=np.int64) for t in target_lists]
        if len(lists) != pids.size:
            raise ValueError("one target list per profile")
        offs = np.zeros(pids.size + 1, np.int64)
        offs[1:] = np.cumsum([t.size for t in lists], dtype=np.int64)
        t = np.concatenate(lists) if lists else np.zeros(0, np.int64)
        sc = np.zeros(max(t.size, 1), np.float32)
        st = np.zeros(max(t.size, 1), np.uint8)
        self._chk(self.lib.msvg_vit_search_many(self._c, pids.size, _p(pids), did, _p(offs), _p(t), _p(sc), _p(st)))
        return [(sc[offs[i]:offs[i + 1]].copy(), st[offs[i]:offs[i + 1]].copy()) for i in range(pids.size)]

    def vit_filter(self, pid, x):
        """p7_ViterbiFilter on one target (codes, no sentinels): (status, sc)."""
        dsq = np.concatenate([[255], np.asarray(x, np.uint8), [255]]).astype(np.uint8)
        sc = C.c_float()
        rc = self.lib.msvg_ViterbiFilter(self._c, pid, _p(dsq), len(x), C.byref(sc))
        if rc not in (eslOK, eslERANGE):
            self._chk(rc)
        return rc, sc.value

    def calibrate_vmu(self, pid, lam):
        mu = C.c_float()
        self._chk(self.lib.msvg_calibrate_vmu(self._c, pid, float(lam), C.byref(mu)))
        return mu.value

    def tile_of(self, M, did):
        """(lanes, waves, segmented) of the batched mode for length M."""
        ln, wv, sg = C.c_int32(), C.c_int32(), C.c_int32()
        self._chk(self.lib.msvg_tile_of(self._c, M, did, C.byref(ln), C.byref(wv), C.byref(sg)))
        return ln.value, wv.value, bool(sg.value)

    def probe_lds(self):
        """Sustained LDS read bandwidth of the device, bytes/s."""
        v = C.c_double()
        self._chk(self.lib.msvg_probe_lds(self._c, C.byref(v)))
        return v.value

    def probe_forms(self, L=16384):
        """The packed-form probe (msvg_probe_forms): {"ms_f16", "ms_u16",
        "cells", "agree"} for one launch of each form."""
        a, b, n, ok = C.c_double(), C.c_double(), C.c_int64(), C.c_int32()
        self._chk(self.lib.msvg_probe_forms(self._c, L, C.byref(a), C.byref(b), C.byref(n), C.byref(ok)))
        return {"ms_f16": a.value, "ms_u16": b.value, "cells": n.value, "agree": bool(ok.value)}

    def set_tile(self, m_lo, m_hi, lanes, waves=0):
        self._chk(self.lib.msvg_set_tile(self._c, m_lo, m_hi, lanes, waves))

    def clear_tiles(self):
        self._chk(self.lib.msvg_set_tile(self._c, 0, 0, 0, 0))

    def set_vit_tile(self, m_lo, m_hi, lanes, waves=1):
        """ViterbiFilter shape for models of length m_lo..m_hi, applied by the
        next set_viterbi (msvg_set_vit_tile)."""
        self._chk(self.lib.msvg_set_vit_tile(self._c, m_lo, m_hi, lanes, waves))

    def clear_vit_tiles(self):
        self._chk(self.lib.msvg_set_vit_tile(self._c, 0, 0, 0, 0))

    def set_vit_vote(self, mode):
        """-1: profiles set after this vote on the lazy-F test by their
        estimated rate (default); 0: every row closes the delete paths; 1:
        every row votes (msvg_set_vit_vote).  Results do not change."""
        self._chk(self.lib.msvg_set_vit_vote(self._c, mode))

    def set_vit_segments(self, on):
        """True: models beyond one wave's ViterbiFilter tables run as column
        segments (default); False: on the multi-wave kernel
        (msvg_set_vit_segments).  Results do not change."""
        self._chk(self.lib.msvg_set_vit_segments(self._c, 1 if on else 0))

    def set_tail_guard(self, on):
        """Batched mode: widen the lane group where the longest target would
        set the launch time (on by default)."""
        self._chk(self.lib.msvg_set_tail_guard(self._c, 1 if on else 0))


class Server:
    """Continuous-batching MSV server over one resident database (see
    msvg_server_* in include/msv_gpu.h).  While it runs, make no other calls
    on its context."""

    def __init__(self, ctx, did, slots=4, F1=0.02, hbm_budget=0, shared=False, batched=False):
        """Default: persistent launches, one query per search (queries start
        as the running one drains); shared=True: one chunk per workgroup
        (running queries share the CUs from the start); batched=True: a free
        slot takes every waiting query (up to 16) into one multi-profile
        search."""
        if shared and batched:
            raise ValueError("shared and batched serving exclude each other")
        self.ctx, self.lib = ctx, ctx.lib
        self.n = self.lib.msvg_db_size(ctx._c, did)
        if self.n < 0:
            raise MSVError(eslEINVAL, f"no database {did}")
        self._s = C.c_void_p()
        ctx._chk(self.lib.msvg_server_start(ctx._c, did, slots, F1, hbm_budget, 2 if batched else 1 if shared else 0,
                                            C.byref(self._s)))

    def submit(self, prof):
        t = C.c_int64()
        s = prof._struct()
        rc = self.lib.msvg_server_submit(self._s, C.byref(s), C.byref(t))
        if rc != 0:
            raise MSVError(rc, "submit refused")
        return t.value

    def wait(self, ticket, timeout=-1.0, want_hits=True):
        """None if not finished within timeout; else (info dict, idx, usc)."""
        r = _QRes()
        cap = self.n if want_hits else 0
        idx = np.zeros(max(cap, 1), np.int32)
        usc = np.zeros(max(cap, 1), np.float32)
        rc = self.lib.msvg_server_wait(self._s, ticket, timeout, C.byref(r), _p(idx) if cap else None,
                                       _p(usc) if cap else None, cap)
        if rc == eslENORESULT:
            return None
        if rc != 0:
            raise MSVError(rc, "wait failed")
        info = {k: getattr(r, k) for k, _ in _QRes._fields_}
        k = int(r.n_pass) if cap else 0
        return info, idx[:k].copy(), usc[:k].copy()

    def stats(self):
        i = _SInfo()
        self.lib.msvg_server_stats(self._s, C.byref(i))
        return {k: getattr(i, k) for k, _ in _SInfo._fields_}

    def stop(self):
        if self._s:
            self.lib.msvg_server_stop(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.stop()
        except Exception:
            pass
