"""The ViterbiFilter planner's lazy-F estimate (msvh_vit_lazyf_rate, host
code, no GPU): how often p7_ViterbiFilter's test Dmax + ddbound_w > xB
(vitfilter.c:199) holds on a pseudo-random target.  It only chooses how the
kernel's rows treat the test (vote on it, or close the delete paths on every
row); scores do not depend on it (the GPU tests run both ways against the
reference).  Pinned here: sampled models (p7_hmm_Sample's Dirichlet
transitions make delete paths strong) are closed on every row, the
reference's tutorial models vote, and the estimate is a deterministic
fraction."""
import ctypes as C
import gzip
import os

import numpy as np

import cudahmmsearch_amd as ca

HMMER = os.path.join(os.path.dirname(__file__), "golden", "hmmer")


def _rate(v, rows=128):
    lib = ca.load_library()
    f = lib.msvh_vit_lazyf_rate
    f.restype = C.c_double
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    mv = lib.msvh_vit_move(C.c_float(v.scale_w), C.c_float(v.nj), 350)
    return f(v.rwv.ctypes.data, v.twv.ctypes.data, v.M, v.rwv.shape[0], v.e_loop, v.base_w, v.ddbound_w, mv, rows)


def test_sampled_models_close_every_row():
    for seed, M in [(3, 400), (11, 800), (12, 2500)]:
        r = _rate(ca.sample_profile(seed, M).viterbi())
        assert 0.9 <= r <= 1.0, (seed, M, r)


def test_tutorial_models_vote(tmp_path):
    for name in ["globins4", "fn3", "Pkinase", "MADE1"]:
        path = tmp_path / f"{name}.hmm"
        with gzip.open(os.path.join(HMMER, f"{name}.hmm.gz")) as fh:
            path.write_bytes(fh.read())
        (_, _, vp), = ca.read_hmm(str(path))
        r = _rate(vp)
        assert 0.0 <= r < 0.75, (name, r)


def test_estimate_is_a_deterministic_fraction():
    v = ca.sample_profile(5, 60).viterbi()
    a, b = _rate(v, 40), _rate(v, 40)
    assert a == b and 0.0 <= a <= 1.0 and abs(a * 40 - round(a * 40)) < 1e-9
    one = ca.sample_profile(6, 1).viterbi()
    assert 0.0 <= _rate(one, 16) <= 1.0
