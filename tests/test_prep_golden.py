"""Pins the product's host-side preparation (include/msv_prep.h) to the
compiled reference: Easel RNG streams, p7_hmm_Sample -> byte profile, lambda,
xfIID targets, p7_CoreEmit samples, tjb_b and nullsc."""
import numpy as np

import cudahmmsearch_amd as ca
from conftest import golden_profile


def test_rng_streams(golden):
    z, _ = golden
    r = ca.Rng(7)
    assert np.array_equal(np.array([r.random() for _ in range(z["u_mt7"].size)]), z["u_mt7"])
    r = ca.Rng(42, fast=True)
    assert np.array_equal(np.array([r.random() for _ in range(z["u_fast42"].size)]), z["u_fast42"])


def test_sampled_profiles(golden):
    z, man = golden
    for pi, meta in enumerate(man["profiles"]):
        if meta["seed"] is None:
            continue
        p = golden_profile(z, man, pi)
        prof = ca.sample_profile(meta["seed"], meta["M"])
        assert np.array_equal(prof.rbv, p["rbv"]), meta["name"]
        assert (prof.tbm, prof.tec, prof.base, prof.bias) == (p["tbm"], p["tec"], p["base"], p["bias"])
        assert np.float32(prof.scale).view(np.uint32) == meta["scale_bits"]
        assert np.float32(prof.lam).view(np.uint32) == meta["lambda_bits"], meta["name"]


def test_iid_and_core_emit(golden):
    z, _ = golden
    assert np.array_equal(ca.iid(2, 16, 300), z["iid_s2"])
    hmm = ca.sample_hmm(1, 200)
    seqs = ca.core_emit(hmm, 5, 20)
    assert [len(s) for s in seqs] == list(z["core_s5_len"])
    assert np.array_equal(np.concatenate(seqs), z["core_s5_res"])


def test_length_tables(golden):
    z, _ = golden
    assert np.array_equal(np.array([ca.tjb(L) for L in range(z["tjb"].size)], np.uint8), z["tjb"])
    ns = np.array([np.float32(ca.nullsc(L)).view(np.uint32) for L in range(1, z["tjb"].size)], np.uint32)
    assert np.array_equal(ns, z["nullsc_bits"])


def test_tjb_closed_form():
    """tjb_b(L) = round(3/ln2 * ln((L+3)/3)).  For L <= 40000 (every target
    length of the workloads) no exact value lies within 3e-5 of a rounding
    tie, far above the float evaluation's error (< 1e-5), so the float
    expression and the exact one agree; beyond that the reference's float
    expression itself (glibc logf, used on both sides) is the definition."""
    L = np.arange(0, 40_001, dtype=np.float64)
    exact = 3.0 / np.log(2.0) * np.log((L + 3.0) / 3.0)
    frac = exact - np.floor(exact)
    assert np.min(np.abs(frac - 0.5)) > 3e-5
    rnd = np.floor(exact + 0.5).astype(np.int64)
    sample = np.arange(0, 40_001)
    got = np.array([ca.tjb(int(v)) for v in sample])
    assert np.array_equal(got, np.minimum(rnd[sample], 255))


def test_generators_deterministic():
    a = ca.gen_database(9, 500, 300, 0.6, 20, 5000, tail_frac=0.1, nthreads=3)
    b = ca.gen_database(9, 500, 300, 0.6, 20, 5000, tail_frac=0.1, nthreads=7)
    assert np.array_equal(a.lens, b.lens) and np.array_equal(a.res, b.res)
    assert a.res.max() < 20 and a.lens.min() >= 20
