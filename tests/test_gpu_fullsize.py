"""Full-size checks.  C1 (the reference's CPU case) against the compiled
reference itself where oracle/_ref is present; C2 at its full 2,000,000
targets through size-independent properties (plain == batched bit for bit,
a permuted copy gives permuted results, a random sample equals the oracle,
planted homologs pass)."""
import os

import numpy as np
import pytest

import cudahmmsearch_amd as ca
from oracle import REF_SO, Oracle, Reference

pytestmark = pytest.mark.gpu


def test_c1_against_compiled_reference():
    """C1: M=200 (seed 1) vs 20,000 log-normal targets (median 300, sigma
    0.6, clip 30-5000, seed 2) with 1% core emissions: every usc, status and
    pass bit equal to the reference's own p7_MSVFilter + p7_Pipeline rule."""
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libmsvref.so is not built")
    prof = ca.sample_profile(1, 200)
    db = ca.gen_database(2, 20000, 300, 0.6, 30, 5000)
    db, planted = ca.plant(db, ca.core_emit(prof.hmm, 3, 200), 2, 0.01)
    ctx = ca.MSVContext(0)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    usc, st, ps = ctx.search([pid], did)
    R = Reference()
    mat, ins, t = prof.hmm
    rp = R.sample(1, 200, calibrate=True)        # p7_hmm_Sample + p7_Calibrate in the reference
    ref = rp.byte_profile()
    assert np.array_equal(ref["rbv"], prof.rbv)
    assert np.float32(ref["mu"]) == np.float32(prof.mu) and np.float32(ref["lambda"]) == np.float32(prof.lam)
    rst, rusc = rp.msv_batch(db.res, db.off, db.lens, nthreads=16)
    assert np.array_equal(st[0].astype(np.int32), rst)
    assert np.array_equal(usc[0].view(np.uint32), rusc.view(np.uint32))
    rpass = np.array([1 if rst[i] == 16 else rp.passes(int(db.lens[i]), float(rusc[i])) for i in range(db.n)])
    assert np.array_equal(ps[0], rpass)
    print(f"C1: {int(ps.sum())} of {db.n} pass; {int(ps[0, planted].sum())} of {planted.size} planted pass")
    ctx.close()


def test_c2_full_size_properties():
    prof = ca.sample_profile(3, 400)
    db = ca.gen_database(4, 2_000_000, 350, 0.7, 20, 10000, nthreads=16)
    db, planted = ca.plant(db, ca.core_emit(prof.hmm, 5, 10000), 5, 0.005)
    ctx = ca.MSVContext(0)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    ub, sb, pb = ctx.search([pid], did, mode=ca.MODE_BATCHED)
    up, sp, pp = ctx.search([pid], did, mode=ca.MODE_PLAIN)
    assert np.array_equal(ub.view(np.uint32), up.view(np.uint32)) and np.array_equal(pb, pp)
    # a random sample against the oracle
    O = Oracle()
    idx = np.sort(np.random.default_rng(8).choice(db.n, 4000, replace=False))
    idx = np.union1d(idx, planted[:500])
    sub = ca.Database.from_list([db.seq(i) for i in idx])
    ost, ousc = O.msv_batch(prof.as_dict(), sub.res, sub.off, sub.lens, nthreads=16)
    assert np.array_equal(sb[0, idx].astype(np.int32), ost)
    assert np.array_equal(ub[0, idx].view(np.uint32), ousc.view(np.uint32))
    # planted homologs score high; background passes at about the F1 rate
    assert pb[0, planted].mean() > 0.9
    bg = np.setdiff1d(np.arange(db.n), planted)
    assert 0.005 < pb[0, bg].mean() < 0.05
    # a permuted database gives the permuted results
    perm = np.random.default_rng(9).permutation(db.n)
    lens = db.lens[perm]
    off = np.zeros(db.n, np.int64)
    off[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    res = np.empty_like(db.res)
    src_off = db.off[perm]
    for a in range(0, db.n, 65536):   # copy in blocks
        for i in range(a, min(db.n, a + 65536)):
            res[off[i]:off[i] + lens[i]] = db.res[src_off[i]:src_off[i] + lens[i]]
    did2 = ctx.add_database(ca.Database(res, off, lens))
    u2, s2, p2 = ctx.search([pid], did2)
    assert np.array_equal(ub[0, perm].view(np.uint32), u2[0].view(np.uint32))
    assert np.array_equal(pb[0, perm], p2[0])
    ctx.close()
