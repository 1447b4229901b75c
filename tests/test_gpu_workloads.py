"""The benchmark workloads at their configured sizes, against the compiled
reference (oracle/_ref: HMMER's own p7_MSVFilter built from its sources).

* c3: 64 sampled profiles (M uniform in [50, 2500]) x 1,000,000
  length-skewed targets.  Batched equals plain bit for bit on every pair;
  each profile's byte profile equals the reference's conversion of the same
  HMM; usc and status equal the reference's p7_MSVFilter on every target
  over 10,000 residues, 2,000 random targets and the profile's 200
  best-scoring targets (where the J-aware and overflow paths run); pass
  bits follow the reference's F1 rule (p7_pipeline.c:667-677) as restated
  by the oracle.
* c4: the 8 long profiles (M 6000..12000, column-segmented) x 500,000
  targets: batched equals plain at full size; a 20,000-target sample plus
  each profile's 50 best-scoring targets against the reference.
* c5: the serving loop over a reduced resident database (200,000 targets of
  the c5 generator): the 64 query profiles submitted at once run
  overlapped (a query is admitted while another is on the GPU), every
  served pass list equals the reference's pass set, and every result
  equals the same query served alone; with batched serving the 64 go to
  the GPU in a few multi-profile searches and every result is again the
  one served alone.

The reference's own unit test sweeps model and target lengths the same way
(impl_sse/msvfilter.c:1000-1037).
"""
import os

import numpy as np
import pytest

import cudahmmsearch_amd as ca
from cudahmmsearch_amd import workloads
from oracle import REF_SO, Oracle, Reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref/libmsvref.so is not built")]


def _calibrated(ctx, profs):
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    return pids


def _oracle_pass(O, st, usc, lens, mu, lam):
    return np.array([1 if st[i] == 16 else O.passes(usc[i], int(lens[i]), mu, lam) for i in range(st.size)],
                    np.uint8)


def _check_vs_reference(R, O, p, db, idx, ub, sb, pb, tag):
    """usc/status of profile p on targets idx against the reference, and the
    pass bits against the F1 rule; the byte profile against the reference's
    conversion of the same HMM."""
    mat, ins, t = p.hmm
    rp = R.from_probs(mat, t, 100, False)
    rb = rp.byte_profile()
    assert np.array_equal(rb["rbv"], p.rbv), f"{tag}: byte profile differs from the reference's"
    assert (rb["tbm"], rb["tec"], rb["base"], rb["bias"]) == (p.tbm, p.tec, p.base, p.bias), tag
    sub = ca.Database.from_list([db.seq(int(i)) for i in idx])
    rst, rusc = rp.msv_batch(sub.res, sub.off, sub.lens, nthreads=16)
    bad = np.flatnonzero((sb[idx].astype(np.int32) != rst) | (ub[idx].view(np.uint32) != rusc.view(np.uint32)))
    assert bad.size == 0, (f"{tag}: {bad.size} of {idx.size} targets differ from the reference, first target "
                           f"{idx[bad[0]]} (L={sub.lens[bad[0]]}): gpu {sb[idx[bad[0]]]}/{ub[idx[bad[0]]]!r}, "
                           f"ref {rst[bad[0]]}/{rusc[bad[0]]!r}")
    assert np.array_equal(pb[idx], _oracle_pass(O, rst, rusc, sub.lens, p.mu, p.lam)), f"{tag}: pass bits"
    return int((rst == 16).sum())


def test_c3_full_batch_vs_reference():
    profs, db, _ = workloads.make_workload("c3")
    print(f"c3: {len(profs)} profiles, {db.n} targets, {db.residues} residues", flush=True)
    ctx = ca.MSVContext(0)
    pids = _calibrated(ctx, profs)
    did = ctx.add_database(db)
    ub, sb, pb = ctx.search(pids, did, mode=ca.MODE_BATCHED)
    st = ctx.stats()
    assert st["n_batch"] >= 1 and st["n_fast_launches"] <= 4
    up, sp, pp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    assert np.array_equal(ub.view(np.uint32), up.view(np.uint32)), "batched usc differs from plain"
    assert np.array_equal(sb, sp) and np.array_equal(pb, pp), "batched status/pass differs from plain"
    del up, sp, pp
    # a second batched run (graph replay) is identical
    ub2, sb2, pb2 = ctx.search(pids, did, mode=ca.MODE_BATCHED)
    assert np.array_equal(ub.view(np.uint32), ub2.view(np.uint32)) and np.array_equal(pb, pb2)
    del ub2, sb2, pb2
    ctx.close()
    R, O = Reference(), Oracle()
    rs = np.random.default_rng(31)
    common = np.union1d(np.flatnonzero(db.lens > 10000), rs.choice(db.n, 2000, replace=False))
    print(f"c3: {np.count_nonzero(db.lens > 10000)} targets over 10,000 residues", flush=True)
    n_erange = 0
    for j, p in enumerate(profs):
        top = np.argsort(-ub[j], kind="stable")[:200]
        idx = np.union1d(common, top)
        n_erange += _check_vs_reference(R, O, p, db, idx, ub[j], sb[j], pb[j], f"c3 profile {j} (M={p.M})")
    print(f"c3: 64 profiles checked against the reference; {n_erange} overflow pairs among them; "
          f"{int(pb.sum())} passing pairs overall", flush=True)


def test_c4_long_profiles_vs_reference():
    profs, db, _ = workloads.make_workload("c4")
    print(f"c4: M {[p.M for p in profs]}, {db.n} targets, {db.residues} residues", flush=True)
    ctx = ca.MSVContext(0)
    pids = _calibrated(ctx, profs)
    did = ctx.add_database(db)
    for M in sorted({p.M for p in profs}):
        assert ctx.tile_of(M, did)[2] == 1, f"M={M} should take the column-segmented path"
    ub, sb, pb = ctx.search(pids, did, mode=ca.MODE_BATCHED)
    up, sp, pp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    assert np.array_equal(ub.view(np.uint32), up.view(np.uint32)) and np.array_equal(sb, sp)
    assert np.array_equal(pb, pp)
    ctx.close()
    R, O = Reference(), Oracle()
    common = np.sort(np.random.default_rng(32).choice(db.n, 20000, replace=False))
    for j, p in enumerate(profs):
        idx = np.union1d(common, np.argsort(-ub[j], kind="stable")[:50])
        _check_vs_reference(R, O, p, db, idx, ub[j], sb[j], pb[j], f"c4 profile {j} (M={p.M})")
    print("c4: 8 profiles checked against the reference", flush=True)


def test_c5_serving_vs_reference():
    db = workloads.serving_database(200_000)
    profs = workloads.serving_queries()
    ctx = ca.MSVContext(0)
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        ctx.remove_profile(pid)
    did = ctx.add_database(db)
    # every query alone, one at a time
    srv = ca.Server(ctx, did, slots=4)
    alone = []
    for p in profs:
        info, idx, usc = srv.wait(srv.submit(p))
        assert info["status"] == 0
        alone.append((idx, usc))
    srv.stop()
    # all 64 at once: they overlap on the device
    srv = ca.Server(ctx, did, slots=4)
    tickets = [srv.submit(p) for p in profs]
    mixed = [srv.wait(t) for t in tickets]
    srv.stop()
    # all 64 at once, batched serving: folded into shared searches
    srv = ca.Server(ctx, did, slots=2, batched=True)
    tickets = [srv.submit(p) for p in profs]
    batched = [srv.wait(t) for t in tickets]
    bst = srv.stats()
    srv.stop()
    ctx.close()
    print(f"c5 batched: 64 queries in {bst['batches']} searches", flush=True)
    assert bst["batches"] < 16
    for j, (info, idx, usc) in enumerate(batched):
        assert info["status"] == 0
        assert np.array_equal(idx, alone[j][0]) and np.array_equal(usc.view(np.uint32), alone[j][1].view(np.uint32)), \
            f"query {j}: batched serving result differs from the query served alone"
    spans = sorted((m[0]["t_admit"], m[0]["t_done"]) for m in mixed)
    overlapped = sum(1 for k in range(1, len(spans)) if spans[k][0] < spans[k - 1][1])
    print(f"c5: {overlapped} of 63 queries admitted while the one admitted before them was still running",
          flush=True)
    assert overlapped > 0, "no query was admitted while another was on the GPU"
    R, O = Reference(), Oracle()
    for j, (p, (info, idx, usc)) in enumerate(zip(profs, mixed)):
        assert info["status"] == 0
        a_idx, a_usc = alone[j]
        assert np.array_equal(idx, a_idx) and np.array_equal(usc.view(np.uint32), a_usc.view(np.uint32)), \
            f"query {j}: served result differs alone and in the mix"
        mat, ins, t = p.hmm
        rp = R.from_probs(mat, t, 100, False)
        rst, rusc = rp.msv_batch(db.res, db.off, db.lens, nthreads=16)
        rpass = np.flatnonzero(_oracle_pass(O, rst, rusc, db.lens, p.mu, p.lam))
        assert np.array_equal(np.sort(idx), rpass), f"query {j} (M={p.M}): pass list differs from the reference"
        assert np.array_equal(usc.view(np.uint32), rusc[idx].view(np.uint32)), f"query {j}: usc of the hits"
    print(f"c5: 64 served pass lists equal the reference's ({sum(m[1].size for m in mixed)} hits)", flush=True)
