"""GPU parity of the ViterbiFilter stage (vit_kernels.hip through the C ABI).

Bar: bit-identical p7_ViterbiFilter scores and statuses (integer DP; the
float score is the reference's expression of the integer xC).  Checked
against the golden vectors written from the compiled reference, against the
reference itself (oracle/_ref) for model lengths that exercise every shape
(16-, 32- and 64-lane targets, column segments and 2, 4 and 8 waves per
target, M up to 16384) and both ways the rows treat the lazy-F test, and the
drop-in and calibration entry points.  Targets of one wave have
different lengths (some end, some overflow, while the others go on).
"""
import numpy as np
import pytest

import cudahmmsearch_amd as ca
from conftest import vit_golden_cases, vit_golden_profile
from oracle import Oracle

pytestmark = pytest.mark.gpu


def _placeholder_msv(M, Kp):
    """An MSV profile to hang the Viterbi fields on (only the Viterbi stage runs)."""
    return ca.ByteProfile(np.zeros((Kp, M), np.uint8), 0, 3, 190, 0, 1.0, mu=0.0, lam=0.7)


def _ctx_with(vd, Kp):
    ctx = ca.MSVContext(0, tiles=None)
    pid = ctx.add_profile(_placeholder_msv(vd["rwv"].shape[1], Kp))
    ctx.set_viterbi(pid, ca.ViterbiProfile.from_dict(vd))
    return ctx, pid


def _same(sc, st, ref_sc, ref_st):
    return (np.array_equal(np.asarray(st).astype(np.int32), np.asarray(ref_st).astype(np.int32))
            and np.array_equal(np.asarray(sc, np.float32).view(np.uint32), np.asarray(ref_sc, np.float32).view(np.uint32)))


def test_golden_vectors(vit_golden):
    z, man = vit_golden
    for i, meta in enumerate(man["profiles"]):
        v = vit_golden_profile(z, man, i)
        cases = vit_golden_cases(z, i)
        print(f"profile {meta['name']} M={meta['M']} cases={len(cases)}", flush=True)
        ctx, pid = _ctx_with(v, meta["Kp"])
        did = ctx.add_database(ca.Database.from_list([x for _, _, x in cases]))
        sc, st = ctx.vit_search(pid, did, np.arange(len(cases)))
        idx = [c for c, _, _ in cases]
        assert _same(sc, st, z["case_bits"][idx].view(np.float32), z["case_status"][idx]), meta["name"]
        ctx.close()


@pytest.mark.parametrize("M", [1, 7, 33, 128, 129, 200, 256, 257, 400, 480, 481, 512, 513, 640, 1024, 1025, 1536,
                               2048, 2049, 2304, 2432, 2433, 4096, 4097, 8192, 8193, 9000, 16384])
def test_every_wave_count_vs_reference(M):
    """Model lengths at the edges of each (lanes, waves, registers) shape
    against the compiled reference's p7_ViterbiFilter: four 16-lane targets
    per wave up to M = 512, two 32-lane targets up to 1024, one wave per
    target up to 2432 (1 to 16 registers per lane, 19 for one amino wave,
    whose MM, IM and DM transitions come from LDS beyond 16), then column
    segments of up to 1792 positions (2 to 10 of them here), with the
    core-emitted homologs' pairs scored again by the multi-wave kernel."""
    from oracle import Reference
    R = Reference()
    rp = R.sample(500 + M, M, calibrate=False)
    f = rp.vit_fields()
    vd = {"rwv": f["rwv_pad"][:, :M], "twv": f["twv_pad"][:, :M], "e_loop": f["e_loop"], "e_move": f["e_move"],
          "base_w": f["base_w"], "ddbound_w": f["ddbound_w"], "scale_w": f["scale_w"], "nj": f["nj"]}
    rng = np.random.default_rng(M)
    n = 60 if M <= 4097 else 24
    lens = rng.integers(1, 1200 if M <= 4097 else 500, n)
    seqs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in lens]
    emit, elens = rp.core_emit(7 + M, 4)
    o = 0
    for L in elens:
        seqs.append(emit[o:o + L])
        o += L
    seqs.append(rng.integers(0, 29, 300).astype(np.uint8))
    vp = ca.ViterbiProfile.from_dict(vd)
    for w in (8, 30):                                  # local alignments that raise J
        k0 = int(rng.integers(0, max(1, M - w)))
        seqs.append(_best_residues(vp, k0, min(M, k0 + w), rng))
    db = ca.Database.from_list(seqs)
    print(f"M={M} targets={db.n}", flush=True)
    rst, rsc = rp.vit_batch(db.res, db.off, db.lens, nthreads=8)
    # beyond one wave: column segments (default) and the multi-wave kernel
    for seg in ((True, False) if M > 2432 else (True,)):
        ctx, pid = _ctx_with(vd, 29)
        ctx.set_vit_segments(seg)
        ctx.set_viterbi(pid, vp)
        did = ctx.add_database(db)
        sc, st = ctx.vit_search(pid, did, np.arange(db.n))
        assert _same(sc, st, rsc, rst), seg
        ctx.close()


def _skewed_targets(seed, n, n_long):
    """n i.i.d. amino targets, log-normal lengths (median 350, sigma 0.7,
    clipped 20-3000), then n_long of 3,000-9,000 residues and one empty."""
    rng = np.random.default_rng(seed)
    lens = np.clip(np.exp(rng.normal(np.log(350), 0.7, n)), 20, 3000).astype(int)
    lens = np.concatenate([lens, rng.integers(3000, 9000, n_long), [0]])
    return [rng.integers(0, 20, int(L)).astype(np.uint8) for L in lens]


@pytest.mark.parametrize("vote", [-1, 0, 1])
def test_skewed_batch_vs_reference_every_vote_mode(vote):
    """More wave units than the launch has slots (several units per slot,
    dealt greedily by rows), targets of very different lengths, an empty one:
    one sampled profile per band shape (16, 32 and 64 lanes) and one in
    column segments, in one multi-profile call, against the compiled
    reference's p7_ViterbiFilter, with the rows voting on the lazy-F test
    (1), closing the delete paths on every row (0), or chosen by the
    profile's estimated rate (-1)."""
    from oracle import Reference
    R = Reference()
    seqs = _skewed_targets(11, 12000, 12)
    db = ca.Database.from_list(seqs)
    ctx = ca.MSVContext(0, tiles=None)
    ctx.set_vit_vote(vote)
    did = ctx.add_database(db)
    refs, pids, lists = [], [], []
    rng = np.random.default_rng(5)
    for k, M in enumerate([400, 900, 1700, 2300, 4100]):
        rp = R.sample(700 + k, M, calibrate=False)
        f = rp.vit_fields()
        vd = {"rwv": f["rwv_pad"][:, :M], "twv": f["twv_pad"][:, :M], "e_loop": f["e_loop"], "e_move": f["e_move"],
              "base_w": f["base_w"], "ddbound_w": f["ddbound_w"], "scale_w": f["scale_w"], "nj": f["nj"]}
        pid = ctx.add_profile(_placeholder_msv(M, 29))
        ctx.set_viterbi(pid, ca.ViterbiProfile.from_dict(vd))
        t = np.arange(db.n) if k == 0 else np.sort(rng.choice(db.n, 3000, replace=False))
        if k > 0:
            t = np.concatenate([t, np.arange(12000, db.n)])     # every long target and the empty one
        refs.append(rp); pids.append(pid); lists.append(t.astype(np.int64))
    print(f"vote={vote} targets={db.n} lists={[len(t) for t in lists]}", flush=True)
    many = ctx.vit_search_many(pids, did, lists)
    for rp, t, (sc, st) in zip(refs, lists, many):
        sub = ca.Database.from_list([seqs[i] for i in t])
        rst, rsc = rp.vit_batch(sub.res, sub.off, sub.lens, nthreads=8)
        assert _same(sc, st, rsc, rst)
    ctx.close()


def _best_residues(vp, k0, k1, rng, pad=30):
    """Random residues around the best-scoring residue of every position
    k0..k1-1 of the model: a local alignment that scores well above the
    E->J loop cost (J fires), and past the 16-bit range when long."""
    best = np.argmax(vp.rwv[:20, k0:k1], axis=0).astype(np.uint8)
    return np.concatenate([rng.integers(0, 20, pad).astype(np.uint8), best, rng.integers(0, 20, pad).astype(np.uint8)])


@pytest.mark.parametrize("M", [2433, 3841, 6000, 12000])
def test_column_segments_equal_multiwave(M):
    """Models beyond one wave's tables run as column segments (vit_seg): the
    same scores and statuses, bit for bit, as the multi-wave kernel on
    targets of every length 0..40, longer random ones, repeats,
    core-emitted samples, and runs of each position's best residue (short
    ones placed across a segment edge raise J, long ones overflow: those
    pairs are scored again by the multi-wave kernel, counted in n_slow)."""
    prof = ca.sample_profile(3000 + M, M)
    vp = prof.viterbi()
    rng = np.random.default_rng(M)
    seqs = [rng.integers(0, 20, L).astype(np.uint8) for L in range(41)]
    seqs += [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(41, 3000, 60)]
    seqs += ca.core_emit(prof.hmm, M + 7, 3)
    nseg = -(-M // 1792)                                # the runtime's segment width rule
    edge = 128 * -(-(-(-M // nseg)) // 128)
    hot = [_best_residues(vp, max(0, edge - w // 2), min(M, edge + w // 2), rng) for w in (6, 10, 20, 40)]
    hot += [_best_residues(vp, 0, M, rng), _best_residues(vp, M - 200, M, rng)]
    seqs += hot
    db = ca.Database.from_list(seqs)
    t = np.concatenate([np.arange(db.n), rng.integers(0, db.n, 20)]).astype(np.int64)
    print(f"segments M={M} targets={db.n} pairs={t.size}", flush=True)
    out = []
    for on in (True, False):
        ctx = ca.MSVContext(0, tiles=None)
        ctx.set_vit_segments(on)
        pid = ctx.add_profile(prof)
        ctx.set_viterbi(pid, prof.viterbi())
        did = ctx.add_database(db)
        out.append(ctx.vit_search(pid, did, t) + (ctx.stats()["n_slow"],))
        ctx.close()
    (sc1, st1, nd1), (sc0, st0, nd0) = out
    assert _same(sc1, st1, sc0, st0)
    # the long best-residue runs at least (short ones may stay below the
    # B->M entry cost), not the random targets
    assert nd0 == 0 and 2 <= nd1 < t.size // 4
    assert np.all(st1[db.n - 2:db.n] == 16)              # the long runs overflow


def test_many_profiles_equal_separate_searches():
    """msvg_vit_search_many over profiles of every kernel shape (16- and
    32-lane groups, one wave per target, column segments), one of them with
    an empty list, equals one msvg_vit_search per profile, bit for bit; the
    targets include each long model's best-residue runs, so several
    profiles hand pairs back to the multi-wave kernel in the same call."""
    Ms = [60, 300, 700, 1500, 2300, 120, 3000, 5200]
    ctx = ca.MSVContext(0, tiles=None)
    rng = np.random.default_rng(17)
    profs = [ca.sample_profile(60 + k, M) for k, M in enumerate(Ms)]
    seqs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(1, 2500, 300)]
    hot = []
    for prof in profs[6:]:
        vp = prof.viterbi()
        for w in (40, 400):
            k0 = int(rng.integers(0, prof.M - w))
            hot.append(len(seqs))
            seqs.append(_best_residues(vp, k0, k0 + w, rng))
    db = ca.Database.from_list(seqs)
    did = ctx.add_database(db)
    pids, lists = [], []
    for k, prof in enumerate(profs):
        pid = ctx.add_profile(prof)
        ctx.set_viterbi(pid, prof.viterbi())
        pids.append(pid)
        t = np.zeros(0, np.int64) if k == 5 else rng.choice(300, 120 + 10 * k, replace=False)
        if k >= 6:
            t = np.concatenate([t, hot])
        lists.append(np.asarray(t, np.int64))
    print(f"profiles M={Ms} targets={db.n}", flush=True)
    many = ctx.vit_search_many(pids, did, lists)
    assert ctx.stats()["n_slow"] >= 2
    for pid, t, (sc, st) in zip(pids, lists, many):
        sc1, st1 = ctx.vit_search(pid, did, t)
        assert _same(sc, st, sc1, st1)
    ctx.close()


def test_subset_order_and_repeats(vit_golden):
    """Any target list: a subset in scrambled order with repeats gives each
    target its own score."""
    z, man = vit_golden
    i = 1
    v = vit_golden_profile(z, man, i)
    cases = vit_golden_cases(z, i)
    ctx, pid = _ctx_with(v, man["profiles"][i]["Kp"])
    did = ctx.add_database(ca.Database.from_list([x for _, _, x in cases]))
    rng = np.random.default_rng(3)
    pick = rng.integers(0, len(cases), 3 * len(cases))
    sc, st = ctx.vit_search(pid, did, pick)
    idx = np.array([cases[k][0] for k in pick])
    assert _same(sc, st, z["case_bits"][idx].view(np.float32), z["case_status"][idx])
    sc0, st0 = ctx.vit_search(pid, did, np.zeros(0, np.int64))
    assert sc0.size == 0
    ctx.close()


def test_dropin_single_target(vit_golden):
    z, man = vit_golden
    O = Oracle()
    for i in (0, 4, 15):
        v = vit_golden_profile(z, man, i)
        ctx, pid = _ctx_with(v, man["profiles"][i]["Kp"])
        for c, _, x in vit_golden_cases(z, i)[:12]:
            st, sc = ctx.vit_filter(pid, x)
            ost, osc = O.vit(v, x)
            assert st == ost and np.float32(sc).view(np.uint32) == np.float32(osc).view(np.uint32)
        ctx.close()


def test_calibrated_vmu(vit_golden):
    """p7_ViterbiMu on the GPU gives the reference's calibrated mu (sampled,
    p7_Calibrate-d profiles of the golden set)."""
    z, man = vit_golden
    done = 0
    for i, meta in enumerate(man["profiles"]):
        if not meta["name"].startswith("sample-") or meta["M"] > 1000:
            continue
        v = vit_golden_profile(z, man, i)
        ctx, pid = _ctx_with(v, meta["Kp"])
        vmu = ctx.calibrate_vmu(pid, v["vlambda"])
        assert np.float32(vmu).view(np.uint32) == meta["vmu_bits"], meta["name"]
        ctx.close()
        done += 1
    assert done >= 6


def test_invalid_arguments(vit_golden):
    z, man = vit_golden
    v = vit_golden_profile(z, man, 0)
    ctx = ca.MSVContext(0, tiles=None)
    pid = ctx.add_profile(_placeholder_msv(v["rwv"].shape[1], 29))
    did = ctx.add_database(ca.Database.from_list([np.array([1, 2, 3], np.uint8)]))
    with pytest.raises(ca.MSVError) as e:
        ctx.vit_search(pid, did, [0])                       # no Viterbi fields yet
    assert e.value.code == 11
    other = ca.ViterbiProfile.from_dict(vit_golden_profile(z, man, 1))
    with pytest.raises(ca.MSVError):
        ctx.set_viterbi(pid, other)                          # M differs
    ctx.set_viterbi(pid, ca.ViterbiProfile.from_dict(v))
    with pytest.raises(ca.MSVError):
        ctx.vit_search(pid, did, [1])                        # out of range
    ctx.close()
