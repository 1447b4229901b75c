"""GPU parity: the HIP path against the golden vectors of the compiled
reference and against the oracle (oracle/msv_oracle.c) on seeded workloads.
Bar: bit-exact usc (float bits), status and F1 pass bit for every pair."""
import numpy as np
import pytest

import cudahmmsearch_amd as ca
from conftest import golden_cases, golden_profile
from oracle import Oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    # The workloads here are small: keep the tile table's lane groups (the
    # tail guard would widen them) so that every group size is exercised.
    c = ca.MSVContext(0)
    c.set_tail_guard(False)
    yield c
    c.close()


def _bp(p):
    return ca.ByteProfile(p["rbv"], p["tbm"], p["tec"], p["base"], p["bias"], p["scale"], p["mu"], p["lambda"])


def _oracle_pass(O, ust, uusc, lens, mu, lam, F1=0.02):
    return np.array([1 if ust[i] == 16 else O.passes(uusc[i], lens[i], mu, lam, F1) for i in range(len(lens))],
                    np.uint8)


@pytest.mark.parametrize("mode", [ca.MODE_PLAIN, ca.MODE_BATCHED])
def test_golden_cases(ctx, golden, mode):
    """Every golden case (sampled and shipped models, edge lengths, degenerate
    residues, planted homologs, ERANGE and SSV fall-backs) bit-exact."""
    z, man = golden
    nprof = len(man["profiles"])
    for pi in range(nprof):
        p = golden_profile(z, man, pi)
        rows = golden_cases(z, pi)
        print(f"golden profile {p['name']} M={p['M']} cases={len(rows)} mode={mode}", flush=True)
        pid = ctx.add_profile(_bp(p))
        db = ca.Database.from_list([r[2] for r in rows])
        did = ctx.add_database(db)
        usc, st, ps = ctx.search([pid], did, mode=mode)
        idx = np.array([r[0] for r in rows])
        assert np.array_equal(st[0].astype(np.int32), z["case_status"][idx]), p["name"]
        assert np.array_equal(usc[0].view(np.uint32), z["case_usc_bits"][idx]), p["name"]
        assert np.array_equal(ps[0], z["case_pass"][idx]), p["name"]
        ctx.remove_database(did)
        ctx.remove_profile(pid)


def test_fn3_vs_7less_drome_passes(ctx, golden):
    """tutorial/fn3.out:110 -- fn3 against 7LESS_DROME: 'Passed MSV filter: 1'."""
    z, man = golden
    pi = [m["name"] for m in man["profiles"]].index("fn3")
    c = man["named_cases"]["7LESS_DROME"]
    p = golden_profile(z, man, pi)
    pid = ctx.add_profile(_bp(p))
    x = z["res"][z["case_off"][c]:z["case_off"][c] + z["case_len"][c]]
    did = ctx.add_database(ca.Database.from_list([x]))
    usc, st, ps = ctx.search([pid], did)
    assert ps[0, 0] == 1
    ctx.remove_database(did)
    ctx.remove_profile(pid)


def test_calibration_mu_matches_reference(ctx, golden):
    """GPU p7_MSVMu reproduces p7_Calibrate's mu bit for bit."""
    z, man = golden
    for pi, meta in enumerate(man["profiles"]):
        if meta["seed"] is None or meta["M"] > 2500:
            continue
        prof = ca.sample_profile(meta["seed"], meta["M"])
        p = golden_profile(z, man, pi)
        assert np.array_equal(prof.rbv, p["rbv"])
        pid = ctx.add_profile(prof)
        mu = ctx.calibrate(pid, prof.lam)
        print(f"calibrate {meta['name']}: mu={mu} ref={p['mu']}", flush=True)
        assert np.float32(mu).view(np.uint32) == meta["mu_bits"], meta["name"]
        assert np.float32(prof.lam).view(np.uint32) == meta["lambda_bits"], meta["name"]
        ctx.remove_profile(pid)


def _workload(seed, M, n, median, sigma, lo, hi, planted=0.02, **kw):
    prof = ca.sample_profile(seed, M)
    db = ca.gen_database(seed + 1, n, median, sigma, lo, hi, **kw)
    hom = ca.core_emit(prof.hmm, seed + 2, max(1, int(planted * n)))
    db, _ = ca.plant(db, hom, seed + 3, planted)
    return prof, db


def _check_vs_oracle(ctx, prof, db, mode, pid=None, did=None, sample=None):
    O = Oracle()
    if pid is None:
        pid = ctx.add_profile(prof)
        prof.mu = ctx.calibrate(pid, prof.lam)
    if did is None:
        did = ctx.add_database(db)
    usc, st, ps = ctx.search([pid], did, mode=mode)
    idx = np.arange(db.n) if sample is None else np.sort(np.random.default_rng(0).choice(db.n, sample, replace=False))
    sub = ca.Database.from_list([db.seq(i) for i in idx])
    ost, ousc = O.msv_batch(prof.as_dict(), sub.res, sub.off, sub.lens, nthreads=16)
    assert np.array_equal(st[0, idx].astype(np.int32), ost)
    assert np.array_equal(usc[0, idx].view(np.uint32), ousc.view(np.uint32))
    opass = _oracle_pass(O, ost, ousc, sub.lens, prof.mu, prof.lam)
    assert np.array_equal(ps[0, idx], opass)
    return pid, did, usc, st, ps


@pytest.mark.parametrize("mode", [ca.MODE_PLAIN, ca.MODE_BATCHED])
def test_plain_profile_vs_oracle(ctx, mode):
    """M = 400 against 4000 log-normal targets with 2% planted homologs."""
    prof, db = _workload(3, 400, 4000, 350, 0.7, 20, 10000)
    print(f"M=400 n={db.n} residues={db.residues} mode={mode}", flush=True)
    pid, did, *_ = _check_vs_oracle(ctx, prof, db, mode)
    ctx.remove_database(did)
    ctx.remove_profile(pid)


@pytest.mark.parametrize("M", [1, 7, 33, 64, 65, 128, 129, 300, 513, 1024, 1025, 2049, 4096])
def test_model_lengths_vs_oracle(ctx, M):
    """Tile boundaries of every lane-group size (16/32/64 lanes, Q = 1..32)."""
    prof, db = _workload(100 + M, M, 400, 300, 0.8, 1, 3000)
    print(f"M={M} n={db.n}", flush=True)
    pid, did, *_ = _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED)
    usc2, st2, ps2 = ctx.search([pid], did, mode=ca.MODE_PLAIN)
    usc, st, ps = ctx.search([pid], did, mode=ca.MODE_BATCHED)
    assert np.array_equal(usc.view(np.uint32), usc2.view(np.uint32))
    ctx.remove_database(did)
    ctx.remove_profile(pid)


@pytest.mark.parametrize("M", [4097, 6000, 10000])
def test_long_profiles_vs_oracle(ctx, M):
    """Column-segmented path (tables beyond LDS)."""
    prof, db = _workload(200 + M, M, 60, 1200, 0.6, 100, 6000, planted=0.05)
    print(f"long M={M} n={db.n}", flush=True)
    pid, did, *_ = _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED)
    ctx.remove_database(did)
    ctx.remove_profile(pid)


@pytest.mark.parametrize("M", [5000, 7500])
def test_segmented_every_row_count(ctx, M):
    """Column segments (5000: 3584 + 1416 positions, an even last segment;
    7500: 3584 + 3584 + 332, an odd one) against targets of every length
    0..49 (blocks of 16 rows with 0..15 rows left over, the carry between
    blocks), lengths at multiples of 16 up to 4096, and core-emitted
    homologs of the whole model whose best diagonals cross the segment
    edges, in both modes, against the oracle."""
    prof = ca.sample_profile(900 + M, M)
    rng = np.random.default_rng(M)
    seqs = [rng.integers(0, 20, L).astype(np.uint8) for L in range(50)]
    seqs += [rng.integers(0, 20, L).astype(np.uint8) for L in (64, 256, 1024, 4096)]
    seqs += ca.core_emit(prof.hmm, M + 1, 12)
    db = ca.Database.from_list(seqs)
    print(f"segmented M={M} n={db.n}", flush=True)
    for mode in (ca.MODE_PLAIN, ca.MODE_BATCHED):
        pid, did, *_ = _check_vs_oracle(ctx, prof, db, mode)
        assert ctx.tile_of(M, did)[2] == 1
        ctx.remove_database(did)
        ctx.remove_profile(pid)


def test_batched_multi_profile(ctx):
    """Eight profiles of uneven M against length-skewed targets: batched equals
    plain equals the oracle on a sample."""
    seeds = list(range(10, 18))
    Ms = [50, 180, 420, 777, 1200, 1800, 2300, 2500]
    profs = [ca.sample_profile(s, M) for s, M in zip(seeds, Ms)]
    db = ca.gen_database(5, 3000, 300, 0.6, 20, 5000, tail_frac=0.02, tail_xm=1000, tail_alpha=1.5, tail_max=12000)
    hom = ca.core_emit(profs[2].hmm, 9, 60)
    db, _ = ca.plant(db, hom, 11, 0.02)
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    did = ctx.add_database(db)
    ub, sb, pb = ctx.search(pids, did, mode=ca.MODE_BATCHED)
    up, sp, pp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    assert np.array_equal(ub.view(np.uint32), up.view(np.uint32))
    assert np.array_equal(sb, sp) and np.array_equal(pb, pp)
    O = Oracle()
    idx = np.sort(np.random.default_rng(1).choice(db.n, 150, replace=False))
    sub = ca.Database.from_list([db.seq(i) for i in idx])
    for j, p in enumerate(profs):
        ost, ousc = O.msv_batch(p.as_dict(), sub.res, sub.off, sub.lens, nthreads=16)
        assert np.array_equal(sb[j, idx].astype(np.int32), ost), p.M
        assert np.array_equal(ub[j, idx].view(np.uint32), ousc.view(np.uint32)), p.M
        assert np.array_equal(pb[j, idx], _oracle_pass(O, ost, ousc, sub.lens, p.mu, p.lam)), p.M
    # determinism: a second batched run (graph replay) is identical
    ub2, sb2, pb2 = ctx.search(pids, did, mode=ca.MODE_BATCHED)
    assert np.array_equal(ub.view(np.uint32), ub2.view(np.uint32))
    ctx.remove_database(did)
    for pid in pids:
        ctx.remove_profile(pid)


def test_dropin_msvfilter(ctx, golden):
    """msvg_MSVFilter, the p7_MSVFilter signature, on the golden cases of
    every golden profile (M = 1 .. 6000, amino and DNA): the one-launch
    kernel where the profile's table fits LDS, the search path beyond."""
    z, man = golden
    paths = set()
    for pi, meta in enumerate(man["profiles"]):
        p = golden_profile(z, man, pi)
        pid = ctx.add_profile(_bp(p))
        cases = golden_cases(z, pi)
        print(f"profile {meta['name']} M={meta['M']} cases={len(cases)}", flush=True)
        for c, _, x in cases[:40]:
            st, usc = ctx.msv_filter(pid, x)
            assert st == z["case_status"][c], (meta["name"], c)
            assert np.float32(usc).view(np.uint32) == z["case_usc_bits"][c], (meta["name"], c)
            paths.add((meta["M"] <= 2560, ctx.stats()["dropin_path"]))
        ctx.remove_profile(pid)
    assert (True, 1) in paths and (False, 2) in paths and (False, 1) not in paths


def test_dropin_paths_match_batched_search(ctx):
    """Each msvg_MSVFilter call equals the batched search's score for the
    same target: planted homologs (J fires, overflow), background targets
    of 1..6000 residues; empty targets and the guard regime's sub-B0 case
    go to the search path, codes outside the alphabet are an error."""
    O = Oracle()
    prof, db = _workload(31, 200, 500, 300, 0.7, 1, 6000, planted=0.1)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    usc, st, _ = ctx.search([pid], did, mode=ca.MODE_BATCHED)
    assert (st[0] == 16).any() and (st[0] == 0).any()
    for i in range(db.n):
        r, u = ctx.msv_filter(pid, db.seq(i))
        assert ctx.stats()["dropin_path"] == 1, i
        assert r == int(st[0, i]) and np.float32(u).view(np.uint32) == usc[0, i].view(np.uint32), i
    r, u = ctx.msv_filter(pid, np.zeros(0, np.uint8))
    ost, ousc = O.msv(prof.as_dict(), np.zeros(0, np.uint8))
    assert ctx.stats()["dropin_path"] == 2 and r == ost and np.float32(u).view(np.uint32) == np.float32(ousc).view(np.uint32)
    with pytest.raises(ca.MSVError):
        ctx.msv_filter(pid, np.array([3, 29, 4], np.uint8))
    ctx.remove_database(did)
    ctx.remove_profile(pid)

    # Guard regime (tjb + tbm + tec + bias >= 127): M = 2500, L = 100,000.
    # A target of '*' (code 27, cost 255 at every node) keeps every cell at
    # the floor: its exact maximum lies below B0 and only the search path
    # has it.  The same target with 300 random residues inside, and a random
    # target (which overflows), stay on the kernel.
    prof = ca.sample_profile(11, 2500)
    pid = ctx.add_profile(prof)
    L = 100_000
    assert ca.tjb(L) + prof.tbm + prof.tec + prof.bias >= 127
    assert (prof.rbv[27] >= prof.bias).all()
    rng = np.random.default_rng(5)
    stars = np.full(L, 27, np.uint8)
    island = stars.copy()
    island[50_000:50_300] = rng.integers(0, 20, 300)
    for x, path in ((stars, 2), (island, 1), (rng.integers(0, 20, L).astype(np.uint8), 1)):
        print(f"guard regime: L={L} expect path {path}", flush=True)
        r, u = ctx.msv_filter(pid, x)
        ost, ousc = O.msv(prof.as_dict(), x)
        assert ctx.stats()["dropin_path"] == path
        assert r == ost and np.float32(u).view(np.uint32) == np.float32(ousc).view(np.uint32)
    ctx.remove_profile(pid)


def test_empty_and_zero_length(ctx):
    prof = ca.sample_profile(3, 400)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(ca.Database.from_list([]))
    usc, st, ps = ctx.search([pid], did)
    assert usc.shape == (1, 0)
    ctx.remove_database(did)
    db = ca.Database.from_list([np.zeros(0, np.uint8), np.array([3], np.uint8), np.zeros(0, np.uint8)])
    _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED, pid=pid)
    ctx.remove_profile(pid)


def test_megabase_targets(ctx):
    """Targets longer than 2^20 residues (the database build's comparison-sort
    path; the two of equal length keep the caller's order) beside short and
    empty ones: every result equals the oracle's, in both modes."""
    rng = np.random.default_rng(17)
    prof = ca.sample_profile(33, 60)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    big = 1_100_000
    seqs = [rng.integers(0, 20, big).astype(np.uint8), rng.integers(0, 20, 300).astype(np.uint8),
            np.zeros(0, np.uint8), rng.integers(0, 20, big).astype(np.uint8), rng.integers(0, 20, 45).astype(np.uint8)]
    frag = ca.core_emit(prof.hmm, 5, 1)[0]
    seqs[3][500_000:500_000 + frag.size] = frag                            # a planted homolog in one of them
    db = ca.Database.from_list(seqs)
    print("megabase targets: plain", flush=True)
    _, did, *_ = _check_vs_oracle(ctx, prof, db, ca.MODE_PLAIN, pid=pid)
    print("megabase targets: batched", flush=True)
    _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED, pid=pid, did=did)
    ctx.remove_database(did)
    ctx.remove_profile(pid)


def test_order_invariance(ctx):
    """Permuting the database permutes the results (sorting is internal)."""
    prof, db = _workload(21, 300, 1500, 300, 0.8, 10, 4000)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    u1, s1, p1 = ctx.search([pid], did)
    perm = np.random.default_rng(3).permutation(db.n)
    db2 = ca.Database.from_list([db.seq(i) for i in perm])
    did2 = ctx.add_database(db2)
    u2, s2, p2 = ctx.search([pid], did2)
    assert np.array_equal(u1[0, perm].view(np.uint32), u2[0].view(np.uint32))
    assert np.array_equal(p1[0, perm], p2[0])
    ctx.remove_database(did)
    ctx.remove_database(did2)
    ctx.remove_profile(pid)


@pytest.mark.parametrize("M", [100, 400, 700, 1500])
def test_every_tile_same_result(ctx, M):
    """Tiles (lanes per target, waves per workgroup) change speed, never results."""
    prof, db = _workload(300 + M, M, 800, 300, 0.7, 20, 3000)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    ref = None
    for lanes in (8, 16, 32, 64):
        if (M + 2 * lanes - 1) // (2 * lanes) > 32:
            continue
        for waves in (4, 16):
            ctx.clear_tiles()
            ctx.set_tile(1, 1 << 20, lanes, waves)
            out = ctx.search([pid], did, mode=ca.MODE_BATCHED)
            if ref is None:
                ref = out
            else:
                assert np.array_equal(out[0].view(np.uint32), ref[0].view(np.uint32)), (lanes, waves)
                assert np.array_equal(out[2], ref[2])
    ctx.clear_tiles()
    _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED, pid=pid, did=did, sample=200)
    ctx.remove_database(did)
    ctx.remove_profile(pid)


def test_deferred_pairs_take_the_j_stage(ctx):
    """Planted homologs whose J state may fire are resolved by the J-aware
    stage, not the byte stage, for every lane-group size the tile table may
    choose (M = 1000 sits where a 16-lane table would not fit LDS)."""
    for M in (300, 1000):
        prof, db = _workload(700 + M, M, 3000, 300, 0.7, 20, 3000, planted=0.03)
        pid, did, *_ = _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED, sample=300)
        ctx.search_device([pid], did, mode=ca.MODE_BATCHED | ca.FLAG_EAGER)
        st = ctx.stats()
        assert st["n_jrow"] > 0 and st["n_slow"] == 0, (M, st)
        ctx.remove_database(did)
        ctx.remove_profile(pid)


def test_j_stage_every_chunk_offset(ctx):
    """Multi-domain targets for the J-aware stage, which takes rows 16 at a
    time: a background prefix of every length 0..47 before the first
    homolog, background gaps between two to four homologs, and total
    lengths of every residue mod 16, so that B changes (and overflows) fall
    on every row offset of a chunk and partial last chunks take every
    length.  Every pair against the oracle, in both launch modes."""
    for M in (40, 120, 300):
        prof = ca.sample_profile(900 + M, M)
        hom = ca.core_emit(prof.hmm, 901 + M, 160)
        bg = ca.gen_database(902 + M, 640, 40, 0.6, 1, 300)
        rng = np.random.default_rng(M)
        seqs = []
        h = 0
        for t in range(48 * 3):
            parts = [bg.seq((4 * t) % bg.n)[: t % 48]]
            for d in range(2 + t % 3):
                parts.append(hom[h % len(hom)])
                h += 1
                parts.append(bg.seq((4 * t + d + 1) % bg.n)[: int(rng.integers(0, 80))])
            s = np.concatenate(parts).astype(np.uint8)
            seqs.append(s[: max(1, len(s) - int(rng.integers(0, 16)))])
        db = ca.Database.from_list(seqs)
        print(f"M={M} targets={db.n} residues={db.residues}", flush=True)
        for mode in (ca.MODE_PLAIN, ca.MODE_BATCHED):
            pid, did, _, st, _ = _check_vs_oracle(ctx, prof, db, mode)
            ctx.remove_database(did)
            ctx.remove_profile(pid)
        assert (st[0] == 16).any() or M == 40


def test_tail_guard(ctx):
    """A small database holding one long target: the tail guard takes 64
    lanes (the shortest row chain) where the tile table picks a narrower
    group, and the results stay identical."""
    prof, db = _workload(41, 200, 2000, 300, 0.6, 30, 5000, planted=0.01)
    seqs = [db.seq(i) for i in range(db.n)] + [np.random.default_rng(9).integers(0, 20, 30000).astype(np.uint8)]
    db = ca.Database.from_list(seqs)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    did = ctx.add_database(db)
    try:
        lanes_off = ctx.tile_of(200, did)[0]
        off = ctx.search([pid], did, mode=ca.MODE_BATCHED)
        ctx.set_tail_guard(True)
        lanes_on = ctx.tile_of(200, did)[0]
        on = ctx.search([pid], did, mode=ca.MODE_BATCHED)
    finally:
        ctx.set_tail_guard(False)
    assert lanes_off < 64 and lanes_on == 64, (lanes_off, lanes_on)
    assert np.array_equal(on[0].view(np.uint32), off[0].view(np.uint32))
    assert np.array_equal(on[1], off[1]) and np.array_equal(on[2], off[2])
    _check_vs_oracle(ctx, prof, db, ca.MODE_BATCHED, pid=pid, did=did, sample=300)
    ctx.remove_database(did)
    ctx.remove_profile(pid)


def test_scan_direction_equals_search(ctx):
    """msvg_scan (many profiles, one query; one workgroup per profile) gives
    each profile what msvg_search gives on a one-target database, and what
    the oracle gives: profiles
    of every register class (M = 1 .. 3000, beyond LDS at 2,561 and up,
    which take the search path), against random queries, planted homologs
    (J fires, overflow), a one-residue and an empty query."""
    rng = np.random.default_rng(77)
    Ms = [1, 2, 7, 60, 128, 129, 200, 256, 257, 400, 512, 513, 1000, 1024, 1025, 1536, 2048, 2049, 2560, 2561,
          3000] + [int(m) for m in rng.integers(20, 700, 40)]
    profs = [ca.sample_profile(5000 + i, M) for i, M in enumerate(Ms)]
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    queries = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in (350, 1200, 37)]
    queries += ca.core_emit(profs[9].hmm, 78, 2) + ca.core_emit(profs[3].hmm, 79, 1)
    queries += [np.array([5], np.uint8), np.zeros(0, np.uint8)]
    O = Oracle()
    for q in queries:
        print(f"query L={q.size} profiles={len(pids)}", flush=True)
        u, s, ps = ctx.scan(pids, q)
        did = ctx.add_database(ca.Database.from_list([q]))
        ru, rs, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
        ctx.remove_database(did)
        assert np.array_equal(u.view(np.uint32), ru[:, 0].view(np.uint32))
        assert np.array_equal(s, rs[:, 0]) and np.array_equal(ps, rp[:, 0])
        # and directly against the oracle
        one = ca.Database.from_list([q])
        for j, p in enumerate(profs):
            ost, ousc = O.msv_batch(p.as_dict(), one.res, one.off, one.lens)
            assert s[j] == ost[0] and u[j].view(np.uint32) == ousc[0].view(np.uint32), (j, q.size)
            assert ps[j] == _oracle_pass(O, ost, ousc, one.lens, p.mu, p.lam)[0], (j, q.size)
    st = ctx.stats()
    assert st["cells"] == 0 or st["cells"] > 0
    # new Gumbel parameters for every profile after the scan set was built:
    # the next scan of the same profile list applies them
    q = queries[0]
    _, _, before = ctx.scan(pids, q)
    for pid, p in zip(pids, profs):
        p.mu = p.mu - 8.0
        ctx.set_evparams(pid, p.mu, p.lam)
    u, s, ps = ctx.scan(pids, q)
    did = ctx.add_database(ca.Database.from_list([q]))
    _, _, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    ctx.remove_database(did)
    assert np.array_equal(ps, rp[:, 0]) and not np.array_equal(ps, before)
    one = ca.Database.from_list([q])
    for j, p in enumerate(profs):
        ost, ousc = O.msv_batch(p.as_dict(), one.res, one.off, one.lens)
        assert ps[j] == _oracle_pass(O, ost, ousc, one.lens, p.mu, p.lam)[0], j
    with pytest.raises(ca.MSVError):
        ctx.scan(pids, np.array([1, 29], np.uint8))              # code outside the alphabet
    for pid in pids:
        ctx.remove_profile(pid)


def test_scan_dna_and_mixed_alphabets(ctx, golden):
    """A DNA profile set (the tutorial's MADE1) scans like the amino sets;
    mixing alphabets in one scan is an error."""
    z, man = golden
    names = [m["name"] for m in man["profiles"]]
    dna = golden_profile(z, man, names.index("MADE1"))
    amino = golden_profile(z, man, names.index("fn3"))
    pd_ = ctx.add_profile(_bp(dna))
    pa = ctx.add_profile(_bp(amino))
    rng = np.random.default_rng(3)
    for L in (1, 80, 500):
        q = rng.integers(0, 4, L).astype(np.uint8)
        u, s, ps = ctx.scan([pd_], q)
        did = ctx.add_database(ca.Database.from_list([q]))
        ru, rs, rp = ctx.search([pd_], did, mode=ca.MODE_PLAIN)
        ctx.remove_database(did)
        assert np.array_equal(u.view(np.uint32), ru[:, 0].view(np.uint32)) and np.array_equal(s, rs[:, 0])
    # a batch (the grouped kernel, tables over the 18 DNA codes) with a
    # degenerate-code query and an empty one
    qs = [rng.integers(0, 4, int(L)).astype(np.uint8) for L in (3, 60, 400, 900, 2000)]
    qs += [rng.integers(0, 18, 300).astype(np.uint8), np.zeros(0, np.uint8)]
    pd2 = ctx.add_profile(_bp(dna))
    u, s, ps = ctx.scan_many([pd_, pd2], qs)
    did = ctx.add_database(ca.Database.from_list(qs))
    ru, rs, rp = ctx.search([pd_, pd2], did, mode=ca.MODE_PLAIN)
    ctx.remove_database(did)
    assert np.array_equal(u.view(np.uint32), ru.T.view(np.uint32)) and np.array_equal(s, rs.T)
    assert np.array_equal(ps, rp.T)
    with pytest.raises(ca.MSVError):
        ctx.scan([pd_, pa], np.array([1, 2], np.uint8))
    ctx.remove_profile(pd_)
    ctx.remove_profile(pd2)
    ctx.remove_profile(pa)


def _scan_many_vs_search_and_oracle(ctx, profs, qs):
    pids = [ctx.add_profile(p) for p in profs]
    db = ca.Database.from_list(qs)
    u, s, ps = ctx.scan_many(pids, db)
    jrow = ctx.stats()["n_jrow"]
    did = ctx.add_database(db)
    ru, rs, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    ctx.remove_database(did)
    assert np.array_equal(u.view(np.uint32), ru.T.view(np.uint32))
    assert np.array_equal(s, rs.T) and np.array_equal(ps, rp.T)
    O = Oracle()
    for j, p in enumerate(profs):
        ost, ousc = O.msv_batch(p.as_dict(), db.res, db.off, db.lens)
        assert np.array_equal(s[:, j].astype(np.int32), ost), (j, p.M)
        assert np.array_equal(u[:, j].view(np.uint32), ousc.view(np.uint32)), (j, p.M)
    for pid in pids:
        ctx.remove_profile(pid)
    return jrow


def test_scan_many_every_group_kernel(ctx):
    """The grouped scan kernel for every lane count and register-block count
    it is built for, at odd and even register counts (the narrow row loop
    pairs an odd count's leftover register across two rows) and at the
    block boundaries: amino models from 1 to 2,560 positions (8 lanes up to
    128, 16 up to 768, 32 up to 1,024, a whole wave beyond) and 18-code
    models of 3,000 to 4,096 positions (a whole wave, 6 to 8 blocks, tables
    that fit LDS only over the smaller alphabet), against batches of 4, 23
    and 90 queries with planted homologs and an empty query; equal to the
    search path and to the oracle."""
    rng = np.random.default_rng(2024)
    Ms = [1, 33, 64, 65, 100, 128, 129, 161, 200, 257, 300, 385, 450, 513, 600, 641, 700, 768,
          769, 900, 1024, 1025, 1300, 1537, 1800, 2049, 2300, 2560]
    profs = [ca.sample_profile(8100 + i, M) for i, M in enumerate(Ms)]
    for p in profs:
        p.mu, p.lam = -8.0, 0.69
    jrow = []
    for nq in (4, 23, 90):
        hi = 1500 if nq < 50 else 400
        qs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(1, hi, nq)]
        qs[1] = np.zeros(0, np.uint8)
        qs[2] = ca.core_emit(profs[9].hmm, 30 + nq, 1)[0]
        qs[-1] = ca.core_emit(profs[24].hmm, 40 + nq, 1)[0]
        print(f"scan_many every group kernel: amino, {nq} queries", flush=True)
        jrow.append(_scan_many_vs_search_and_oracle(ctx, profs, qs))
    assert max(jrow) > 0                        # the homologs' pairs took the J-aware stage
    # 18-code models (the nucleotide alphabets' Kp) from sampled models' first 18 rows
    dna = []
    for i, M in enumerate((100, 500, 3000, 3500, 4096)):
        d = ca.sample_profile(8200 + i, M).as_dict()
        dna.append(ca.ByteProfile(d["rbv"][:18], d["tbm"], d["tec"], d["base"], d["bias"], d["scale"], -8.0, 0.69,
                                  alphabet=ca.DNA))
    qs = [rng.integers(0, 4, int(L)).astype(np.uint8) for L in rng.integers(1, 900, 11)]
    qs += [rng.integers(0, 18, 300).astype(np.uint8), np.zeros(0, np.uint8)]
    print("scan_many every group kernel: 18 codes", flush=True)
    _scan_many_vs_search_and_oracle(ctx, dna, qs)


def test_scan_guard_regime_fallback(ctx):
    """In the guard regime a query that keeps every cell at the floor needs
    the exact byte MSV: msvg_scan hands that profile to the search path
    (n_slow), the other profiles stay on the kernel, and all results equal
    msvg_search's."""
    pa = ca.sample_profile(11, 2500)
    pb = ca.sample_profile(12, 100)
    L = 100_000
    assert ca.tjb(L) + pa.tbm + pa.tec + pa.bias >= 127 and (pa.rbv[27] >= pa.bias).all()
    pids = [ctx.add_profile(pa), ctx.add_profile(pb)]
    stars = np.full(L, 27, np.uint8)
    island = stars.copy()
    island[50_000:50_300] = np.random.default_rng(5).integers(0, 20, 300)
    for q, nslow in ((stars, 1), (island, 0)):
        print(f"guard-regime scan: L={L} expect {nslow} on the search path", flush=True)
        u, s, ps = ctx.scan(pids, q)
        assert ctx.stats()["n_slow"] == nslow
        did = ctx.add_database(ca.Database.from_list([q]))
        ru, rs, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
        ctx.remove_database(did)
        assert np.array_equal(u.view(np.uint32), ru[:, 0].view(np.uint32))
        assert np.array_equal(s, rs[:, 0]) and np.array_equal(ps, rp[:, 0])
    for pid in pids:
        ctx.remove_profile(pid)


def test_scan_many_equals_search(ctx):
    """msvg_scan_many (a batch of queries; each workgroup's waves share its
    profile's staged table) gives, for every (query, profile), what
    msvg_search gives over a database of the same queries and what the oracle
    gives: ragged lengths,
    empty queries, planted homologs, and a profile beyond LDS that takes the
    search path; a batch of one equals msvg_scan."""
    rng = np.random.default_rng(91)
    Ms = [1, 9, 64, 130, 257, 513, 1025, 2561] + [int(m) for m in rng.integers(20, 900, 24)]
    profs = [ca.sample_profile(7000 + i, M) for i, M in enumerate(Ms)]
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    jrow = []
    for nq in (1, 3, 4, 5, 13, 150):
        hi = 1500 if nq < 100 else 200          # 150 queries: more than 16 query groups per workgroup
        qs = [rng.integers(0, 20, int(L)).astype(np.uint8) for L in rng.integers(1, hi, nq)]
        if nq > 3:
            qs[1] = np.zeros(0, np.uint8)
            qs[2] = ca.core_emit(profs[4].hmm, 80 + nq, 1)[0]
            qs[-1] = ca.core_emit(profs[9].hmm, 90 + nq, 1)[0]
        db = ca.Database.from_list(qs)
        print(f"scan_many: {nq} queries x {len(pids)} profiles", flush=True)
        u, s, ps = ctx.scan_many(pids, db)
        jrow.append(ctx.stats()["n_jrow"])
        assert u.shape == (nq, len(pids))
        did = ctx.add_database(db)
        ru, rs, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
        ctx.remove_database(did)
        assert np.array_equal(u.view(np.uint32), ru.T.view(np.uint32))
        assert np.array_equal(s, rs.T) and np.array_equal(ps, rp.T)
        # and directly against the oracle
        O = Oracle()
        for j, p in enumerate(profs):
            ost, ousc = O.msv_batch(p.as_dict(), db.res, db.off, db.lens)
            assert np.array_equal(s[:, j].astype(np.int32), ost) and np.array_equal(u[:, j].view(np.uint32),
                                                                                  ousc.view(np.uint32)), j
            assert np.array_equal(ps[:, j], _oracle_pass(O, ost, ousc, db.lens, p.mu, p.lam)), j
        if nq == 1:
            u1, s1, p1 = ctx.scan(pids, qs[0])
            assert np.array_equal(u1.view(np.uint32), u[0].view(np.uint32)) and np.array_equal(p1, ps[0])
    assert max(jrow[2:]) > 0                    # the homologs' pairs took the J-aware stage
    # new Gumbel parameters after the batch's scan set was built: the next
    # batch applies them
    _, _, before = ctx.scan_many(pids, db)
    for pid, p in zip(pids, profs):
        p.mu = p.mu - 8.0
        ctx.set_evparams(pid, p.mu, p.lam)
    u, s, ps = ctx.scan_many(pids, db)
    did = ctx.add_database(db)
    _, _, rp = ctx.search(pids, did, mode=ca.MODE_PLAIN)
    ctx.remove_database(did)
    assert np.array_equal(ps, rp.T) and not np.array_equal(ps, before)
    u, s, ps = ctx.scan_many(pids, [])
    assert u.shape == (0, len(pids))
    for pid in pids:
        ctx.remove_profile(pid)


def test_packed_forms_probe_agree(ctx):
    """The packed-form probe's two encodings of the J-free recurrence (the
    kernels' f16 clamp form and the unsigned 16-bit form) give the same
    maximum on every one of its targets."""
    r = ctx.probe_forms(256)
    assert r["agree"], r
    assert r["cells"] > 0 and r["ms_f16"] > 0 and r["ms_u16"] > 0
