"""Serving: every query's pass set and scores are the same when it runs
alone, when it runs in a mix of concurrent queries, and through msvg_search
-- in each of the three serving modes: persistent (one query per search),
shared (one chunk per workgroup) and batched (the queries waiting for a slot
go to the GPU together in one multi-profile search)."""
import numpy as np
import pytest

import cudahmmsearch_amd as ca

pytestmark = pytest.mark.gpu


MODES = ["persistent", "shared", "batched"]


def _server(ctx, did, mode, slots=4):
    return ca.Server(ctx, did, slots=slots, shared=(mode == "shared"), batched=(mode == "batched"))


@pytest.mark.parametrize("mode", MODES)
def test_alone_equals_mixed(mode):
    ctx = ca.MSVContext(0)
    db = ca.gen_database(31, 20000, 300, 0.7, 20, 6000, nthreads=8)
    profs = [ca.sample_profile(500 + i, M) for i, M in enumerate([60, 180, 333, 512, 900, 1500])]
    hom = ca.core_emit(profs[2].hmm, 3, 200)
    db, _ = ca.plant(db, hom, 4, 0.01)
    did = ctx.add_database(db)
    direct = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        usc, st, ps = ctx.search([pid], did)
        direct.append((np.flatnonzero(ps[0]).astype(np.int32), usc[0][ps[0] == 1]))
        ctx.remove_profile(pid)
    srv = _server(ctx, did, mode)
    alone = []
    for p in profs:
        t = srv.submit(p)
        info, idx, usc = srv.wait(t)
        assert info["status"] == 0
        alone.append((idx, usc))
    # a burst: every profile three times, all queued at once
    tickets = [(j, srv.submit(profs[j])) for _ in range(3) for j in range(len(profs))]
    batches = []
    for j, t in tickets:
        info, idx, usc = srv.wait(t)
        assert info["status"] == 0
        batches.append(info["batch"])
        assert np.array_equal(idx, alone[j][0]), j
        assert np.array_equal(usc.view(np.uint32), alone[j][1].view(np.uint32)), j
    for j in range(len(profs)):
        assert np.array_equal(alone[j][0], direct[j][0])
        assert np.array_equal(alone[j][1].view(np.uint32), direct[j][1].view(np.uint32))
    st = srv.stats()
    assert st["completed"] == 4 * len(profs) and st["pending"] == 0
    if mode == "batched":
        assert max(batches) > 1, "the burst never shared a search"
        assert st["batches"] < st["admitted"]
    else:
        assert max(batches) == 1 and st["batches"] == st["admitted"]
    srv.stop()
    ctx.close()


@pytest.mark.parametrize("mode", MODES)
def test_same_new_profile_in_several_slots(mode):
    """A profile the server has not seen, submitted four times at once: the
    first admission builds its tables on its slot's stream and waits for
    them before other slots may run the same profile, so all four results
    equal msvg_search's."""
    ctx = ca.MSVContext(0)
    db = ca.gen_database(33, 30000, 350, 0.7, 20, 8000, nthreads=8)
    profs = [ca.sample_profile(700 + i, M) for i, M in enumerate([90, 420, 1300, 2400])]
    did = ctx.add_database(db)
    want = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        usc, st, ps = ctx.search([pid], did)
        want.append((np.flatnonzero(ps[0]).astype(np.int32), usc[0][ps[0] == 1]))
        ctx.remove_profile(pid)
    for j, p in enumerate(profs):
        srv = _server(ctx, did, mode)
        tickets = [srv.submit(p) for _ in range(4)]
        for t in tickets:
            info, idx, usc = srv.wait(t)
            assert info["status"] == 0
            assert np.array_equal(idx, want[j][0]), (j, p.M)
            assert np.array_equal(usc.view(np.uint32), want[j][1].view(np.uint32)), (j, p.M)
        srv.stop()
    ctx.close()


def test_batched_search_mixes_lane_groups_and_duplicates():
    """Batched serving with one slot: a burst of queries of every kind the
    batched search takes apart -- lane-group sizes 8 to 64, a column-segmented
    long model (M = 5000) and the same profile three times -- goes to the GPU
    in shared searches, and every result equals msvg_search's for that
    profile alone."""
    ctx = ca.MSVContext(0)
    db = ca.gen_database(35, 40000, 320, 0.7, 20, 9000, nthreads=8)
    profs = [ca.sample_profile(7100 + i, M) for i, M in enumerate([35, 150, 700, 2000, 5000])]
    hom = ca.core_emit(profs[1].hmm, 6, 300)
    db, _ = ca.plant(db, hom, 7, 0.01)
    did = ctx.add_database(db)
    want = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        usc, st, ps = ctx.search([pid], did)
        want.append((np.flatnonzero(ps[0]).astype(np.int32), usc[0][ps[0] == 1]))
        ctx.remove_profile(pid)
    order = [0, 1, 1, 2, 3, 4, 1, 0, 4, 2, 3, 1]
    srv = _server(ctx, did, "batched", slots=1)
    tickets = [(j, srv.submit(profs[j])) for j in order]
    sizes = []
    for j, t in tickets:
        info, idx, usc = srv.wait(t)
        assert info["status"] == 0
        sizes.append(info["batch"])
        assert np.array_equal(idx, want[j][0]), (j, profs[j].M)
        assert np.array_equal(usc.view(np.uint32), want[j][1].view(np.uint32)), (j, profs[j].M)
    st = srv.stats()
    srv.stop()
    ctx.close()
    print(f"batched serving: {st['admitted']} queries in {st['batches']} searches, sizes {sizes}", flush=True)
    assert max(sizes) >= 6 and st["batches"] <= 3


def test_profile_cache_stays_bounded():
    """More distinct query profiles than the server keeps: the cache stays
    at its limit (least recently used profiles are removed) and a profile
    that was removed and comes back scores as before."""
    ctx = ca.MSVContext(0)
    db = ca.gen_database(34, 2000, 200, 0.5, 20, 1000, nthreads=8)
    did = ctx.add_database(db)
    profs = [ca.sample_profile(9000 + i, 40 + (i % 7) * 10) for i in range(300)]
    srv = ca.Server(ctx, did, slots=4)
    first = srv.wait(srv.submit(profs[0]))
    for p in profs[1:]:
        srv.wait(srv.submit(p), want_hits=False)
    st = srv.stats()
    assert st["completed"] == 300
    assert st["cached_profiles"] == 256, st
    again = srv.wait(srv.submit(profs[0]))      # evicted long ago: added again
    assert np.array_equal(first[1], again[1]) and np.array_equal(first[2].view(np.uint32), again[2].view(np.uint32))
    assert srv.stats()["cached_profiles"] == 256
    srv.stop()
    ctx.close()
