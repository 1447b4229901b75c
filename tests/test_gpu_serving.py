"""Serving: every query's pass set and scores are the same when it runs
alone, when it runs in a mix of concurrent queries, and through msvg_search."""
import numpy as np
import pytest

import cudahmmsearch_amd as ca

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shared", [False, True])
def test_alone_equals_mixed(shared):
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
    srv = ca.Server(ctx, did, slots=4, shared=shared)
    alone = []
    for p in profs:
        t = srv.submit(p)
        info, idx, usc = srv.wait(t)
        assert info["status"] == 0
        alone.append((idx, usc))
    # a burst: every profile three times, all queued at once
    tickets = [(j, srv.submit(profs[j])) for _ in range(3) for j in range(len(profs))]
    for j, t in tickets:
        info, idx, usc = srv.wait(t)
        assert info["status"] == 0
        assert np.array_equal(idx, alone[j][0]), j
        assert np.array_equal(usc.view(np.uint32), alone[j][1].view(np.uint32)), j
    for j in range(len(profs)):
        assert np.array_equal(alone[j][0], direct[j][0])
        assert np.array_equal(alone[j][1].view(np.uint32), direct[j][1].view(np.uint32))
    st = srv.stats()
    assert st["completed"] == 4 * len(profs) and st["pending"] == 0
    srv.stop()
    ctx.close()


@pytest.mark.parametrize("shared", [False, True])
def test_same_new_profile_in_several_slots(shared):
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
        srv = ca.Server(ctx, did, slots=4, shared=shared)
        tickets = [srv.submit(p) for _ in range(4)]
        for t in tickets:
            info, idx, usc = srv.wait(t)
            assert info["status"] == 0
            assert np.array_equal(idx, want[j][0]), (j, p.M)
            assert np.array_equal(usc.view(np.uint32), want[j][1].view(np.uint32)), (j, p.M)
        srv.stop()
    ctx.close()


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
