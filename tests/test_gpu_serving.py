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
