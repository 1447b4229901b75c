"""Error paths on the GPU, forced from the host: a device allocation that
fails surfaces as a status code with a message, leaves the context usable,
and the next search is exact again.  Invalid arguments are refused before
any device work."""
import numpy as np
import pytest

import cudahmmsearch_amd as ca
from oracle import Oracle

pytestmark = pytest.mark.gpu


def test_allocation_failure_is_reported_and_recoverable():
    lib = ca.load_library()
    ctx = ca.MSVContext(0, tiles=None)
    prof = ca.sample_profile(41, 300)
    pid = ctx.add_profile(prof)
    prof.mu = ctx.calibrate(pid, prof.lam)
    db = ca.gen_database(42, 500, 300, 0.7, 20, 3000)
    did = ctx.add_database(db)
    lib.msvg_test_alloc_budget(0)                 # the next device allocation fails
    try:
        db2 = ca.gen_database(43, 300, 300, 0.7, 20, 3000)
        with pytest.raises(ca.MSVError) as e:
            ctx.add_database(db2)
        assert e.value.code == 100 and "out of memory" in str(e.value).lower()
    finally:
        lib.msvg_test_alloc_budget(-1)
    usc, st, ps = ctx.search([pid], did)           # the context still works
    ost, ousc = Oracle().msv_batch(prof.as_dict(), db.res, db.off, db.lens)
    assert np.array_equal(usc[0].view(np.uint32), ousc.view(np.uint32))
    ctx.close()


def test_invalid_arguments():
    ctx = ca.MSVContext(0, tiles=None)
    prof = ca.sample_profile(44, 50)
    bad = ca.ByteProfile(prof.rbv, prof.tbm, prof.tec, 190, 70, prof.scale, 0.0, 0.7)   # bias + base >= 255
    with pytest.raises(ca.MSVError) as e:
        ctx.add_profile(bad)
    assert e.value.code == 11
    pid = ctx.add_profile(prof)
    with pytest.raises(ca.MSVError) as e:
        ctx.search([pid], 12345)                    # unknown database
    assert e.value.code == 11
    with pytest.raises(ca.MSVError) as e:
        ctx.search_device([pid], 12345)             # the same, refused by the library itself
    assert e.value.code == 11 and "database" in str(e.value)
    good = ctx.add_database(ca.Database.from_list([np.array([0, 1, 2], np.uint8)]))
    with pytest.raises(ca.MSVError) as e:
        ctx.search([pid + 1000], good)              # unknown profile
    assert e.value.code == 11 and "profile" in str(e.value)
    # residue codes beyond the profile's alphabet
    x = np.array([0, 1, 40], np.uint8)
    did = ctx.add_database(ca.Database.from_list([x]))
    with pytest.raises(ca.MSVError) as e:
        ctx.search([pid], did)
    assert e.value.code == 11
    ctx.close()


def test_batched_all_residue_codes():
    """A database holding every code 0..28 (degenerate residues, gaps,
    nonresidue and missing-data symbols) under several profiles at once."""
    rs = np.random.default_rng(45)
    seqs = [rs.integers(0, 29, int(L)).astype(np.uint8) for L in rs.integers(1, 1500, 600)]
    db = ca.Database.from_list(seqs)
    profs = [ca.sample_profile(46 + i, M) for i, M in enumerate([33, 250, 900, 2200])]
    ctx = ca.MSVContext(0)
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    did = ctx.add_database(db)
    usc, st, ps = ctx.search(pids, did)
    O = Oracle()
    for j, p in enumerate(profs):
        ost, ousc = O.msv_batch(p.as_dict(), db.res, db.off, db.lens)
        assert np.array_equal(st[j].astype(np.int32), ost)
        assert np.array_equal(usc[j].view(np.uint32), ousc.view(np.uint32))
    ctx.close()
