"""Real models on the GPU: the reference's tutorial and testsuite HMMER3 text
models (amino and DNA, M = 1 to 1142), read by read_hmm and searched in both
modes, against the compiled reference's own p7_MSVFilter and
p7_ViterbiFilter (oracle/_ref: p7_hmmfile_Read, p7_ProfileConfig,
p7_oprofile_Convert, then the SSE filters per target) on the same targets:
random residues with degenerate codes, the model's own core emissions
(p7_CoreEmit through the reference) and empty targets.  Scores, statuses
and pass bits must be bit-identical."""
import gzip
import os

import numpy as np
import pytest

import cudahmmsearch_amd as ca

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "hmmer")
MODELS = ["tutorial_" + m for m in ("globins4", "fn3", "Pkinase", "MADE1")] + \
         ["testsuite_" + m for m in ("20aa", "3box", "Caudal_act", "LuxC", "M1", "PSE", "Patched", "RRM_1",
                                     "SMC_N", "XYPPX", "ecori")]


@pytest.fixture(scope="module")
def ctx():
    c = ca.MSVContext(0)
    yield c
    c.close()


def _model_file(tmp_path, name):
    kind, base = name.split("_", 1)
    src = os.path.join(INPUTS, (base if kind == "tutorial" else f"testsuite_{base}") + ".hmm.gz")
    path = os.path.join(str(tmp_path), base + ".hmm")
    with gzip.open(src, "rb") as fh, open(path, "wb") as out:
        out.write(fh.read())
    return path


@pytest.mark.gpu
@pytest.mark.parametrize("name", MODELS)
def test_real_model_filters_equal_reference(ctx, tmp_path, name):
    from oracle import Reference
    try:
        R = Reference()
    except FileNotFoundError:
        pytest.skip("oracle/_ref is not built (make -C oracle ref)")
    path = _model_file(tmp_path, name)
    (mname, prof, vit), = ca.read_hmm(path)
    rp = R.from_file(path, 0)
    Kp = prof.rbv.shape[0]
    K = 20 if Kp == 29 else 4
    rng = np.random.default_rng(sum(map(ord, mname)))
    seqs = []
    for L in rng.integers(1, 1500, 150):
        x = rng.integers(0, K, int(L)).astype(np.uint8)
        if L % 5 == 0:                                 # degenerate codes K+1 .. Kp-3
            x[rng.integers(0, L, 3)] = rng.integers(K + 1, Kp - 2, 3)
        seqs.append(x)
    emit, elens = rp.core_emit(7, 12)
    o = 0
    for L in elens:
        seqs.append(emit[o:o + L].copy())
        o += L
    seqs += [np.zeros(0, np.uint8), np.array([0], np.uint8)]
    db = ca.Database.from_list(seqs)
    print(f"{mname}: M={prof.M} Kp={Kp}, {db.n} targets", flush=True)
    rst, rusc = rp.msv_batch(db.res, db.off, db.lens, nthreads=16)
    rpass = np.array([1 if rst[i] == 16 else rp.passes(int(db.lens[i]), float(rusc[i])) for i in range(db.n)],
                     np.uint8)
    pid = ctx.add_profile(prof)
    ctx.set_viterbi(pid, vit)
    did = ctx.add_database(db)
    for mode in (ca.MODE_PLAIN, ca.MODE_BATCHED):
        usc, st, ps = ctx.search([pid], did, mode=mode)
        assert np.array_equal(st[0].astype(np.int32), rst), (mname, mode)
        assert np.array_equal(usc[0].view(np.uint32), rusc.view(np.uint32)), (mname, mode)
        assert np.array_equal(ps[0], rpass), (mname, mode)
    live = np.flatnonzero(db.lens > 0)
    vsc, vst = ctx.vit_search(pid, did, live)
    sub = ca.Database.from_list([db.seq(i) for i in live])
    rvst, rvsc = rp.vit_batch(sub.res, sub.off, sub.lens, nthreads=16)
    assert np.array_equal(np.asarray(vst).astype(np.int32), rvst), mname
    assert np.array_equal(np.asarray(vsc, np.float32).view(np.uint32), rvsc.view(np.uint32)), mname
    ctx.remove_database(did)
    ctx.remove_profile(pid)
