"""hmmscan's direction on the GPU: real query sequences against the models
of an hmmpress'ed database (tutorial/minifam.h3f/.h3p read by
cudahmmsearch_amd.read_pressed), through the batched MSV search (all models
in one call), the scan entry point (all models against one query per call)
and the ViterbiFilter stage, against the oracle -- scores, statuses and pass
bits bit for bit."""
import gzip
import os
import shutil

import numpy as np
import pytest

import cudahmmsearch_amd as ca
from oracle import Oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "hmmer")
AMINO = "ACDEFGHIKLMNPQRSTVWY-BJZOUX*~"           # Easel's amino alphabet order (esl_alphabet.c:275)


def _text(name):
    with gzip.open(os.path.join(INPUTS, name + ".gz"), "rt") as fh:
        return fh.read()


def _unpack(name, dst):
    out = os.path.join(str(dst), name)
    with gzip.open(os.path.join(INPUTS, name + ".gz"), "rb") as src, open(out, "wb") as fh:
        shutil.copyfileobj(src, fh)
    return out


def _digitize(s):
    return np.array([AMINO.index(ch) for ch in s.upper() if ch.isalpha() or ch in "-*~"], np.uint8)


def _queries():
    seqs = []
    lines = _text("7LESS_DROME").splitlines()
    on, buf = False, []
    for line in lines:
        if line.startswith("SQ"):
            on = True
        elif line.startswith("//"):
            break
        elif on:
            buf.append("".join(ch for ch in line if ch.isalpha()))
    seqs.append("".join(buf))
    for block in _text("globins45.fa").split(">")[1:]:
        seqs.append("".join(block.splitlines()[1:]))
    seqs.append("".join(_text("HBB_HUMAN").splitlines()[1:]))
    return [_digitize(s) for s in seqs]


def test_scan_queries_against_pressed_models(tmp_path):
    models = ca.read_pressed(_unpack("minifam.h3f", tmp_path), _unpack("minifam.h3p", tmp_path))
    queries = _queries()
    db = ca.Database.from_list(queries)
    ctx = ca.MSVContext(0)
    pids = []
    for name, prof, vit in models:
        pid = ctx.add_profile(prof)
        ctx.set_viterbi(pid, vit)
        pids.append(pid)
    did = ctx.add_database(db)
    usc, st, ps = ctx.search(pids, did)
    O = Oracle()
    for j, (name, prof, vit) in enumerate(models):
        ost, ousc = O.msv_batch(prof.as_dict(), db.res, db.off, db.lens)
        assert np.array_equal(st[j].astype(np.int32), ost), name
        assert np.array_equal(usc[j].view(np.uint32), ousc.view(np.uint32)), name
        opass = [1 if ost[i] == 16 else O.passes(ousc[i], db.lens[i], prof.mu, prof.lam) for i in range(db.n)]
        assert np.array_equal(ps[j], np.array(opass, np.uint8)), name
        vsc, vst = ctx.vit_search(pids[j], did, np.arange(db.n))
        for i, x in enumerate(queries):
            s, sc = O.vit(vit.as_dict() | {"M": vit.M}, x)
            assert vst[i] == s and np.float32(vsc[i]).view(np.uint32) == np.float32(sc).view(np.uint32), (name, i)
    # the known answer of tutorial/fn3.out: 7LESS_DROME passes fn3's MSV filter
    assert ps[[m[0] for m in models].index("fn3")][0] == 1
    # msvg_scan, one query at a time over all models (hmmscan's loop), gives
    # the same scores, statuses and pass bits
    for i, x in enumerate(queries):
        su, ss, sp = ctx.scan(pids, x)
        assert np.array_equal(su.view(np.uint32), usc[:, i].view(np.uint32)), i
        assert np.array_equal(ss, st[:, i]) and np.array_equal(sp, ps[:, i]), i
    ctx.close()
