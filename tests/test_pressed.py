"""The pressed-profile reader (cudahmmsearch_amd/csrc/pressed.cpp, format 3/f
of p7_oprofile_Write, src/impl_sse/io.c:89-180) on the reference's own
hmmpress output (tutorial/minifam.h3f/.h3p, tutorial/MADE1.hmm.h3f/.h3p):
every MSV and ViterbiFilter field, and the calibrated E-value parameters,
equal what the compiled reference builds from the same models' text files
(the golden vectors), and corrupt or mismatched files are refused."""
import gzip
import os
import shutil

import numpy as np
import pytest

import cudahmmsearch_amd as ca
from conftest import golden_profile, vit_golden_profile

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "hmmer")


def _unpack(name, dst):
    out = os.path.join(str(dst), name)
    with gzip.open(os.path.join(INPUTS, name + ".gz"), "rb") as src, open(out, "wb") as fh:
        shutil.copyfileobj(src, fh)
    return out


def _models(tmp_path):
    return (ca.read_pressed(_unpack("minifam.h3f", tmp_path), _unpack("minifam.h3p", tmp_path))
            + ca.read_pressed(_unpack("MADE1.hmm.h3f", tmp_path), _unpack("MADE1.hmm.h3p", tmp_path)))


def _f32(x):
    return np.float32(x).view(np.uint32)


def test_fields_equal_the_reference_conversion(tmp_path, golden, vit_golden):
    z, man = golden
    zv, manv = vit_golden
    mi = {m["name"]: i for i, m in enumerate(man["profiles"])}
    vi = {m["name"]: i for i, m in enumerate(manv["profiles"])}
    models = _models(tmp_path)
    assert [m[0] for m in models] == ["globins4", "fn3", "Pkinase", "MADE1"]
    for name, prof, vit in models:
        g = golden_profile(z, man, mi[name])
        assert np.array_equal(prof.rbv, g["rbv"]), name
        assert (prof.tbm, prof.tec, prof.base, prof.bias) == (g["tbm"], g["tec"], g["base"], g["bias"])
        assert _f32(prof.scale) == _f32(g["scale"]) and _f32(prof.mu) == _f32(g["mu"])
        assert _f32(prof.lam) == _f32(g["lambda"])
        v = vit_golden_profile(zv, manv, vi[name])
        assert np.array_equal(vit.rwv, v["rwv"]) and np.array_equal(vit.twv, v["twv"]), name
        assert (vit.e_loop, vit.e_move, vit.base_w, vit.ddbound_w) == (v["e_loop"], v["e_move"], v["base_w"],
                                                                       v["ddbound_w"])
        assert _f32(vit.scale_w) == _f32(v["scale_w"]) and _f32(vit.nj) == _f32(v["nj"])
        assert _f32(vit.vmu) == _f32(v["vmu"]) and _f32(vit.vlambda) == _f32(v["vlambda"])


def test_msv_part_alone(tmp_path):
    models = ca.read_pressed(_unpack("minifam.h3f", tmp_path))
    assert len(models) == 3 and all(v is None for _, _, v in models)


def test_refuses_bad_files(tmp_path):
    f = _unpack("minifam.h3f", tmp_path)
    p = _unpack("MADE1.hmm.h3p", tmp_path)
    with pytest.raises(ca.MSVError) as e:              # parts of different databases
        ca.read_pressed(f, p)
    assert e.value.code == 7
    data = open(f, "rb").read()
    cut = os.path.join(str(tmp_path), "cut.h3f")
    with open(cut, "wb") as fh:                          # truncated inside the first model
        fh.write(data[:3000])
    with pytest.raises(ca.MSVError) as e:
        ca.read_pressed(cut)
    assert e.value.code == 7
    bad = os.path.join(str(tmp_path), "bad.h3f")
    with open(bad, "wb") as fh:                          # an older format's magic (3/e)
        fh.write((0xb3e5e6f3).to_bytes(4, "little") + data[4:])
    with pytest.raises(ca.MSVError) as e:
        ca.read_pressed(bad)
    assert e.value.code == 7
    with pytest.raises(ca.MSVError) as e:
        ca.read_pressed(os.path.join(str(tmp_path), "missing.h3f"))
    assert e.value.code == 6                    # eslENOTFOUND
