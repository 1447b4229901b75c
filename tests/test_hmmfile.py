"""The HMMER3 ASCII reader (cudahmmsearch_amd/csrc/hmmfile.cpp, the parser of
p7_hmmfile_Read, src/p7_hmmfile.c:1295-1630) and the alphabet-general
conversions (msvh_profile_convert_abc / msvh_vit_convert_abc) on the
reference's own tutorial models: every MSV and ViterbiFilter field, and the
E-value parameters, equal what the compiled reference builds from the same
text files (the golden vectors), for amino models and the DNA model MADE1;
a multi-model file equals the hmmpress'ed database of the same models; the
older 3/a..3/e layouts read the same; malformed files are refused."""
import gzip
import os

import numpy as np
import pytest

import cudahmmsearch_amd as ca
from conftest import golden_profile, vit_golden_profile

HERE = os.path.dirname(os.path.abspath(__file__))
INPUTS = os.path.join(HERE, "golden", "hmmer")
MODELS = ["globins4", "fn3", "Pkinase", "MADE1"]


def _text(name):
    with gzip.open(os.path.join(INPUTS, name + ".hmm.gz"), "rt") as fh:
        return fh.read()


def _write(tmp_path, name, text):
    p = os.path.join(str(tmp_path), name)
    with open(p, "w") as fh:
        fh.write(text)
    return p


def _f32(x):
    return np.float32(x).view(np.uint32)


def _same_profiles(a, b):
    (na, pa, va), (nb, pb, vb) = a, b
    assert na == nb
    assert np.array_equal(pa.rbv, pb.rbv) and (pa.tbm, pa.tec, pa.base, pa.bias) == (pb.tbm, pb.tec, pb.base, pb.bias)
    assert _f32(pa.scale) == _f32(pb.scale) and _f32(pa.mu) == _f32(pb.mu) and _f32(pa.lam) == _f32(pb.lam)
    assert np.array_equal(va.rwv, vb.rwv) and np.array_equal(va.twv, vb.twv)
    assert (va.e_loop, va.e_move, va.base_w, va.ddbound_w) == (vb.e_loop, vb.e_move, vb.base_w, vb.ddbound_w)
    assert _f32(va.vmu) == _f32(vb.vmu) and _f32(va.vlambda) == _f32(vb.vlambda)


def test_models_equal_the_reference_conversion(tmp_path, golden, vit_golden):
    z, man = golden
    zv, manv = vit_golden
    mi = {m["name"]: i for i, m in enumerate(man["profiles"])}
    vi = {m["name"]: i for i, m in enumerate(manv["profiles"])}
    for name in MODELS:
        models = ca.read_hmm(_write(tmp_path, name + ".hmm", _text(name)))
        assert [m[0] for m in models] == [name]
        _, prof, vit = models[0]
        g = golden_profile(z, man, mi[name])
        assert prof.rbv.shape == (g["Kp"], g["M"])
        assert np.array_equal(prof.rbv, g["rbv"]), name
        assert (prof.tbm, prof.tec, prof.base, prof.bias) == (g["tbm"], g["tec"], g["base"], g["bias"])
        assert prof.tjb_config == ca.tjb(100)
        assert _f32(prof.scale) == _f32(g["scale"]) and _f32(prof.mu) == _f32(g["mu"])
        assert _f32(prof.lam) == _f32(g["lambda"])
        v = vit_golden_profile(zv, manv, vi[name])
        assert np.array_equal(vit.rwv, v["rwv"]) and np.array_equal(vit.twv, v["twv"]), name
        assert (vit.e_loop, vit.e_move, vit.base_w, vit.ddbound_w) == (v["e_loop"], v["e_move"], v["base_w"],
                                                                       v["ddbound_w"])
        assert _f32(vit.scale_w) == _f32(v["scale_w"]) and _f32(vit.nj) == _f32(v["nj"])
        assert _f32(vit.vmu) == _f32(v["vmu"]) and _f32(vit.vlambda) == _f32(v["vlambda"])


def test_multi_model_file_equals_pressed_database(tmp_path):
    """tutorial/minifam holds globins4, fn3 and Pkinase in one HMMER3/e file;
    its hmmpress'ed .h3f/.h3p (the reference's own output) hold the same
    models.  The 3/f files of the same models, concatenated, are refused: a
    file keeps its first model's format tag (p7_hmmfile.c:1322-1330)."""
    with gzip.open(os.path.join(INPUTS, "minifam.gz"), "rt") as fh:
        models = ca.read_hmm(_write(tmp_path, "minifam", fh.read()))
    with pytest.raises(ca.MSVError) as e:
        ca.read_hmm(_write(tmp_path, "mixed", "".join(_text(n) for n in ["globins4", "fn3", "Pkinase"])))
    assert e.value.code == 7
    for ext in ("h3f", "h3p"):
        with gzip.open(os.path.join(INPUTS, "minifam." + ext + ".gz"), "rb") as src:
            with open(os.path.join(str(tmp_path), "minifam." + ext), "wb") as dst:
                dst.write(src.read())
    pressed = ca.read_pressed(os.path.join(str(tmp_path), "minifam.h3f"), os.path.join(str(tmp_path), "minifam.h3p"))
    assert len(models) == len(pressed) == 3
    for a, b in zip(models, pressed):
        _same_profiles(a, b)


def test_lambda_from_probabilities_matches_stats_line(tmp_path):
    """Without STATS lines, lambda is p7_Lambda of the model's probabilities
    (ln 2 + 1.44 / (M H), H the mean match relative entropy against the
    alphabet's background); hmmbuild wrote the same quantity, rounded to
    five decimals, on the STATS lines."""
    for name in MODELS:
        text = _text(name)
        stats = [ln for ln in text.splitlines() if ln.startswith("STATS")]
        lam_file = float(stats[0].split()[4])
        bare = "\n".join(ln for ln in text.splitlines() if not ln.startswith("STATS")) + "\n"
        (_, prof, vit), = ca.read_hmm(_write(tmp_path, name + ".bare", bare))
        assert prof.mu is None
        assert abs(prof.lam - lam_file) < 1e-4, name
        (_, full, _), = ca.read_hmm(_write(tmp_path, name + ".hmm", text))
        assert np.array_equal(prof.rbv, full.rbv)


def _older(text, fmt):
    """The same model in an older layout: 3/e drops the MM column and header,
    3/d..3/b also the CONS column, 3/b the MAXL line, 3/a writes STATS as
    LOCAL VLAMBDA / VMU / FTAU."""
    out = []
    in_body = False
    K = 0
    for ln in text.splitlines():
        tok = ln.split()
        if ln.startswith("HMMER3/f"):
            ln = ln.replace("HMMER3/f", "HMMER3/" + fmt)
        elif not in_body and tok and tok[0] == "MM":
            continue
        elif not in_body and tok and tok[0] == "MAXL" and fmt < "c":
            continue
        elif not in_body and tok and tok[0] == "STATS" and fmt == "a":
            if tok[2] == "MSV":
                continue
            if tok[2] == "VITERBI":
                out += [f"STATS LOCAL VLAMBDA {tok[4]}", f"STATS LOCAL VMU {tok[3]}"]
                continue
            ln = f"STATS LOCAL FTAU {tok[3]}"
        elif tok and tok[0] == "HMM":
            in_body = True
            K = len(tok) - 1                   # the alphabet's residues head the columns
        elif in_body and tok and tok[0].isdigit() and len(tok) == K + 6:
            rest = tok[K + 1:]                 # MAP CONS RF MM CS
            keep = [rest[0]] + ([rest[1]] if fmt >= "e" else []) + [rest[2], rest[4]]
            ln = "  ".join(tok[:K + 1] + keep)
        out.append(ln)
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("fmt", ["e", "d", "c", "b", "a"])
def test_older_formats_read_the_same(tmp_path, fmt):
    for name in ["fn3", "MADE1"]:
        text = _text(name)
        (n0, p0, v0), = ca.read_hmm(_write(tmp_path, name + ".f", text))
        old = _older(text, fmt)
        assert sum(len(ln.split()) == 4 + {"fn3": 20, "MADE1": 4}[name] + (fmt >= "e")
                   for ln in old.splitlines()) >= {"fn3": 86, "MADE1": 80}[name]   # every match line rewritten
        (n1, p1, v1), = ca.read_hmm(_write(tmp_path, name + "." + fmt, old))
        assert n1 == n0 and np.array_equal(p1.rbv, p0.rbv) and np.array_equal(v1.rwv, v0.rwv)
        assert np.array_equal(v1.twv, v0.twv)
        if fmt == "a":                         # 3/a has one mu for MSV and Viterbi: the VMU line's
            assert _f32(p1.mu) == _f32(v0.vmu) and _f32(v1.vmu) == _f32(v0.vmu)
            assert _f32(p1.lam) == _f32(v0.vlambda)
        else:
            assert _f32(p1.mu) == _f32(p0.mu) and _f32(v1.vmu) == _f32(v0.vmu)


def test_refuses_malformed_files(tmp_path):
    text = _text("fn3")
    lines = text.splitlines()
    body = next(i for i, ln in enumerate(lines) if ln.split()[:1] == ["1"])

    def code(t, name):
        with pytest.raises(ca.MSVError) as e:
            ca.read_hmm(_write(tmp_path, name, t))
        return e.value.code

    assert code("# not a profile\nNAME x\n", "notahmm") == 7                       # no HMMER3 tag
    assert code(text.replace("HMMER3/f", "HMMER2.0", 1), "h2") == 7                 # HMMER2 is not read
    with pytest.raises(ca.MSVError) as e:
        ca.read_hmm(os.path.join(str(tmp_path), "missing.hmm"))
    assert e.value.code == 6                                                       # eslENOTFOUND
    assert code("\n".join(lines[:body + 10]) + "\n", "cut") == 3                     # cut short: eslEOF, with a message
    short = lines[:]
    short[body] = " ".join(short[body].split()[:10])
    assert code("\n".join(short) + "\n", "short") == 7                             # too few fields on a match line
    renum = lines[:]
    renum[body] = renum[body].replace(" 1 ", " 2 ", 1)
    assert code("\n".join(renum) + "\n", "renum") == 7                             # node number out of order
    nostats = [ln for ln in lines if "STATS LOCAL FORWARD" not in ln]
    assert code("\n".join(nostats) + "\n", "twostats") == 7                        # two of the three STATS lines
    noend = [ln for ln in lines if ln.strip() != "//"] + ["xx"]
    assert code("\n".join(noend) + "\n", "noend") == 7                             # no closing //
    badyn = text.replace("CONS  yes", "CONS  maybe", 1)
    assert code(badyn, "badyn") == 7
    assert code(text + _text("MADE1"), "mixed") == 10                              # eslEINCOMPAT: amino then DNA


def test_comments_and_blank_lines_are_skipped(tmp_path):
    text = _text("globins4")
    noisy = "\n".join(("\n# a comment\n" + ln) if i % 7 == 3 else ln for i, ln in enumerate(text.splitlines())) + "\n"
    (n0, p0, v0), = ca.read_hmm(_write(tmp_path, "g.hmm", text))
    (n1, p1, v1), = ca.read_hmm(_write(tmp_path, "g.noisy", noisy))
    assert n0 == n1 and np.array_equal(p0.rbv, p1.rbv) and np.array_equal(v0.twv, v1.twv)


TS_MODELS = ["20aa", "3box", "Caudal_act", "LuxC", "M1", "PSE", "Patched", "RRM_1", "SMC_N", "XYPPX", "ecori"]


@pytest.mark.parametrize("name", TS_MODELS)
def test_testsuite_models_equal_compiled_reference(tmp_path, name):
    """The reference's testsuite models (HMMER3/e; amino and DNA; M = 1 to
    1142): every MSV and ViterbiFilter field from read_hmm equals what the
    compiled reference builds from the same file (p7_hmmfile_Read,
    p7_ProfileConfig at L = 100, p7_oprofile_Convert; oracle/_ref)."""
    from oracle import Reference, vprofile_from_fields
    try:
        R = Reference()
    except FileNotFoundError:
        pytest.skip("oracle/_ref is not built (make -C oracle ref)")
    with gzip.open(os.path.join(INPUTS, f"testsuite_{name}.hmm.gz"), "rt") as fh:
        path = _write(tmp_path, name + ".hmm", fh.read())
    (n, prof, vit), = ca.read_hmm(path)
    rp = R.from_file(path, 0)
    g = rp.byte_profile()
    assert prof.rbv.shape == g["rbv"].shape and np.array_equal(prof.rbv, g["rbv"]), name
    assert (prof.tbm, prof.tec, prof.base, prof.bias) == (g["tbm"], g["tec"], g["base"], g["bias"])
    assert _f32(prof.scale) == _f32(g["scale"]) and _f32(prof.mu) == _f32(g["mu"])
    assert _f32(prof.lam) == _f32(g["lambda"])
    v = vprofile_from_fields(rp.vit_fields())
    assert np.array_equal(vit.rwv, v["rwv"]) and np.array_equal(vit.twv, v["twv"]), name
    assert (vit.e_loop, vit.e_move, vit.base_w, vit.ddbound_w) == (v["e_loop"], v["e_move"], v["base_w"],
                                                                   v["ddbound_w"])
    assert _f32(vit.scale_w) == _f32(v["scale_w"]) and _f32(vit.nj) == _f32(v["nj"])
    assert _f32(vit.vmu) == _f32(v["vmu"]) and _f32(vit.vlambda) == _f32(v["vlambda"])
