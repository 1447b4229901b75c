"""The drop-in check at the caller.

p7_Pipeline (src/p7_pipeline.c:646-...) calls p7_MSVFilter for every target
(p7_pipeline.c:671) and p7_ViterbiFilter for the MSV survivors (:701), and
derives each filter's pass from its score.  The serial search of
oracle/pipeline_driver.c runs that pipeline over the reference's own library
three times: oracle/_ref/pipeline_ref with the reference's SSE filters,
oracle/_ref/pipeline_mi355x with cudahmmsearch_amd/hmmer/filters_mi355x.c
(libmsvgpu on the GPU) in their place, nothing else changed, and
oracle/_ref/pipeline_block, which hands the whole database to
p7_mi355x_PipelineBlock (both filters batched on the GPU, then the rest of
the pipeline per sequence through p7_PipelineShort, as cudaHmmsearch's
block step does, src/GPUFunction.cu:470-490) -- once as one block in one
thread, once through p7_mi355x_ThreadLoop with 3 worker threads over blocks
of 16 sequences (hmmsearch's threaded loop, src/hmmsearch.c:1317-1425, its
workers' hits and pipelines merged).  Their output
files -- hit tables, domain alignments, tabular output, the pipeline's
filter pass counts and a line per target with both filters' statuses and
score bits -- must be byte-identical.  Inputs: the reference's tutorial
models and sequences (tests/golden/hmmer/, copied by make_inputs.py) and a
seeded synthetic protein database with planted fragments of real family
members, zero-length and long targets.
"""
import gzip
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFBIN = os.path.join(ROOT, "oracle", "_ref", "pipeline_ref")
GPUBIN = os.path.join(ROOT, "oracle", "_ref", "pipeline_mi355x")
BLOCKBIN = os.path.join(ROOT, "oracle", "_ref", "pipeline_block")
INPUTS = os.path.join(ROOT, "tests", "golden", "hmmer")
AMINO = "ACDEFGHIKLMNPQRSTVWY"
# Robinson & Robinson background frequencies, as p7_AminoFrequencies (hmmer.c:163-186) lists them.
BGF = np.array([0.0787945, 0.0151600, 0.0535222, 0.0668298, 0.0397062, 0.0695071, 0.0229198, 0.0590092,
                0.0594422, 0.0963728, 0.0237718, 0.0414386, 0.0482904, 0.0395639, 0.0540978, 0.0683364,
                0.0540687, 0.0673417, 0.0114135, 0.0304133])

needs_bins = pytest.mark.skipif(not all(os.path.exists(b) for b in (REFBIN, GPUBIN, BLOCKBIN)),
                                reason="pipeline binaries not built (make -C oracle ref dropin)")


def _unpack(name, dst):
    out = os.path.join(dst, name)
    with gzip.open(os.path.join(INPUTS, name + ".gz"), "rb") as src, open(out, "wb") as fh:
        shutil.copyfileobj(src, fh)
    return out


def _read_fasta(path):
    seqs, cur = [], None
    with open(path) as fh:
        for line in fh:
            if line.startswith(">"):
                cur = []
                seqs.append(cur)
            elif cur is not None:
                cur.append(line.strip())
    return ["".join(s) for s in seqs]


def _embl_sequence(path):
    out, on = [], False
    with open(path) as fh:
        for line in fh:
            if line.startswith("SQ"):
                on = True
            elif line.startswith("//"):
                break
            elif on:
                out.append("".join(ch for ch in line if ch.isalpha()).upper())
    return "".join(out)


def synthetic_fasta(path, n, seed, donors):
    """n targets: i.i.d. background residues (log-normal lengths, median 300),
    every 7th with a fragment of a real family member planted, a few
    zero-length targets, a few of 5000-9000 residues, and a sprinkling of
    the ambiguity codes X, B, Z."""
    rs = np.random.default_rng(seed)
    lens = np.clip(rs.lognormal(np.log(300), 0.6, n), 1, 4000).astype(int)
    with open(path, "w") as fh:
        for i in range(n):
            L = int(lens[i])
            if i % 97 == 5:
                L = 0
            elif i % 211 == 7:
                L = int(rs.integers(5000, 9000))
            s = "".join(np.array(list(AMINO))[rs.choice(20, L, p=BGF / BGF.sum())])
            if L and i % 7 == 3:
                d = donors[int(rs.integers(len(donors)))]
                w = int(rs.integers(40, min(300, len(d)) + 1))
                a = int(rs.integers(0, len(d) - w + 1))
                at = int(rs.integers(0, L + 1))
                s = s[:at] + d[a:a + w] + s[at:]
            if s and i % 13 == 0:
                s = list(s)
                for _ in range(3):
                    s[int(rs.integers(len(s)))] = "XBZ"[int(rs.integers(3))]
                s = "".join(s)
            fh.write(f">t{i} synthetic target {i}\n")
            for k in range(0, len(s), 60):
                fh.write(s[k:k + 60] + "\n")
            if not s:
                fh.write("\n")


def _run_all(tmp, hmm, seqs, extra=()):
    """Outputs of the SSE build, the per-call GPU build, the one-block build
    and the threaded block build (3 workers, blocks of 16 sequences)."""
    full = list(extra) + ["-", "bias"][len(extra):]
    runs = [(REFBIN, list(extra)), (GPUBIN, list(extra)), (BLOCKBIN, list(extra)),
            (BLOCKBIN, full + ["lines", "3", "16"])]
    outs = []
    for k, (b, args) in enumerate(runs):
        out = os.path.join(tmp, f"{os.path.basename(b)}_{k}.out")
        r = subprocess.run([b, hmm, seqs, out, *args], capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, f"{os.path.basename(b)} {args} failed: {r.stderr[-2000:]}"
        with open(out) as fh:
            outs.append(fh.read())
    return outs


def _check_same(ref, gpu):
    assert "Passed MSV filter:" in ref
    if ref != gpu:
        rl, gl = ref.splitlines(), gpu.splitlines()
        for i, (a, b) in enumerate(zip(rl, gl)):
            assert a == b, f"first difference at line {i + 1}:\n ref: {a}\n gpu: {b}"
        assert len(rl) == len(gl), "outputs differ in length"


@pytest.mark.gpu
@needs_bins
def test_tutorial_fn3_vs_7less(tmp_path):
    hmm = _unpack("fn3.hmm", tmp_path)
    seq = _unpack("7LESS_DROME", tmp_path)
    ref, *others = _run_all(str(tmp_path), hmm, seq)
    for o in others:
        _check_same(ref, o)
    # the shipped answer tutorial/fn3.out:110: the one target passes the MSV filter
    assert any(l.startswith("Passed MSV filter:") and l.split()[3] == "1" for l in ref.splitlines())


@pytest.mark.gpu
@needs_bins
def test_tutorial_globins(tmp_path):
    ref, *others = _run_all(str(tmp_path), _unpack("globins4.hmm", tmp_path), _unpack("globins45.fa", tmp_path))
    for o in others:
        _check_same(ref, o)


@pytest.mark.gpu
@needs_bins
def test_tutorial_dna(tmp_path):
    ref, *others = _run_all(str(tmp_path), _unpack("MADE1.hmm", tmp_path), _unpack("dna_target.fa", tmp_path))
    for o in others:
        _check_same(ref, o)


def _models(tmp, names):
    """One HMM file holding the named models in order (a multi-query search
    when there are several; the models in one file share a format)."""
    path = os.path.join(tmp, "_".join(n.split(".")[0] for n in names) + ".hmm")
    with open(path, "w") as fh:
        for m in names:
            with open(_unpack(m, tmp)) as src:
                fh.write(src.read())
    return path


def _donors(tmp):
    return _read_fasta(_unpack("globins45.fa", tmp)) + [_embl_sequence(_unpack("7LESS_DROME", tmp))]


@pytest.mark.gpu
@needs_bins
@pytest.mark.parametrize("models,extra", [
    (("fn3.hmm", "globins4.hmm"), ()),
    (("fn3.hmm", "globins4.hmm"), ("0.3", "nobias")),
    (("Pkinase.hmm",), ()),
], ids=["fn3+globins4", "fn3+globins4-F1=0.3-nobias", "Pkinase-3e"])
def test_synthetic_targets(tmp_path, models, extra):
    tmp = str(tmp_path)
    seqs = os.path.join(tmp, "targets.fa")
    synthetic_fasta(seqs, 700, 2024, _donors(tmp))
    ref, *others = _run_all(tmp, _models(tmp, models), seqs, extra)
    for o in others:
        _check_same(ref, o)
    assert ref.count("\nmsv ") == len(models) * 700
    passed = [int(l.split()[3]) for l in ref.splitlines() if l.startswith("Passed MSV filter:")]
    assert len(passed) == len(models) and all(p > 0 for p in passed)
    vit = [int(l.split()[3]) for l in ref.splitlines() if l.startswith("Passed Vit filter:")]
    assert len(vit) == len(models)
    if "globins4.hmm" in models:                  # planted globins: Viterbi survivors and overflows
        assert sum(vit) > 0 and ref.count(" vit 16 ") > 0 and ref.count(" vit 0 ") > 0


@pytest.mark.skipif(os.path.exists("/dev/kfd") or not os.path.exists(GPUBIN),
                    reason="needs a machine without a GPU and the built drop-in")
def test_dropin_fails_loudly_without_gpu(tmp_path):
    """No silent CPU fallback: without a device the drop-in filters raise
    through Easel's exception handler and the search stops."""
    for b, args in ((GPUBIN, []), (BLOCKBIN, []), (BLOCKBIN, ["-", "bias", "lines", "2", "4"])):
        if not os.path.exists(b):
            continue
        r = subprocess.run([b, _unpack("fn3.hmm", tmp_path), _unpack("HBB_HUMAN", tmp_path),
                            os.path.join(str(tmp_path), "o.txt"), *args], capture_output=True, text=True, timeout=120)
        assert r.returncode != 0
        assert "no MI355X device" in r.stderr


def fast_fasta(path, n, seed, donors):
    """n targets (log-normal lengths, median 350, sigma 0.7, clipped 20..10000)
    of i.i.d. background residues, every 7th with a fragment of a real family
    member planted; generated in bulk for large n."""
    rs = np.random.default_rng(seed)
    lens = np.clip(np.round(np.exp(np.log(350.0) + 0.7 * rs.standard_normal(n))), 20, 10000).astype(np.int64)
    cdf = np.cumsum(BGF / BGF.sum())
    letters = np.frombuffer(AMINO.encode(), np.uint8)
    total = int(lens.sum())
    res = np.empty(total, np.uint8)
    for a in range(0, total, 1 << 24):
        b = min(total, a + (1 << 24))
        res[a:b] = letters[np.minimum(np.searchsorted(cdf, rs.random(b - a)), 19)]
    off = np.concatenate([[0], np.cumsum(lens)])
    with open(path, "wb") as fh:
        for i in range(n):
            s = res[off[i]:off[i + 1]].tobytes()
            if i % 7 == 3:
                d = donors[int(rs.integers(len(donors)))].encode()
                w = int(rs.integers(40, min(300, len(d)) + 1))
                a = int(rs.integers(0, len(d) - w + 1))
                at = int(rs.integers(0, len(s) + 1))
                s = s[:at] + d[a:a + w] + s[at:]
            fh.write(b">t%d\n" % i)
            fh.write(b"\n".join(s[k:k + 60] for k in range(0, len(s), 60)) + b"\n")


def _seconds(stderr):
    return [float(l.split()[2]) for l in stderr.splitlines() if l.startswith("search_seconds")]


@pytest.mark.gpu
@needs_bins
def test_threaded_block_step_500k(tmp_path):
    """500,000 targets, each model in a fresh process: the SSE pipeline
    serial and as hmmsearch --cpu 4 and --cpu 16 (the driver's reader and
    worker threads), the block step as one block and through
    p7_mi355x_ThreadLoop (4 workers x blocks of 10,000, 8 x 5,000): every
    output byte-identical to the serial SSE one; the wall times (FASTA
    reading included) are printed."""
    tmp = str(tmp_path)
    seqs = os.path.join(tmp, "db500k.fa")
    fast_fasta(seqs, 500_000, 2025, _donors(tmp))
    runs = [("SSE serial", REFBIN, []), ("SSE 4 threads", REFBIN, ["4", "10000"]),
            ("SSE 16 threads", REFBIN, ["16", "10000"]), ("block step, one block", BLOCKBIN, []),
            ("block step 4 x 10,000", BLOCKBIN, ["4", "10000"]), ("block step 8 x 5,000", BLOCKBIN, ["8", "5000"])]
    for m in ("fn3.hmm", "globins4.hmm", "Pkinase.hmm"):
        hmm = _unpack(m, tmp)
        outs, secs = [], []
        for k, (_, b, args) in enumerate(runs):
            out = os.path.join(tmp, f"{m}_{k}.out")
            r = subprocess.run([b, hmm, seqs, out, "-", "bias", "nolines", *args], capture_output=True, text=True,
                               timeout=300)
            assert r.returncode == 0, f"{os.path.basename(b)} {args} failed: {r.stderr[-2000:]}"
            secs.append(_seconds(r.stderr)[0])
            with open(out) as fh:
                outs.append(fh.read())
        print(f"{m}: " + ", ".join(f"{name} {t:.3f} s" for (name, _, _), t in zip(runs, secs)), flush=True)
        assert "targets 500000" in outs[0]
        for o in outs[1:]:
            _check_same(outs[0], o)


@pytest.mark.skipif(not os.path.exists(GPUBIN), reason="needs the built drop-in")
@pytest.mark.parametrize("which", ["msv", "vit"])
def test_dropin_refuses_a_mismatched_length_model(tmp_path, which):
    """The library derives tjb_b and the N/C/J moves from the target length;
    an om configured for another length (hmmsim --x-no-lengthmodel elides
    the length model, hmmsim.c:946-952) is refused with eslEINVAL through
    Easel's exception handler, before any device work, instead of being
    scored differently from the SSE filters.  The SSE build scores it."""
    hmm, seq = _unpack("fn3.hmm", tmp_path), _unpack("HBB_HUMAN", tmp_path)
    out = os.path.join(str(tmp_path), "o.txt")
    r = subprocess.run([GPUBIN, hmm, seq, out, "-", "bias", f"badlength-{which}"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "length model is not that of L" in r.stderr, r.stderr[-500:]
    if os.path.exists(REFBIN):
        r = subprocess.run([REFBIN, hmm, seq, out, "-", "bias", f"badlength-{which}"], capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0 and "filter status" in r.stderr
