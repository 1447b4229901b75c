"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

tests/sanitize_host.cpp links, with g++ -fsanitize=address,undefined, the
HMMER3 ASCII reader (csrc/hmmfile.cpp), the pressed-file reader
(csrc/pressed.cpp), the profile preparation (csrc/msv_host.cpp) and the CPU
oracle (oracle/msv_oracle.c, oracle/vit_oracle.c), and runs them on the
golden model files (tutorial and testsuite models, amino and DNA), on a
pressed database, on truncated and byte-corrupted copies of every model file,
and on sampled models with targets of edge lengths.  Any sanitizer finding
aborts its run; the driver's own consistency checks (one-target oracle
calls equal the batch calls, statuses are valid, readers end with a status)
make it exit 1.  The device runtime is not in this build: it needs a GPU.
The reference runs its own suite under valgrind (testsuite/testsuite.sqc).
"""
import glob
import gzip
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "hmmer")
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")


def _have_asan(tmp):
    src = os.path.join(tmp, "probe.c")
    with open(src, "w") as fh:
        fh.write("int main(void) { return 0; }\n")
    r = subprocess.run(["gcc", *SAN, src, "-o", os.path.join(tmp, "probe")], capture_output=True)
    return r.returncode == 0


@pytest.fixture(scope="module")
def binary(tmp_path_factory):
    tmp = str(tmp_path_factory.mktemp("san"))
    if not _have_asan(tmp):
        pytest.skip("this compiler has no address/undefined sanitizer runtime")
    objs = []
    for c in ("msv_oracle.c", "vit_oracle.c"):
        o = os.path.join(tmp, c + ".o")
        subprocess.run(["gcc", "-std=c99", "-pthread", *SAN, "-c", os.path.join(ROOT, "oracle", c), "-o", o],
                       check=True)
        objs.append(o)
    csrc = os.path.join(ROOT, "cudahmmsearch_amd", "csrc")
    out = os.path.join(tmp, "sanitize_host")
    subprocess.run(["g++", "-std=c++17", "-pthread", *SAN, "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests", "sanitize_host.cpp"),
                    os.path.join(csrc, "msv_host.cpp"), os.path.join(csrc, "hmmfile.cpp"),
                    os.path.join(csrc, "pressed.cpp"), *objs, "-o", out], check=True)
    return out, tmp


def _unpack(name, tmp):
    out = os.path.join(tmp, name)
    with gzip.open(os.path.join(GOLD, name + ".gz"), "rb") as src, open(out, "wb") as fh:
        shutil.copyfileobj(src, fh)
    return out


def _run(args):
    r = subprocess.run(args, capture_output=True, text=True, env=ENV, timeout=600)
    assert r.returncode == 0, f"{' '.join(os.path.basename(a) for a in args[:2])} failed:\n{r.stderr[-3000:]}"
    return r.stdout


def _models(tmp):
    names = sorted(os.path.basename(p)[:-3] for p in glob.glob(os.path.join(GOLD, "*.hmm.gz")))
    assert len(names) >= 14
    return [_unpack(n, tmp) for n in names]


def test_readers_and_oracle_on_golden_models(binary):
    exe, tmp = binary
    out = _run([exe, "hmm", *_models(tmp)])
    assert "models read" in out


def test_pressed_reader(binary):
    exe, tmp = binary
    for base in ("minifam", "MADE1.hmm"):
        out = _run([exe, "pressed", _unpack(base + ".h3f", tmp), _unpack(base + ".h3p", tmp)])
        assert "pressed:" in out


def test_damaged_model_files(binary):
    exe, tmp = binary
    for path in _models(tmp):
        _run([exe, "mangle", path, os.path.join(tmp, "mangled.hmm")])


def test_preparation_and_generators(binary):
    exe, _ = binary
    out = _run([exe, "prep"])
    assert "prep:" in out
