"""The C-ABI library loads and exports every function include/*.h declares
(no GPU work is started here)."""
import ctypes
import os
import re

import cudahmmsearch_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b((?:msvg|msvh)_[a-z_0-9]+)\s*\(", txt)))


def test_exports():
    lib = ca.load_library()
    names = [n for h in sorted(os.listdir(os.path.join(ROOT, "include"))) if h.endswith(".h") for n in declared(h)]
    assert len(names) > 25 and "msvh_hmmfile_next" in names and "msvh_pressed_next" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_status_codes_match_easel():
    txt = open(os.path.join(ROOT, "include", "msv_gpu.h")).read()
    codes = dict(re.findall(r"#define (MSVG_E?[A-Z]+)\s+(\d+)", txt))
    assert codes["MSVG_OK"] == "0" and codes["MSVG_EINVAL"] == "11" and codes["MSVG_ERANGE"] == "16"
    assert codes["MSVG_EMEM"] == "5" and codes["MSVG_ENORESULT"] == "19"


def test_profile_validation_without_gpu():
    """Argument checks run before any device call: a NULL context is EINVAL."""
    lib = ca.load_library()
    rc = lib.msvg_profile_add(None, None, None)
    assert rc == 11
    assert ctypes.c_int(lib.msvg_db_size(None, 0)).value == -1
