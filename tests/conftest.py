import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


@pytest.fixture(scope="session")
def golden():
    import json
    z = np.load(os.path.join(GOLDEN, "msv_golden.npz"))
    with open(os.path.join(GOLDEN, "manifest.json")) as fh:
        man = json.load(fh)
    return z, man


def golden_profile(z, man, i):
    """The i-th golden profile as a dict the oracle and the product accept."""
    meta = man["profiles"][i]
    M, Kp = meta["M"], meta["Kp"]
    off = int(z["prof_rbv_off"][i])
    rbv = z["prof_rbv"][off:off + M * Kp].reshape(Kp, M)
    f = lambda b: float(np.array([b], np.uint32).view(np.float32)[0])  # noqa: E731
    return {"rbv": rbv, "tbm": meta["tbm"], "tec": meta["tec"], "base": meta["base"], "bias": meta["bias"],
            "scale": f(meta["scale_bits"]), "mu": f(meta["mu_bits"]), "lambda": f(meta["lambda_bits"]),
            "name": meta["name"], "seed": meta["seed"], "M": M, "Kp": Kp}


def golden_cases(z, pi=None):
    """Rows (index, profile, seq) of the golden cases, optionally of one profile."""
    out = []
    for c in range(z["case_len"].size):
        p = int(z["case_prof"][c])
        if pi is not None and p != pi:
            continue
        o, L = int(z["case_off"][c]), int(z["case_len"][c])
        out.append((c, p, z["res"][o:o + L]))
    return out


@pytest.fixture(scope="session")
def vit_golden():
    import json
    z = np.load(os.path.join(GOLDEN, "vit_golden.npz"))
    with open(os.path.join(GOLDEN, "vit_manifest.json")) as fh:
        man = json.load(fh)
    return z, man


def _bits_to_float(b):
    return float(np.array([b], np.uint32).view(np.float32)[0])


def vit_golden_profile(z, man, i):
    """The i-th Viterbi golden profile as a dict (model order, k = 1..M)."""
    metas = man["profiles"]
    ro = sum(m["M"] * m["Kp"] for m in metas[:i])
    to = sum(m["M"] * 8 for m in metas[:i])
    m = metas[i]
    M, Kp = m["M"], m["Kp"]
    return {"M": M, "Kp": Kp, "rwv": z["rwv"][ro:ro + M * Kp].reshape(Kp, M),
            "twv": z["twv"][to:to + 8 * M].reshape(8, M), "e_loop": m["e_loop"], "e_move": m["e_move"],
            "base_w": m["base_w"], "ddbound_w": m["ddbound_w"], "scale_w": _bits_to_float(m["scale_w_bits"]),
            "nj": _bits_to_float(m["nj_bits"]), "vmu": _bits_to_float(m["vmu_bits"]),
            "vlambda": _bits_to_float(m["vlambda_bits"]), "name": m["name"], "seed": m["seed"]}


def vit_golden_cases(z, pi=None):
    out = []
    for c in range(z["case_len"].size):
        p = int(z["case_prof"][c])
        if pi is not None and p != pi:
            continue
        o, L = int(z["case_off"][c]), int(z["case_len"][c])
        out.append((c, p, z["residues"][o:o + L]))
    return out
