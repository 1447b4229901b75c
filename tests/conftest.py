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
