"""Pins the oracle (oracle/msv_oracle.c) to the compiled reference's golden
vectors: every case's p7_MSVFilter status and usc bits, p7_SSVFilter status,
F1 pass bit, and the tjb_b / nullsc tables."""
import numpy as np

from conftest import golden_cases, golden_profile
from oracle import Oracle


def test_msv_cases_bit_exact(golden):
    z, man = golden
    O = Oracle()
    bad = []
    for pi in range(len(man["profiles"])):
        p = golden_profile(z, man, pi)
        for c, _, x in golden_cases(z, pi):
            st, usc = O.msv(p, x)
            sst, _ = O.ssv(p, x)
            ok = (st == z["case_status"][c] and np.float32(usc).view(np.uint32) == z["case_usc_bits"][c]
                  and sst == z["case_ssv"][c])
            if st == 0:
                ok = ok and O.passes(usc, len(x), p["mu"], p["lambda"]) == z["case_pass"][c]
            if not ok:
                bad.append((p["name"], c, len(x)))
    assert not bad, bad[:10]


def test_golden_covers_every_path(golden):
    """The fixture holds OK, ERANGE, SSV-OK and SSV-deferred cases, both pass
    outcomes, L = 0 and M = 1."""
    z, man = golden
    assert {0, 16} <= set(np.unique(z["case_status"]).tolist())
    assert {0, 16, 19} <= set(np.unique(z["case_ssv"]).tolist())
    assert {0, 1} <= set(np.unique(z["case_pass"]).tolist())
    assert 0 in z["case_len"] and 1 in [m["M"] for m in man["profiles"]]


def test_tjb_and_nullsc_tables(golden):
    z, _ = golden
    O = Oracle()
    tj = np.array([O.tjb(L) for L in range(z["tjb"].size)], np.uint8)
    assert np.array_equal(tj, z["tjb"])
    ns = np.array([np.float32(O.nullsc(L)).view(np.uint32) for L in range(1, z["tjb"].size)], np.uint32)
    assert np.array_equal(ns, z["nullsc_bits"])


def test_pass_rule_grid(golden):
    z, man = golden
    O = Oracle()
    for i in range(z["pass_L"].size):
        p = golden_profile(z, man, int(z["pass_prof"][i]))
        usc = z["pass_usc_bits"][i:i + 1].view(np.float32)[0]
        assert O.passes(usc, int(z["pass_L"][i]), p["mu"], p["lambda"]) == z["pass_bit"][i], i


def test_fn3_7less_drome_passes(golden):
    """tutorial/fn3.out:110: fn3 vs 7LESS_DROME passes the MSV filter."""
    z, man = golden
    c = man["named_cases"]["7LESS_DROME"]
    assert z["case_pass"][c] == 1
