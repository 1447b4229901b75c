"""CPU check of the decomposition the HIP path relies on: the composite
p7_MSVFilter result follows from the J-free maximum alone (orc_decide) unless
the J state may fire, in which case the byte MSV decides; and the deferred
pairs are rare on background targets."""
import numpy as np

import cudahmmsearch_amd as ca
from conftest import golden_cases, golden_profile
from oracle import Oracle


def _check(O, p, x):
    st, usc = O.msv(p, x)
    code, xJ = O.decide(p, O.jfree_max(p, x), len(x))
    if code == 19:
        return True, True
    if code == 16:
        return st == 16, False
    ok = st == 0 and np.float32(usc).view(np.uint32) == np.float32(
        O.score(xJ, O.tjb(len(x)), p["base"], p["scale"])).view(np.uint32)
    return ok, False


def test_golden_cases(golden):
    z, man = golden
    O = Oracle()
    deferred = total = 0
    for pi in range(len(man["profiles"])):
        p = golden_profile(z, man, pi)
        for c, _, x in golden_cases(z, pi):
            if p["M"] * len(x) > 3_000_000:
                continue
            ok, d = _check(O, p, x)
            assert ok, (p["name"], c)
            deferred += d
            total += 1
    # the fixture is rich in planted homologs; background targets rarely defer
    assert deferred < total // 4


def test_random_regimes():
    """Guard-true regimes (large M x L), tiny models, degenerate residues."""
    O = Oracle()
    rs = np.random.default_rng(5)
    n = 0
    for seed, M in [(31, 1), (32, 3), (33, 64), (34, 500), (35, 2500)]:
        prof = ca.sample_profile(seed, M)
        p = prof.as_dict()
        for L in [0, 1, 5, 64, 700, 6000]:
            if M * L > 4_000_000:
                continue
            for alphabet in (20, 29):
                x = rs.integers(0, alphabet, L).astype(np.uint8)
                ok, _ = _check(O, p, x)
                assert ok, (M, L, alphabet)
                n += 1
        for h in ca.core_emit(prof.hmm, seed, 4):
            ok, _ = _check(O, p, h)
            assert ok
    assert n > 40
