"""Pins the oracle's byte MSV loop against an independent formulation: the
best multi-segment path, enumerated over all chains of ungapped diagonal
segments for tiny M and L (the MSV model is N -> B -> M_k..M_k' -> E -> J ...
-> C, with the byte costs of impl_sse/p7_oprofile.c:602-640).  Also pins the
SSV path's J-free value (one segment, floor at the begin score)."""
import itertools

import numpy as np

from oracle import Oracle


def segments(M, L, rbv, x, bias):
    """All (i0, i1, value) of ungapped diagonal runs rows i0..i1 (0-based),
    with value the summed biased emissions, for every start column."""
    out = []
    for i0 in range(L):
        for k0 in range(M):
            s = 0
            for d in range(min(L - i0, M - k0)):
                s += bias - int(rbv[x[i0 + d], k0 + d])
                out.append((i0, i0 + d, s, k0))
    return out


def brute_msv(p, x):
    """(overflow, xJ) by enumerating segment chains."""
    M = p["rbv"].shape[1]
    L = len(x)
    base, bias, tec, tbm = p["base"], p["bias"], p["tec"], p["tbm"]
    tjb = Oracle().tjb(L)
    tjbm = (tjb + tbm) & 255
    segs = segments(M, L, p["rbv"], x, bias)
    # best J value available after row i (J >= 0), by dynamic closure over chains
    bestJ = [0] * (L + 1)     # bestJ[i]: J after rows < i
    maxcell = -10**9
    # chains: process segments ordered by end row; B at start row i0 uses J after row i0-1
    J_after = [0] * L
    by_end = sorted(segs, key=lambda t: t[1])
    cur = 0
    ends = {}
    for seg in by_end:
        ends.setdefault(seg[1], []).append(seg)
    Jrun = 0
    Jhist = []
    for i in range(L):
        # J available to B at row i is Jrun (rows < i)
        Jhist.append(Jrun)
        for (i0, i1, s, k0) in ends.get(i, []):
            pass
        # segments ending at row i: B taken at their start row
        for (i0, i1, s, k0) in ends.get(i, []):
            B = max(0, max(base, Jhist[i0]) - tjbm)
            cell = B + s
            maxcell = max(maxcell, cell)
            Jrun = max(Jrun, cell - tec)
        Jrun = max(Jrun, 0)
    # every partial segment is also a cell (prefixes are segments themselves)
    overflow = maxcell + bias >= 255
    return overflow, Jrun


def test_brute_force_tiny():
    O = Oracle()
    rs = np.random.default_rng(11)
    checked = 0
    for trial in range(400):
        M = int(rs.integers(1, 5))
        L = int(rs.integers(1, 6))
        bias = int(rs.integers(0, 20))
        rbv = rs.integers(0, 60, (29, M)).astype(np.uint8)
        rbv[:, rs.random(M) < 0.2] = 255
        tbm = int(rs.integers(0, 60))
        base = 190
        p = {"rbv": rbv, "tbm": tbm, "tec": 3, "base": base, "bias": bias, "scale": float(np.float32(3 / np.log(2)))}
        x = rs.integers(0, 20, L).astype(np.uint8)
        st, usc = O.msv_bytes(p, x)
        ov, xJ = brute_msv(p, x)
        if ov:
            assert st == 16, (trial, M, L)
        else:
            assert st == 0
            assert np.float32(usc) == np.float32(O.score(xJ, O.tjb(L), base, p["scale"])), (trial, M, L)
        checked += 1
    assert checked == 400


def test_ssv_single_segment_value():
    """When SSV returns a score, xJ = max(B0, best single segment) - tec."""
    O = Oracle()
    rs = np.random.default_rng(12)
    n_ok = 0
    for trial in range(300):
        M = int(rs.integers(1, 6))
        L = int(rs.integers(1, 7))
        rbv = rs.integers(0, 40, (29, M)).astype(np.uint8)
        p = {"rbv": rbv, "tbm": int(rs.integers(10, 40)), "tec": 3, "base": 190, "bias": int(rs.integers(5, 20)),
             "scale": float(np.float32(3 / np.log(2)))}
        x = rs.integers(0, 20, L).astype(np.uint8)
        st, usc = O.ssv(p, x)
        if st != 0:
            continue
        tjb = O.tjb(L)
        B0 = p["base"] - tjb - p["tbm"]
        best = max([B0] + [B0 + s for (_, _, s, _) in segments(M, L, rbv, x, p["bias"])])
        assert np.float32(usc) == np.float32(O.score(best - p["tec"], tjb, p["base"], p["scale"]))
        n_ok += 1
    assert n_ok > 100
