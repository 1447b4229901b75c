"""Autotuner for the batched MSV launch.

Chooses the lane-group size G (lanes per target: 8, 16, 32 or 64, so 8, 4,
2 or 1 targets per wavefront) of the batched J-free pass for every model
length, and the model-length bin edges where the choice changes, from
measurements on a sample of the database: each probe length M on a
log-spaced grid is timed at every G whose batched table fits (the batched
kernel's limits: registers per lane, LDS, residue rows), and between two
neighbouring probes with different winners the edge is found by bisection
on M (both winners timed at each midpoint).  A G within ``tol`` of the best
keeps the previous bin's choice, so noise does not split bins.  Installs
the bins with ``msvg_set_tile`` and can persist the table (JSON) for later
contexts.  Tiles never change results, only speed (the GPU tests check
every tile against the oracle).  The workgroup size of the batched launch
follows from its residency (msv_runtime.cpp), so a bin's waves are 0.

    python -m cudahmmsearch_amd.autotune [--targets 300000] [--out tiles.json]
"""
import argparse
import json
import math
import os
import time

import numpy as np

import cudahmmsearch_amd as ca

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles_gfx950.json")
LANES = (8, 16, 32, 64)
LDS_BYTES = 160 * 1024


def feasible(M, lanes, rows=21):
    """Does the batched kernel take a model of length M at this lane-group
    size over a database of rows - 1 residue codes (plus the null row)?  The
    kernel's limits: registers per lane (32 at G = 64, 24 at 16, 16 at 8
    and 32) and the table within LDS (G = 64: rows x ceil(Q/4) KiB; narrow
    layouts: 32 rows x 16 or 32 lanes x 16 bytes per register block, at
    most 32 rows)."""
    Q = (M + 2 * lanes - 1) // (2 * lanes)
    qmax = {64: 32, 16: 24}.get(lanes, 16)
    if Q > qmax:
        return False
    if lanes == 64:
        return rows * ((Q + 3) // 4) * 1024 + 16 <= LDS_BYTES
    tw = 16 if lanes < 32 else 32
    return rows <= 32 and ((Q + 3) // 4) * 32 * tw * 16 + 16 <= LDS_BYTES


def time_tile(ctx, pid, did, lanes, reps=3):
    ctx.clear_tiles()
    ctx.set_tile(1, 1 << 20, lanes, 0)
    best = float("inf")
    for _ in range(reps):
        ctx.search_device([pid], did, mode=ca.MODE_BATCHED | ca.FLAG_EAGER)
        best = min(best, ctx.stats()["ms_fast"])
    return best


def tune(ctx, db, m_lo=16, m_hi=4096, probes=24, tol=0.02, seed=900, log=print):
    """Returns [{"lo", "hi", "lanes", "waves", "probes"}]: bins of model
    length covering m_lo..m_hi (the first bin from M = 1), each with its
    lane-group size and the probe timings (ms of the J-free pass) that
    decided it.  The tail guard is off while timing, so each candidate runs
    the lane-group size it is labelled with."""
    ctx.set_tail_guard(False)
    did = ctx.add_database(db)
    rows = len(np.unique(db.res[:min(db.res.size, 1 << 22)])) + 1
    timings = {}

    def measure(M, lanes_set):
        out = timings.setdefault(M, {})
        todo = [g for g in lanes_set if g not in out and feasible(M, g, rows)]
        if todo:
            prof = ca.sample_profile(seed + M, M)
            pid = ctx.add_profile(prof)
            prof.mu = ctx.calibrate(pid, prof.lam)
            for g in todo:
                out[g] = time_tile(ctx, pid, did, g)
            ctx.remove_profile(pid)
        return {g: out[g] for g in lanes_set if g in out}

    def winner(t, prev=None):
        best = min(t, key=t.get)
        if prev in t and t[prev] <= (1.0 + tol) * t[best]:
            return prev
        return best

    grid = sorted({int(round(math.exp(x))) for x in np.linspace(math.log(m_lo), math.log(m_hi), probes)})
    wins, prev = [], None
    for M in grid:
        t = measure(M, LANES)
        if not t:
            wins.append(None)
            continue
        prev = winner(t, prev)
        wins.append(prev)
        log(f"M {M}: " + " ".join(f"{g}:{t[g]:.3f}" for g in sorted(t)) + f" -> {prev}")
    # bins: runs of equal winners; edges by bisection between neighbours
    bins = []
    k = 0
    while k < len(grid):
        if wins[k] is None:
            k += 1
            continue
        j = k
        while j + 1 < len(grid) and wins[j + 1] == wins[k]:
            j += 1
        lo = 1 if not bins else bins[-1]["hi"] + 1
        hi = grid[j]
        if j + 1 < len(grid) and wins[j + 1] is not None:
            a, b, ga, gb = grid[j], grid[j + 1], wins[k], wins[j + 1]
            while b - a > max(2, a // 32):
                mid = (a + b) // 2
                t = measure(mid, (ga, gb))
                if ga not in t:          # the left bin's lane count stops fitting: the edge is below
                    b = mid
                elif gb not in t or winner(t, ga) == ga:
                    a = mid
                else:
                    b = mid
            hi = a
        elif j + 1 == len(grid):
            hi = m_hi
        bins.append({"lo": lo, "hi": hi, "lanes": wins[k], "waves": 0,
                     "probes": {str(M): {str(g): round(v, 4) for g, v in sorted(timings[M].items())}
                                for M in sorted(timings) if lo <= M <= hi}})
        log(f"bin M {lo}-{hi}: {wins[k]} lanes")
        k = j + 1
    ctx.clear_tiles()
    ctx.remove_database(did)
    ctx.set_tail_guard(True)
    return bins


def install(ctx, table):
    ctx.clear_tiles()
    for t in table:
        ctx.set_tile(t["lo"], t["hi"], t["lanes"], t["waves"])


def load(path=DEFAULT_TABLE):
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)["tiles"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--targets", type=int, default=300000)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    args = ap.parse_args()
    db = ca.gen_database(77, args.targets, 350, 0.7, 20, 10000, nthreads=16)
    ctx = ca.MSVContext(0)
    t0 = time.time()
    table = tune(ctx, db)
    with open(args.out, "w") as fh:
        json.dump({"device": "gfx950", "sample_targets": db.n, "sample_residues": db.residues,
                   "seconds": round(time.time() - t0, 1), "tiles": table}, fh, indent=1)
    print(f"wrote {args.out}")
    ctx.close()


if __name__ == "__main__":
    main()
