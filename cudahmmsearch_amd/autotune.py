"""Autotuner for the batched MSV launch and the ViterbiFilter shapes.

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

The ViterbiFilter shapes (lanes and waves per target) are chosen the same
way (vit_tune), timing the stage on a sample of the database's targets, and
installed with msvg_set_vit_tile.

    python -m cudahmmsearch_amd.autotune [--targets 300000] [--out tiles.json] [--only msv|vit]
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


def _bins(measure, cands, grid, m_hi, tol, log, fits=None):
    """Model-length bins from timings: measure(M, keys) returns {key: ms}
    for the keys that fit at M.  The winner at each probe of the grid (a key
    within tol of the best keeps the previous probe's winner), runs of equal
    winners as bins, and between neighbouring probes with different winners
    the edge by bisection on M.  fits(M, key), when given, places an edge
    where the left bin's choice stops fitting exactly (if that choice still
    wins at its last length).  Returns [(lo, hi, key)]."""
    def winner(t, prev=None):
        best = min(t, key=t.get)
        if prev in t and t[prev] <= (1.0 + tol) * t[best]:
            return prev
        return best

    wins, prev = [], None
    for M in grid:
        t = measure(M, cands)
        if not t:
            wins.append(None)
            continue
        prev = winner(t, prev)
        wins.append(prev)
        log(f"M {M}: " + " ".join(f"{g}:{t[g]:.3f}" for g in sorted(t)) + f" -> {prev}")
    bins = []
    k = 0
    while k < len(grid):
        if wins[k] is None:
            k += 1
            continue
        j = k
        while j + 1 < len(grid) and wins[j + 1] == wins[k]:
            j += 1
        lo = 1 if not bins else bins[-1][1] + 1
        hi = grid[j]
        if j + 1 < len(grid) and wins[j + 1] is not None:
            a, b, ga, gb = grid[j], grid[j + 1], wins[k], wins[j + 1]
            edge = None
            if fits is not None and not fits(b, ga):
                lo_f, hi_f = a, b            # ga fits at lo_f, not at hi_f
                while hi_f - lo_f > 1:
                    mid = (lo_f + hi_f) // 2
                    lo_f, hi_f = (mid, hi_f) if fits(mid, ga) else (lo_f, mid)
                t = measure(lo_f, (ga, gb))
                if gb not in t or winner(t, ga) == ga:
                    edge = lo_f              # ga wins up to where it stops fitting
                else:
                    b = lo_f
            while edge is None and b - a > max(2, a // 32):
                mid = (a + b) // 2
                t = measure(mid, (ga, gb))
                if ga not in t:          # the left bin's choice stops fitting: the edge is below
                    b = mid
                elif gb not in t or winner(t, ga) == ga:
                    a = mid
                else:
                    b = mid
            hi = a if edge is None else edge
        elif j + 1 == len(grid):
            hi = m_hi
        elif fits is not None:               # nothing fits at the next probe: up to where this fits
            lo_f, hi_f = grid[j], grid[j + 1]
            while hi_f - lo_f > 1:
                mid = (lo_f + hi_f) // 2
                lo_f, hi_f = (mid, hi_f) if fits(mid, wins[k]) else (lo_f, mid)
            hi = lo_f
        bins.append((lo, hi, wins[k]))
        log(f"bin M {lo}-{hi}: {wins[k]}")
        k = j + 1
    return bins


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

    grid = sorted({int(round(math.exp(x))) for x in np.linspace(math.log(m_lo), math.log(m_hi), probes)})
    bins = [{"lo": lo, "hi": hi, "lanes": g, "waves": 0,
             "probes": {str(M): {str(x): round(v, 4) for x, v in sorted(timings[M].items())}
                        for M in sorted(timings) if lo <= M <= hi}}
            for lo, hi, g in _bins(measure, LANES, grid, m_hi, tol, log, lambda M, g: feasible(M, g, rows))]
    ctx.clear_tiles()
    ctx.remove_database(did)
    ctx.set_tail_guard(True)
    return bins


# ViterbiFilter shapes (lanes per target, waves per target): vit_band with 16
# or 32 lanes (several targets per wave) or a whole wave; beyond one wave's
# tables, the column segments (vit_seg), whose exact fallback runs on the
# multi-wave kernel with the fewest waves (2, 4 or 8) that fit.
VIT_SHAPES = ((16, 1), (32, 1), (64, 1), (64, 2), (64, 4), (64, 8))


def _one_wave_qmax(Kp):
    return min(20, LDS_BYTES // (256 * (Kp + 4)))


def vit_feasible(M, shape, Kp=29):
    """Does the ViterbiFilter take a model of length M in this shape?  At
    most 16 registers per lane, except one wave per target: as many as keep
    its tables (Kp + 1 emission rows and 3 transition rows of 64 lanes x 4
    bytes per register) within 160 KiB of LDS, at most 20.  Several waves
    (column segments, the fallback on that many waves) only beyond one
    wave's limit and only with the fewest waves that fit: the waves change
    nothing but the rare fallback, so the tuner has one such candidate."""
    lanes, waves = shape
    if waves == 1:
        Q = (M + 2 * lanes - 1) // (2 * lanes)
        return Q <= (16 if lanes < 64 else _one_wave_qmax(Kp))
    fits = lambda w: (M + 128 * w - 1) // (128 * w) <= 16    # noqa: E731
    return M > 128 * _one_wave_qmax(Kp) and fits(waves) and (waves == 2 or not fits(waves // 2))


def time_vit_shape(ctx, prof, did, targets, shape, reps=3):
    ctx.clear_vit_tiles()
    ctx.set_vit_tile(1, 1 << 20, shape[0], shape[1])
    pid = ctx.add_profile(prof)
    ctx.set_viterbi(pid, prof.viterbi())
    best = float("inf")
    for _ in range(reps + 1):                       # the first call builds the tables
        ctx.vit_search(pid, did, targets)
        best = min(best, ctx.stats()["ms_fast"])
    ctx.remove_profile(pid)
    return best


def vit_tune(ctx, db, m_lo=16, m_hi=4096, probes=20, tol=0.02, seed=1900, n_targets=40000, max_len=2000,
             log=print):
    """ViterbiFilter shape bins, as tune() does for the MSV pass: each probe
    length is timed in every shape that fits, on a sample of n_targets of
    the database's targets of up to max_len residues (the stage's throughput
    regime; a launch with a much longer target is bound by that target's
    row chain in every shape).  Returns [{"lo", "hi", "lanes", "waves",
    "probes"}] for msvg_set_vit_tile."""
    did = ctx.add_database(db)
    lens = np.asarray(db.lens)
    pool = np.flatnonzero(lens <= max_len)
    rng = np.random.default_rng(seed)
    targets = np.sort(rng.choice(pool, min(n_targets, pool.size), replace=False)).astype(np.int64)
    timings = {}

    def measure(M, shapes):
        out = timings.setdefault(M, {})
        todo = [sh for sh in shapes if sh not in out and vit_feasible(M, sh)]
        if todo:
            prof = ca.sample_profile(seed + M, M)
            for sh in todo:
                out[sh] = time_vit_shape(ctx, prof, did, targets, sh)
        return {sh: out[sh] for sh in shapes if sh in out}

    grid = sorted({int(round(math.exp(x))) for x in np.linspace(math.log(m_lo), math.log(m_hi), probes)})
    bins = [{"lo": lo, "hi": hi, "lanes": sh[0], "waves": sh[1],
             "probes": {str(M): {f"{x[0]}x{x[1]}": round(v, 4) for x, v in sorted(timings[M].items())}
                        for M in sorted(timings) if lo <= M <= hi}}
            for lo, hi, sh in _bins(measure, VIT_SHAPES, grid, m_hi, tol, log, vit_feasible)]
    ctx.clear_vit_tiles()
    ctx.remove_database(did)
    return bins


def install(ctx, table, vit_table=None):
    ctx.clear_tiles()
    for t in table:
        ctx.set_tile(t["lo"], t["hi"], t["lanes"], t["waves"])
    if vit_table is not None:
        ctx.clear_vit_tiles()
        for t in vit_table:
            ctx.set_vit_tile(t["lo"], t["hi"], t["lanes"], t["waves"])


def load(path=DEFAULT_TABLE, key="tiles"):
    """The persisted table: "tiles" (batched MSV pass) or "vit_tiles"
    (ViterbiFilter shapes); None if absent."""
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh).get(key)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--targets", type=int, default=300000)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--only", choices=["msv", "vit"], help="tune one table, keep the other from --out")
    args = ap.parse_args()
    db = ca.gen_database(77, args.targets, 350, 0.7, 20, 10000, nthreads=16)
    ctx = ca.MSVContext(0, tiles=None)
    old = {}
    if os.path.exists(args.out):
        with open(args.out) as fh:
            old = json.load(fh)
    out = dict(old, device="gfx950", sample_targets=db.n, sample_residues=db.residues)
    if args.only != "vit":
        t0 = time.time()
        out["tiles"] = tune(ctx, db)
        out["seconds"] = round(time.time() - t0, 1)
    if args.only != "msv":
        t0 = time.time()
        out["vit_tiles"] = vit_tune(ctx, db)
        out["vit_seconds"] = round(time.time() - t0, 1)
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(f"wrote {args.out}")
    ctx.close()


if __name__ == "__main__":
    main()
