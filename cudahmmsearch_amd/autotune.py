"""Autotuner for the batched MSV launch.

For each bin of model lengths it times every tile the kernels are built for
-- lanes per target G in {8, 16, 32, 64} (so 8, 4, 2 or 1 targets per wavefront)
and waves per workgroup in {4, 8, 16} -- on a sample of the database, picks
the fastest, installs it with ``msvg_set_tile`` and can persist the table
(JSON) for later contexts.  Tiles never change results, only speed (the
GPU tests check every tile against the oracle).

    python -m cudahmmsearch_amd.autotune [--targets 200000] [--out tiles.json]
"""
import argparse
import json
import os
import time

import numpy as np

import cudahmmsearch_amd as ca

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tiles_gfx950.json")
# Model-length bins: the edges where the register count per lane of a tile
# crosses a multiple of the register block, for the three lane-group sizes.
BIN_EDGES = [1, 65, 129, 257, 385, 449, 513, 769, 1025, 1537, 2049, 3073, 4097]
LANES = (8, 16, 32, 64)
WAVES = (4, 8, 16)


def feasible(M, lanes, rows=21):
    """Q <= 32 registers and the table (rows x ceil(Q/4) KiB) within 160 KiB
    of LDS for a database of 20 residue codes plus the null row."""
    Q = (M + 2 * lanes - 1) // (2 * lanes)
    return Q <= 32 and rows * ((Q + 3) // 4) * 1024 <= 160 * 1024


def time_tile(ctx, pid, did, lanes, waves, reps=3):
    ctx.clear_tiles()
    ctx.set_tile(1, 1 << 20, lanes, waves)
    best = float("inf")
    for _ in range(reps):
        ctx.search_device([pid], did, mode=ca.MODE_BATCHED | ca.FLAG_EAGER)
        best = min(best, ctx.stats()["ms_fast"])
    return best


def tune(ctx, db, edges=BIN_EDGES, seed=900, log=print):
    """Returns [{"lo", "hi", "lanes", "waves", "ms", "candidates"}] for the bins.
    The tail guard is off while timing, so each candidate runs the tile it is
    labelled with."""
    ctx.set_tail_guard(False)
    did = ctx.add_database(db)
    table = []
    for b in range(len(edges) - 1):
        lo, hi = edges[b], edges[b + 1] - 1
        M = (lo + hi) // 2
        prof = ca.sample_profile(seed + b, M)
        pid = ctx.add_profile(prof)
        prof.mu = ctx.calibrate(pid, prof.lam)
        cand = {}
        for lanes in LANES:
            if not feasible(M, lanes) or not feasible(hi, lanes):
                continue
            for waves in WAVES:
                cand[f"{lanes}x{waves}"] = time_tile(ctx, pid, did, lanes, waves)
        if not cand:
            log(f"M {lo}-{hi}: no tile fits LDS over the whole bin; left to the built-in choice")
            ctx.remove_profile(pid)
            continue
        key = min(cand, key=cand.get)
        lanes, waves = (int(v) for v in key.split("x"))
        table.append({"lo": lo, "hi": hi, "lanes": lanes, "waves": waves, "ms": round(cand[key], 4),
                      "candidates": {k: round(v, 4) for k, v in cand.items()}})
        log(f"M {lo}-{hi} (probe M={M}): best {key} {cand[key]:.3f} ms; " +
            " ".join(f"{k}:{v:.3f}" for k, v in sorted(cand.items())))
        ctx.remove_profile(pid)
    ctx.clear_tiles()
    ctx.remove_database(did)
    ctx.set_tail_guard(True)
    return table


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
    ap.add_argument("--targets", type=int, default=200000)
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
