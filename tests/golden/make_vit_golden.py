"""Writes vit_golden.npz / vit_manifest.json from the compiled HMMER reference:
the ViterbiFilter stage (impl_sse/vitfilter.c:84-250).

Run in a container that has the reference tree (``make -C oracle ref`` first):

    python tests/golden/make_vit_golden.py

Inputs and expected outputs only: the ViterbiFilter fields of P7_OPROFILE
(de-striped over every padded position), the N/C/J move cost for a table of
target lengths, digitized targets, and p7_ViterbiFilter's status and score
bits for each (profile, target) case.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import Reference  # noqa: E402

TUT = "/root/reference/code/tutorial"
MOVE_L = np.array(list(range(0, 201)) + [250, 300, 350, 447, 500, 1000, 2000, 5000, 10000, 35000, 100000, 1000000],
                  np.int32)


def delete_heavy(R, rng, M, md, dd):
    """An HMM whose delete paths are likely, so the lazy-F delete pass runs."""
    mat = rng.dirichlet(np.ones(20) * 0.3, size=M + 1).astype(np.float32)
    t = np.zeros((M + 1, 7), np.float32)
    mm = 1.0 - md - 0.02
    t[:] = [mm, 0.02, md, 0.6, 0.4, 1.0 - dd, dd]
    t[0] = [mm, 0.02, md, 0.6, 0.4, 1.0, 0.0]
    t[M] = [0.98, 0.02, 0.0, 0.6, 0.4, 1.0, 0.0]
    return R.from_probs(mat, t, calibrate=True)


def main():
    R = Reference()
    rng = np.random.default_rng(20261204)
    profiles = []
    for seed, M in [(1, 200), (3, 400), (2, 1), (5, 16), (6, 17), (7, 33), (10, 50), (12, 129), (13, 1000),
                    (11, 2500), (14, 4100)]:
        profiles.append((f"sample-s{seed}-M{M}", R.sample(seed, M, calibrate=True), seed))
    for fname, which, name in [("minifam", 0, "globins4"), ("minifam", 1, "fn3"), ("minifam", 2, "Pkinase"),
                               ("MADE1.hmm", 0, "MADE1")]:
        profiles.append((name, R.from_file(os.path.join(TUT, fname), which), None))
    for i, (M, md, dd) in enumerate([(120, 0.3, 0.9), (120, 0.05, 0.5), (300, 0.45, 0.99)]):
        profiles.append((f"delete-heavy-{i}-M{M}", delete_heavy(R, rng, M, md, dd), None))

    rwv_all, twv_all, meta, moves = [], [], [], []
    c_prof, c_off, c_len, c_status, c_bits = [], [], [], [], []
    residues = []
    pos = 0
    for pi, (name, rp, seed) in enumerate(profiles):
        f = rp.vit_fields()
        M, Kp = rp.M, rp.Kp
        rwv_all.append(np.ascontiguousarray(f["rwv_pad"][:, :M]).ravel())
        twv_all.append(np.ascontiguousarray(f["twv_pad"][:, :M]).ravel())
        pad_ok = bool((f["rwv_pad"][:, M:] == -32768).all() and (f["twv_pad"][:4, M:] == -32768).all()
                      and (f["twv_pad"][4:, M - 1:] == -32768).all())
        meta.append({"name": name, "seed": seed, "M": M, "Kp": Kp, "e_loop": f["e_loop"], "e_move": f["e_move"],
                     "base_w": f["base_w"], "ddbound_w": f["ddbound_w"],
                     "scale_w_bits": int(np.float32(f["scale_w"]).view(np.uint32)),
                     "nj_bits": int(np.float32(f["nj"]).view(np.uint32)),
                     "vmu_bits": int(np.float32(f["vmu"]).view(np.uint32)),
                     "vlambda_bits": int(np.float32(f["vlambda"]).view(np.uint32)),
                     "loops_zero": f["n_loop"] == 0 and f["c_loop"] == 0 and f["j_loop"] == 0,
                     "padding_is_min": pad_ok})
        moves.append([rp.vit_move(int(L)) for L in MOVE_L])

        K = 4 if Kp == 18 else 20
        targets = [np.zeros(0, np.uint8), rng.integers(0, K, 1), rng.integers(0, K, 2), rng.integers(0, K, 3)]
        for L in rng.integers(4, 2000, 14):
            targets.append(rng.integers(0, K, int(L)))
        targets.append(rng.integers(0, Kp, 400))                      # every code, degenerate ones too
        emit, lens = rp.core_emit(1000 + pi, 8)
        o = 0
        for j, L in enumerate(lens):
            hom = emit[o:o + L]
            o += L
            if j % 2 == 0:
                targets.append(hom)
            else:
                bg = rng.integers(0, K, 300)
                targets.append(np.concatenate([bg[:150], hom, bg[150:]]))
        targets.append(np.concatenate([emit] * 3)[:12000])            # long homolog run
        for x in targets:
            x = np.asarray(x, np.uint8)
            st, sc = rp.vit(x)
            c_prof.append(pi)
            c_off.append(pos)
            c_len.append(x.size)
            c_status.append(st)
            c_bits.append(int(np.float32(sc).view(np.uint32)))
            residues.append(x)
            pos += x.size

    np.savez_compressed(
        os.path.join(HERE, "vit_golden.npz"),
        rwv=np.concatenate(rwv_all).astype(np.int16), twv=np.concatenate(twv_all).astype(np.int16),
        move_L=MOVE_L, moves=np.array(moves, np.int32),
        case_prof=np.array(c_prof, np.int32), case_off=np.array(c_off, np.int64),
        case_len=np.array(c_len, np.int32), case_status=np.array(c_status, np.int32),
        case_bits=np.array(c_bits, np.uint32), residues=np.concatenate(residues).astype(np.uint8))
    with open(os.path.join(HERE, "vit_manifest.json"), "w") as fh:
        json.dump({"generator": "tests/golden/make_vit_golden.py",
                   "reference": "forestcheng/cudahmmsearch code/ (HMMER 3.1b1), p7_ViterbiFilter via oracle/ref_harness.c",
                   "profiles": meta, "cases": len(c_prof),
                   "status_counts": {str(s): int((np.array(c_status) == s).sum()) for s in set(c_status)}},
                  fh, indent=1)
    print(f"{len(profiles)} profiles, {len(c_prof)} cases, statuses "
          f"{ {s: c_status.count(s) for s in set(c_status)} }, {pos} residues")


if __name__ == "__main__":
    main()
