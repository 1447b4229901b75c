"""Writes the golden vectors in this folder from the compiled HMMER reference.

Run in a container that has the reference tree (``make -C oracle ref`` first):

    python tests/golden/make_golden.py

Inputs and expected outputs only (no reference source): byte profiles and
scalars of sampled and shipped models, digitized targets, and the reference's
p7_MSVFilter / p7_SSVFilter statuses and scores, pass bits, tjb_b and nullsc
tables, RNG streams and generator samples.  Provenance goes to manifest.json.
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


def fasta_text(path):
    """Sequence text of a one-record FASTA or EMBL/UniProt file."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    if lines and lines[0].startswith(">"):
        return "".join(x.strip() for x in lines[1:])
    out, inseq = [], False
    for line in lines:
        if line.startswith("SQ "):
            inseq = True
        elif line.startswith("//"):
            break
        elif inseq:
            out.append("".join(ch for ch in line if ch.isalpha()))
    return "".join(out)


def main():
    R = Reference()
    rng = np.random.default_rng(20261203)

    profiles = []      # (name, RefProfile)
    # Sampled models (p7_hmm_Sample, esl_randomness_Create(seed)) with p7_Calibrate.
    for seed, M in [(1, 200), (3, 400), (10, 50), (11, 2500), (2, 1), (5, 16), (6, 17), (7, 33), (12, 129),
                    (13, 1000), (14, 4100), (15, 6000)]:
        profiles.append((f"sample-s{seed}-M{M}", R.sample(seed, M, calibrate=True), seed))
    # Shipped models with their own STATS lines.
    for fname, which, name in [("minifam", 0, "globins4"), ("minifam", 1, "fn3"), ("minifam", 2, "Pkinase"),
                               ("MADE1.hmm", 0, "MADE1")]:
        profiles.append((name, R.from_file(os.path.join(TUT, fname), which), None))

    prof_rbv, prof_meta = [], []
    case_prof, case_off, case_len, case_status, case_usc, case_ssv, case_pass = [], [], [], [], [], [], []
    residues = []
    pos = 0

    def add_case(pi, rp, x):
        nonlocal pos
        x = np.asarray(x, np.uint8)
        st, usc = rp.msv(x)
        sst, _ = rp.ssv(x)
        ps = rp.passes(len(x), usc) if st == 0 else 1
        case_prof.append(pi)
        case_off.append(pos)
        case_len.append(len(x))
        case_status.append(st)
        case_usc.append(np.float32(usc).view(np.uint32))
        case_ssv.append(sst)
        case_pass.append(ps)
        residues.append(x)
        pos += len(x)

    seqs_named = {}
    for pi, (name, rp, seed) in enumerate(profiles):
        b = rp.byte_profile()
        K = 20 if rp.Kp == 29 else 4
        prof_rbv.append(b["rbv"].reshape(-1))
        prof_meta.append({"name": name, "seed": seed, "M": rp.M, "Kp": rp.Kp, "tbm": b["tbm"], "tec": b["tec"],
                          "base": b["base"], "bias": b["bias"],
                          "scale_bits": int(np.float32(b["scale"]).view(np.uint32)),
                          "mu_bits": int(np.float32(b["mu"]).view(np.uint32)),
                          "lambda_bits": int(np.float32(b["lambda"]).view(np.uint32))})
        lens = [0, 1, 2, 3, 15, 16, 17, 31, 100, 300, 1000]
        if rp.M <= 2500:
            lens += [4000]
        if rp.M <= 400:
            lens += [30000]
        for L in lens:
            add_case(pi, rp, rng.integers(0, K, L))
        # degenerate and special codes, all codes, a run of one residue
        add_case(pi, rp, rng.integers(0, rp.Kp, 400))
        add_case(pi, rp, np.full(60, K, np.uint8))                    # gap symbol only: all costs 255
        add_case(pi, rp, np.full(300, 0, np.uint8))
        if rp.Kp == 29:
            # planted core emissions: high scores, ERANGE and J-state paths
            hom, hl = rp.core_emit(1000 + pi, 12)
            o = 0
            for L in hl:
                add_case(pi, rp, hom[o:o + L])
                o += L
            # partial homologs: a core emission embedded in background
            hom2, hl2 = rp.core_emit(2000 + pi, 6)
            o = 0
            for L in hl2:
                h = hom2[o:o + L][: max(1, L // 3)]
                o += L
                flank = rng.integers(0, 20, 200).astype(np.uint8)
                add_case(pi, rp, np.concatenate([flank, h, flank, h, flank]))
        if name == "fn3":
            x = rp.digitize(fasta_text(os.path.join(TUT, "7LESS_DROME")))
            seqs_named["7LESS_DROME"] = len(case_len)
            add_case(pi, rp, x)
        if name == "globins4":
            x = rp.digitize(fasta_text(os.path.join(TUT, "HBB_HUMAN")))
            seqs_named["HBB_HUMAN"] = len(case_len)
            add_case(pi, rp, x)

    res = np.concatenate(residues) if residues else np.zeros(0, np.uint8)
    rbv_off = np.concatenate([[0], np.cumsum([a.size for a in prof_rbv])[:-1]]).astype(np.int64)

    # Length tables: tjb_b and nullsc bits for L = 0..40000.
    p0 = profiles[0][1]
    Ls = np.arange(0, 40001)
    tjb = np.array([p0.tjb(int(L)) for L in Ls], np.uint8)
    nullsc = np.array([np.float32(p0.nullsc(int(L))).view(np.uint32) for L in Ls[1:]], np.uint32)

    # Pass decisions near the threshold: every xJ-equivalent usc for a few L.
    pass_L, pass_usc, pass_bit, pass_prof = [], [], [], []
    for pi in (0, 1, 13):
        rp = profiles[pi][1]
        b = rp.byte_profile()
        for L in (1, 20, 100, 350, 1000, 5000, 35000):
            tj = p0.tjb(L)
            for xJ in range(0, 256):
                sc = np.float32((np.float32(xJ - tj) - np.float32(b["base"])) / np.float32(b["scale"]))
                sc = np.float32(np.float64(sc) - 3.0)
                pass_prof.append(pi)
                pass_L.append(L)
                pass_usc.append(sc.view(np.uint32))
                pass_bit.append(rp.passes(L, float(sc)))

    # RNG streams and generators.
    u_mt = R.uniform(7, 0, 4096)
    u_fast = R.uniform(42, 1, 4096)
    iid2 = R.iid(2, 16, 300)
    ce_res, ce_len = profiles[0][1].core_emit(5, 20)

    np.savez_compressed(
        os.path.join(HERE, "msv_golden.npz"),
        prof_rbv=np.concatenate(prof_rbv), prof_rbv_off=rbv_off,
        case_prof=np.array(case_prof, np.int32), case_off=np.array(case_off, np.int64),
        case_len=np.array(case_len, np.int32), case_status=np.array(case_status, np.int32),
        case_usc_bits=np.array(case_usc, np.uint32), case_ssv=np.array(case_ssv, np.int32),
        case_pass=np.array(case_pass, np.uint8), res=res,
        tjb=tjb, nullsc_bits=nullsc,
        pass_prof=np.array(pass_prof, np.int32), pass_L=np.array(pass_L, np.int32),
        pass_usc_bits=np.array(pass_usc, np.uint32), pass_bit=np.array(pass_bit, np.uint8),
        u_mt7=u_mt, u_fast42=u_fast, iid_s2=iid2, core_s5_res=ce_res, core_s5_len=ce_len,
    )
    manifest = {
        "generator": "tests/golden/make_golden.py",
        "reference": "forestcheng/cudahmmsearch code/ (HMMER 3.1b1 + Easel h3.1b1), compiled by oracle/Makefile",
        "profiles": prof_meta,
        "named_cases": seqs_named,
        "notes": [
            "sampled profiles: p7_hmm_Sample(esl_randomness_Create(seed)) + p7_Calibrate + "
            "p7_ProfileConfig(L=100, p7_LOCAL) + p7_oprofile_Convert",
            "shipped models: tutorial/minifam (globins4, fn3, Pkinase), tutorial/MADE1.hmm; STATS from file",
            "case_status/usc: p7_MSVFilter after p7_oprofile_ReconfigLength(om, L); case_ssv: p7_SSVFilter",
            "case_pass: MSV stage of p7_Pipeline with F1 = 0.02 (ERANGE counts as pass)",
            "u_mt7: esl_random from esl_randomness_Create(7); u_fast42: from esl_randomness_CreateFast(42)",
            "iid_s2: 16 x esl_rsq_xfIID(L=300) from esl_randomness_Create(2) over p7_AminoFrequencies",
            "core_s5: 20 p7_CoreEmit samples of profile 0's HMM from esl_randomness_Create(5)",
        ],
        "counts": {"cases": len(case_len), "residues": int(res.size)},
    }
    with open(os.path.join(HERE, "manifest.json"), "w") as fh:
        json.dump(manifest, fh, indent=1)
    st = np.array(case_status)
    ss = np.array(case_ssv)
    print("cases", len(case_len), "residues", res.size, "status", dict(zip(*np.unique(st, return_counts=True))),
          "ssv", dict(zip(*np.unique(ss, return_counts=True))))


if __name__ == "__main__":
    main()
