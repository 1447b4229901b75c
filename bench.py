"""MSV filter benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c3|c1|c2|c4|c5|scan|forms]

A step is one complete MSV search (fast J-free stage, exact byte-MSV stage for
the queued pairs, finalize to usc / status / F1 pass bit) of the workload's
profiles over its resident target database; inputs are in HBM before the
timed region starts.  Prints one JSON line.  GCUPS = sum over pairs of M * L
per second / 1e9.  The default workload is c3, the length-skewed 64-profile
batch: value = batched GCUPS, beside the plain single-profile kernel on the
same data (plain_gcups, batched_over_plain, bit identity), the roofline of
the dominant batched launch (HIP events), the reference's SSE p7_MSVFilter
on the host (N threads and one), and a shortened c5 serving phase (serving:
persistent serving, one query per search, and batched serving, the queries
waiting for a slot folded into one search; for each the capacity, p50/p99
over a drained sweep at 50/75/90% of it, and p50/p99 at --rate req/s for
--serving-duration s).  For c2, c3 and c4 the line also carries the
ViterbiFilter stage on the MSV survivors (vit_stage: kernel time, GCUPS, roofline, the
reference's SSE p7_ViterbiFilter on a sample with a bit-identity check); the
c3 line also carries it on the c2 survivors (vit_stage_c2).

Workloads (synthetic data generated on the host from seeds by
cudahmmsearch_amd.workloads; Easel-compatible generators):
  c1  the reference's CPU configuration: one sampled profile M=200 (seed 1)
      vs 20,000 targets, log-normal median 300, sigma 0.6, clipped 30..5000
      (seed 2), 1% replaced by p7_CoreEmit samples (seed 5); also times the
      one-target drop-in (msvg_MSVFilter) per call
  c2  one sampled profile M=400 (p7_hmm_Sample seed 3) vs 2,000,000 targets,
      log-normal lengths median 350, sigma 0.7, clipped 20..10000 (length and
      residue seed 4), 0.5% replaced by p7_CoreEmit samples (seed 5)
  c3  64 sampled profiles, M uniform in [50, 2500] (seeds 10..73) vs 1,000,000
      targets, log-normal median 300, sigma 0.6, clipped 20..35000, plus a 2%
      Pareto tail (x_m 1000, alpha 1.5, max 35000), seed 6
  c4  8 sampled profiles, M in {6000, 8000, 10000, 12000} x 2 (seeds 80..87)
      vs 500,000 targets, log-normal median 1200, sigma 0.6, clipped
      100..35000, seed 7
  c5  serving: resident database of 10,000,000 targets (log-normal median
      350, sigma 0.7, clipped 20..10000, seed 8); queries drawn from 64
      sampled profiles with M log-normal (median 250, sigma 0.6, clipped
      30..4000, seeds 1000..1063), calibrated beforehand (as HMM files carry
      their STATS lines); persistent and batched serving, each: capacity
      (192 queries queued at once), p50/p99 over Poisson arrivals at
      50/75/90% of it (drained), and at --rate req/s for --duration s.
      value = batched serving's capacity in queries/s.
  scan  hmmscan direction: --scan-profiles sampled profiles (M log-normal
      median 250, sigma 0.6, clipped 30..2560, seeds 100000..) against
      --scan-queries query sequences (log-normal median 350, sigma 0.7,
      seed 6), one msvg_scan call per query, then all of them as one
      msvg_scan_many batch
  forms  the packed-form probe (msvg_probe_forms): the J-free recurrence in
      the kernels' f16 clamp form and in an unsigned 16-bit form on the same
      cells, value = the f16 form's speed-up
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# VALU issue peak of MI355X: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane-instructions/s.
PEAK_LANE_INSTR = 256 * 64 * 2.4e9
# Algorithmic work per DP cell: max(diag, B), + emission, max into E = 3 ops.
# On packed f16 lanes (2 cells per VGPR) that is 1/2 v_pk_add_f16 (the clamp
# bit is the max with B) + 1/4 v_pk_maximum3_f16 (two E maxima per
# instruction) = 0.75 lane-instructions per cell.
INSTR_PER_CELL = 0.75
# ViterbiFilter stage: per cell the reference's recurrence (vitfilter.c:146-172)
# does 14 int16 operations (M: 5 saturating adds + 3 max + 1 max into xE; D:
# 1 add + 1 max into Dmax; I: 2 adds + 1 max); 2 cells per packed VGPR.
VIT_INSTR_PER_CELL = 7.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_mode(ctx, pids, did, mode, steps, warmup):
    for _ in range(warmup):
        ctx.search_device(pids, did, mode=mode)
    t = []
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.search_device(pids, did, mode=mode)     # returns after the stream is synchronized
        t.append(time.perf_counter() - t0)
    return t, ctx.stats()


def host_facts():
    """CPU model and logical CPUs of this host (the box the GPU sits in)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(profs, db, budget_s=12.0, one_budget_s=6.0):
    """The reference's SSE p7_MSVFilter (oracle/_ref, compiled from the
    reference sources) on bounded samples of the same targets: N host
    threads over a sample sized to budget_s, and one thread over a smaller
    one (the reference's own per-core figure, msvfilter_bench's Mc/s).  The
    profile is the workload's median-M one."""
    import oracle
    import cudahmmsearch_amd as ca
    host = host_facts()
    nthreads = min(16, host["usable_cpus"] or 1)
    kind = "reference"
    try:
        R = oracle.Reference()
    except (FileNotFoundError, OSError):
        R = None
        kind = "port"
    O = None if R is not None else oracle.Oracle()
    p = sorted(profs, key=lambda q: q.M)[len(profs) // 2]
    if R is not None:
        mat, ins, t = p.hmm
        rp = R.from_probs(mat, t, 100, False)
        run = lambda sub, nt: rp.msv_batch(sub.res, sub.off, sub.lens, nt)  # noqa: E731
    else:
        run = lambda sub, nt: O.msv_batch(p.as_dict(), sub.res, sub.off, sub.lens, nt)  # noqa: E731
    rs = np.random.default_rng(0)

    def timed(nt, budget):
        n0 = min(db.n, 2000)
        idx = rs.choice(db.n, n0, replace=False)
        sub = ca.Database.from_list([db.seq(i) for i in idx])
        t0 = time.perf_counter()
        run(sub, nt)
        dt = time.perf_counter() - t0
        n1 = int(min(db.n, max(n0, n0 * budget / max(dt, 1e-3))))
        idx = rs.choice(db.n, n1, replace=False)
        sub = ca.Database.from_list([db.seq(i) for i in idx])
        t0 = time.perf_counter()
        run(sub, nt)
        dt = time.perf_counter() - t0
        return p.M * sub.residues / dt / 1e9, n1, sub.residues, dt

    v, n1, res, dt = timed(nthreads, budget_s)
    v1, m1, res1, dt1 = timed(1, one_budget_s)
    return {"value": round(v, 3), "unit": "GCUPS", "cores": nthreads, "kind": kind,
            "sample": f"{n1} targets ({res} residues) drawn from the workload, profile M={p.M}, "
                      f"p7_MSVFilter on {nthreads} threads, {dt:.1f} s",
            "one_thread": {"value": round(v1, 3), "unit": "GCUPS",
                           "sample": f"{m1} targets ({res1} residues), same profile, 1 thread, {dt1:.1f} s"},
            "host": host}


def vit_stage(ca, ctx, profs, pids, did, db, reps=3, cpu=True):
    """The ViterbiFilter stage on the MSV survivors of the batched search
    (pass = 1), as p7_Pipeline would run it (p7_pipeline.c:699-705, here
    without the CPU bias filter in between): kernel time from HIP events,
    GCUPS over M * L of the survivors, and the reference's SSE
    p7_ViterbiFilter on a sample of the same pairs."""
    n = db.n
    passed = np.zeros((len(pids), n), np.uint8)
    ctx.lib.msvg_results(ctx._c, None, None, passed.ctypes.data)
    lens = np.asarray(db.lens, np.int64)
    # All profiles' survivors in one multi-profile call (msvg_vit_search_many:
    # the launches overlap, so one profile's longest-target tail runs beside
    # the others' work); the stage time is that call's.
    survs = []
    for i, (p, pid) in enumerate(zip(profs, pids)):
        ctx.set_viterbi(pid, p.viterbi())
        survs.append(np.flatnonzero(passed[i]).astype(np.int64))
    res = ctx.vit_search_many(pids, did, survs)             # warm-up (tables)
    tot_ms = None
    for _ in range(reps):
        ctx.vit_search_many(pids, did, survs)
        ms = ctx.stats()["ms_fast"]
        tot_ms = ms if tot_ms is None else min(tot_ms, ms)
    tot_cells = int(sum(p.M * int(lens[sv].sum()) for p, sv in zip(profs, survs)))
    pairs = int(sum(sv.size for sv in survs))
    first = next(((p, sv, r[0], r[1]) for p, sv, r in zip(profs, survs, res) if sv.size), None)
    out = {"pairs": pairs, "cells": tot_cells, "kernel_ms": round(tot_ms, 3),
           "gcups": round(tot_cells / (tot_ms / 1e3) / 1e9, 1) if tot_ms > 0 else None}
    if tot_ms > 0:
        ach = VIT_INSTR_PER_CELL * tot_cells / (tot_ms / 1e3)
        out["roofline"] = {"bound": "valu", "achieved": round(ach / 1e12, 3),
                           "peak": round(PEAK_LANE_INSTR / 1e12, 3), "unit": "T lane-instr/s",
                           "frac": round(ach / PEAK_LANE_INSTR, 4),
                           "kernel": "vit_band (one launch per lanes/registers class of the profiles) / vit_seg (column segments of the longer models, one launch per segment)",
                           "algorithmic": f"14 int16 ops per DP cell = {VIT_INSTR_PER_CELL} packed lane-instructions"}
    if cpu and first is not None:
        out["cpu_baseline"] = vit_cpu_baseline(ca, db, *first)
    return out


def vit_stage_of(name, ca, workloads, cpu=True):
    """vit_stage on another workload's MSV survivors (its own context and
    database, built and searched in batched mode first)."""
    profs, db, desc = workloads.make_workload(name)
    ctx = ca.MSVContext(0)
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    did = ctx.add_database(db)
    ctx.search_device(pids, did, mode=ca.MODE_BATCHED)
    out = {"workload": f"{name}: {desc}"}
    out.update(vit_stage(ca, ctx, profs, pids, did, db, cpu=cpu))
    ctx.close()
    return out


def vit_cpu_baseline(ca, db, p, surv, sc, st, budget_s=8.0):
    """The reference's SSE p7_ViterbiFilter (oracle/_ref) on up to budget_s of
    the first profile's survivors, with a bit-identity check against the GPU."""
    import oracle
    try:
        R = oracle.Reference()
    except (FileNotFoundError, OSError):
        return None
    nthreads = min(16, os.cpu_count() or 1)
    mat, ins, t = p.hmm
    rp = R.from_probs(mat, t, 100, False)
    k = min(surv.size, 4000)
    for _ in range(2):                         # a timing probe, then a sample sized to the budget
        take = surv[:k]
        sub = ca.Database.from_list([db.seq(int(i)) for i in take])
        t0 = time.perf_counter()
        rst, rsc = rp.vit_batch(sub.res, sub.off, sub.lens, nthreads)
        dt = time.perf_counter() - t0
        if k == surv.size:
            break
        k = int(min(surv.size, max(k, k * budget_s / max(dt, 1e-3))))
    k = take.size
    same = bool(np.array_equal(rsc.view(np.uint32), sc[:k].view(np.uint32))
                and np.array_equal(rst.astype(np.uint8), st[:k]))
    return {"value": round(p.M * sub.residues / dt / 1e9, 3), "unit": "GCUPS", "cores": nthreads,
            "kind": "reference", "bit_identical_to_gpu": same,
            "sample": f"{k} MSV survivors ({sub.residues} residues), profile M={p.M}, "
                      f"p7_ViterbiFilter on {nthreads} threads, {dt:.2f} s"}


def dropin_stage(ctx, pid, db, usc_batched, n=2000):
    """The one-target drop-in (msvg_MSVFilter, what p7_MSVFilter calls per
    target): host-to-host latency per call on the first n targets, and
    whether each call returns the batched search's score."""
    n = min(n, db.n)
    seqs = [db.seq(i) for i in range(n)]
    ctx.msv_filter(pid, seqs[0])                            # warm-up
    sc = np.zeros(n, np.float32)
    t0 = time.perf_counter()
    for i, x in enumerate(seqs):
        sc[i] = ctx.msv_filter(pid, x)[1]
    dt = time.perf_counter() - t0
    return {"calls": n, "us_per_call": round(dt / n * 1e6, 1), "calls_per_s": round(n / dt, 1),
            "identical_to_batched": bool(np.array_equal(sc.view(np.uint32), usc_batched[:n].view(np.uint32)))}


def scan_workload(ca, args):
    """hmmscan direction: a database of sampled profiles (M log-normal median
    250, sigma 0.6, clipped 30..2560, seeds 100000+i) against query sequences
    (log-normal median 350, sigma 0.7, clipped 20..10000, seed 6), one
    msvg_scan call per query (host to host: query upload, launches, results,
    the F1 rule), then all queries as one msvg_scan_many batch (median of
    five calls).  Also: the
    search path on the same pairs (msvg_search over
    a one-target database per query) and the reference's SSE p7_MSVFilter on
    a sample of the profiles."""
    nprof, nq = args.scan_profiles, args.scan_queries
    rs = np.random.default_rng(100)
    Ms = np.clip(np.round(250 * np.exp(0.6 * rs.standard_normal(nprof))), 30, 2560).astype(int)
    t0 = time.perf_counter()
    profs = [ca.sample_profile(100000 + i, int(M)) for i, M in enumerate(Ms)]
    qdb = ca.gen_database(6, nq, 350, 0.7, 20, 10000, nthreads=16)
    log(f"scan workload: {nprof} profiles, {nq} queries ({qdb.residues} residues) in {time.perf_counter() - t0:.1f} s")
    ctx = ca.MSVContext(0)
    pids = np.array([ctx.add_profile(p) for p in profs], np.int32)
    queries = [qdb.seq(i) for i in range(nq)]
    sumM = int(Ms.sum())
    cells = sumM * qdb.residues
    for q in queries[:args.warmup + 1]:                        # table builds, warm-up
        ctx.scan(pids, q)
    res = []
    ms_k = 0.0
    t0 = time.perf_counter()
    for q in queries:
        res.append(ctx.scan(pids, q))
        ms_k += ctx.stats()["ms_fast"]
    wall = time.perf_counter() - t0
    # the same queries as one msvg_scan_many batch: each profile's table is
    # staged once per batch and its workgroup's waves share the queries
    ctx.scan_many(pids, qdb)
    walls, kms = [], []
    for _ in range(5):                                        # median of five calls
        t0 = time.perf_counter()
        bu, bst, bps = ctx.scan_many(pids, qdb)
        walls.append(time.perf_counter() - t0)
        st_b = ctx.stats()
        kms.append(st_b["ms_fast"])
    wall_b = float(np.median(walls))
    ms_b = float(np.median(kms))
    same_b = all(np.array_equal(bu[k].view(np.uint32), res[k][0].view(np.uint32)) and np.array_equal(bst[k], res[k][1])
                 and np.array_equal(bps[k], res[k][2]) for k in range(nq))
    # the search path on the first queries, identity checked
    nsp = min(2, nq)
    same = True
    t1 = time.perf_counter()
    for k in range(nsp):
        did = ctx.add_database(ca.Database.from_list([queries[k]]))
        u, st, ps = ctx.search(pids, did, mode=ca.MODE_PLAIN)
        ctx.remove_database(did)
        same = same and np.array_equal(u[:, 0].view(np.uint32), res[k][0].view(np.uint32)) \
            and np.array_equal(st[:, 0], res[k][1]) and np.array_equal(ps[:, 0], res[k][2])
    search_s = (time.perf_counter() - t1) / nsp
    out = {"metric": "MSV scan (hmmscan direction: profiles x one query) GCUPS on 1 MI355X", "value": round(cells / wall / 1e9, 1),
           "unit": "GCUPS", "n_gpus": 1, "steps": nq, "warmup": args.warmup + 1,
           "ms_per_step": round(wall / nq * 1e3, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "u8", "data": "synthetic",
           "config": {"workload": f"scan: {nprof} sampled profiles (M log-normal median 250, sigma 0.6, clip 30-2560; "
                                  f"sum M {sumM}) vs {nq} queries (log-normal median 350, sigma 0.7)",
                      "profiles": nprof, "queries": nq, "query_residues": qdb.residues, "cells": cells},
           "queries_per_s": round(nq / wall, 1), "kernel_ms_per_query": round(ms_k / nq, 3),
           "kernel_gcups": round(cells / (ms_k / 1e3) / 1e9, 1),
           "search_path_ms_per_query": round(search_s * 1e3, 1),
           "scan_over_search": round(search_s / (wall / nq), 1), "identical_to_search": bool(same),
           "batch": {"queries": nq, "calls": 5, "queries_per_s": round(nq / wall_b, 1), "gcups": round(cells / wall_b / 1e9, 1),
                     "kernel_ms": round(ms_b, 3), "kernel_gcups": round(cells / (ms_b / 1e3) / 1e9, 1),
                     "j_aware_pairs": int(st_b["n_jrow"]), "search_path_pairs": int(st_b["n_slow"]),
                     "identical_to_single": bool(same_b)}}
    if not args.no_cpu:
        import oracle
        try:
            R = oracle.Reference()
            nthreads = min(16, os.cpu_count() or 1)
            take = rs.choice(nprof, min(nprof, 400), replace=False)
            rps = []
            for i in take:                                   # profile conversion, not timed
                mat, ins, t = profs[i].hmm
                rps.append(R.from_probs(mat, t, 100, False))
            t2 = time.perf_counter()
            ccells = 0
            for i, rp in zip(take, rps):
                rp.msv_batch(qdb.res, qdb.off, qdb.lens, nthreads)
                ccells += int(Ms[i]) * qdb.residues
            dt = time.perf_counter() - t2
            out["cpu_baseline"] = {"value": round(ccells / dt / 1e9, 3), "unit": "GCUPS", "cores": nthreads,
                                   "kind": "reference",
                                   "sample": f"{take.size} of the profiles against all {nq} queries, p7_MSVFilter "
                                             f"on {nthreads} threads over the queries, one profile at a time, "
                                             f"{dt:.2f} s"}
        except (FileNotFoundError, OSError):
            out["cpu_baseline"] = None
    log(f"scan: {out['queries_per_s']} queries/s, {out['value']} GCUPS host to host, kernels {out['kernel_gcups']} GCUPS; "
        f"search path {out['search_path_ms_per_query']} ms per query; one batch of {nq}: "
        f"{out['batch']['queries_per_s']} queries/s, kernels {out['batch']['kernel_gcups']} GCUPS")
    print(json.dumps(out), flush=True)
    ctx.close()


def packed_forms(ca, args):
    """The packed-form probe (msvg_probe_forms) as its own bench line: the
    J-free recurrence of the 64-lane layout in the kernels' f16 form and in
    an unsigned 16-bit form on the same cells (value = f16 cells/s over u16
    cells/s)."""
    ctx = ca.MSVContext(0)
    r = [ctx.probe_forms(16384) for _ in range(3)]
    ctx.close()
    f16 = min(x["ms_f16"] for x in r)
    u16 = min(x["ms_u16"] for x in r)
    cells = r[0]["cells"]
    out = {
        "metric": "J-free cells per second, f16 clamp form over unsigned 16-bit form (probe on 1 MI355X)",
        "value": round(u16 / f16, 3), "unit": "x", "n_gpus": 1, "steps": 3, "warmup": 1,
        "ms_per_step": round(f16, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16 / u16", "data": "synthetic",
        "config": {"workload": "forms: 64-lane groups, Q = 16 (2048 model positions), 16 waves per CU, 16,384 rows "
                               "per wave, costs and residues from a fixed LCG"},
        "f16": {"ms": round(f16, 3), "gcups": round(cells / f16 / 1e6, 1),
                "lane_instr_per_cell": 0.75, "instructions": "v_pk_add_f16 clamp per 2 cells, v_pk_maximum3_f16 per 4"},
        "u16": {"ms": round(u16, 3), "gcups": round(cells / u16 / 1e6, 1),
                "lane_instr_per_cell": 1.5,
                "instructions": "v_pk_add_u16 + v_pk_sub_u16 clamp + v_pk_max_u16 per 2 cells"},
        "agree": all(x["agree"] for x in r), "cells_per_launch": cells,
    }
    print(json.dumps(out))


def copy_probe(nbytes=256 << 20):
    """One device-to-device copy of a known size, for scaling FETCH_SIZE /
    WRITE_SIZE counters (the copy reads and writes nbytes each)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    hip.hipMalloc(C.byref(a), C.c_size_t(nbytes))
    hip.hipMalloc(C.byref(b), C.c_size_t(nbytes))
    hip.hipMemset(a, 1, C.c_size_t(nbytes))
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)   # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy probe: {nbytes} bytes device to device")


def serving_phase(ca, srv, profs, rate, duration, rs, drain):
    """Open-loop Poisson arrivals; returns latencies (ms) of finished queries
    and the counts of submitted / finished / unfinished ones."""
    t0 = time.perf_counter()
    tickets = []
    nxt = 0.0
    while True:
        nxt += rs.exponential(1.0 / rate)
        if nxt > duration:
            break
        now = time.perf_counter() - t0
        if nxt > now:
            time.sleep(nxt - now)
        tickets.append(srv.submit(profs[int(rs.integers(len(profs)))]))
    lat, cells, unfinished = [], 0, 0
    for t in tickets:
        r = srv.wait(t, timeout=(-1.0 if drain else 0.0), want_hits=False)
        if r is None:
            unfinished += 1
            continue
        info = r[0]
        lat.append((info["t_done"] - info["t_submit"]) * 1e3)
        cells += info["cells"]
    return np.array(lat), len(tickets), unfinished, cells, time.perf_counter() - t0


def serving_mode(ca, ctx, did, profs, mode, slots, rate, duration, sweep_s, rs):
    """One serving mode on the resident database: warm (every query profile
    once), the service capacity (3 x 64 queries queued at once), a drained
    open-loop sweep at 50/75/90% of that capacity, then Poisson arrivals at
    the offered rate for duration seconds (not drained: the server stops
    with the unfinished queries)."""
    kw = {"slots": slots, "shared": mode == "shared", "batched": mode == "batched"}
    srv = ca.Server(ctx, did, **kw)
    for t in [srv.submit(p) for p in profs]:
        srv.wait(t, want_hits=False)
    t0 = time.perf_counter()
    tk = [srv.submit(profs[int(rs.integers(len(profs)))]) for _ in range(3 * len(profs))]
    cap_cells = sum(srv.wait(t, want_hits=False)[0]["cells"] for t in tk)
    cap_s = time.perf_counter() - t0
    cap_qps = len(tk) / cap_s
    log(f"serving {mode}: capacity {cap_qps:.1f} queries/s, {cap_cells / cap_s / 1e9:.0f} GCUPS")
    pct = lambda a, q: round(float(np.percentile(a, q)), 2) if a.size else None  # noqa: E731
    sweep = []
    for f in (0.5, 0.75, 0.9):
        b0 = srv.stats()["batches"]
        a0 = srv.stats()["admitted"]
        lat, sub, _, cells, wall = serving_phase(ca, srv, profs, f * cap_qps, sweep_s, rs, drain=True)
        st = srv.stats()
        sweep.append({"rate_per_s": round(f * cap_qps, 2), "submitted": sub, "p50_ms": pct(lat, 50),
                      "p99_ms": pct(lat, 99), "gcups": round(cells / wall / 1e9, 1),
                      "queries_per_search": round((st["admitted"] - a0) / max(1, st["batches"] - b0), 2)})
        log(f"serving {mode}: {sweep[-1]}")
    bounded = [x for x in sweep if x["p99_ms"] is not None and x["p99_ms"] <= 1000.0]
    st0 = srv.stats()
    lat, sub, unf, cells, wall = serving_phase(ca, srv, profs, rate, duration, rs, drain=False)
    st = srv.stats()
    srv.stop()
    busy = st["busy_s"] - st0["busy_s"]
    log(f"serving {mode}: offered {rate}/s: submitted {sub}, finished {lat.size}, unfinished {unf}")
    return {
        "capacity": {"queries_per_s": round(cap_qps, 2), "gcups": round(cap_cells / cap_s / 1e9, 1)},
        "sweep": sweep,
        "highest_bounded_rate_per_s": max(x["rate_per_s"] for x in bounded) if bounded else None,
        "offered": {"rate_per_s": rate, "duration_s": duration, "submitted": sub, "finished": int(lat.size),
                    "unfinished": unf, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                    "max_queue": st["max_pending"],
                    "queries_per_search": round((st["admitted"] - st0["admitted"]) /
                                                max(1, st["batches"] - st0["batches"]), 2),
                    "gcups_while_serving": round(cells / max(busy, 1e-9) / 1e9, 1)},
    }


def serving_run(ca, n_targets, rate, duration, slots, modes, sweep_s):
    """Serving on a resident database of n_targets (the c5 generator) with
    the 64 c5 query profiles (calibrated first, as model files carry their
    STATS lines), in each of the given modes (serving_mode)."""
    from cudahmmsearch_amd import workloads
    t0 = time.perf_counter()
    db = workloads.serving_database(n_targets)
    profs = workloads.serving_queries()
    Ms = np.array([p.M for p in profs])
    log(f"serving database: {db.n} targets, {db.residues} residues; {len(profs)} query profiles, "
        f"median M {int(np.median(Ms))} ({time.perf_counter() - t0:.1f} s)")
    ctx = ca.MSVContext(0)
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        ctx.remove_profile(pid)
    did = ctx.add_database(db)
    out = {
        "workload": f"c5: {db.n} resident targets ({db.residues} residues), 64 query profiles "
                    f"M log-normal median 250 (mean {float(np.mean(Ms)):.0f}), {slots} slots",
        "cells_per_query_mean": round(float(np.mean(Ms)) * db.residues),
        "note": "capacity: 192 queries queued at once; sweep: Poisson arrivals at 50/75/90% of the mode's "
                f"capacity for {sweep_s:g} s each, drained; highest_bounded_rate: the highest swept rate with "
                f"p99 <= 1 s; offered: Poisson arrivals at {rate:g}/s for {duration:g} s, latencies of the "
                "finished queries only (above capacity the queue grows for the whole run)",
    }
    for m in modes:             # the same query sequence and arrival times in every mode
        out[m] = serving_mode(ca, ctx, did, profs, m, slots, rate, duration, sweep_s, np.random.default_rng(2006))
    if "persistent" in out and "batched" in out:
        out["batched_over_persistent_capacity"] = round(out["batched"]["capacity"]["gcups"] /
                                                        out["persistent"]["capacity"]["gcups"], 3)
    ctx.close()
    return out


def serving(ca, args):
    """The c5 workload as its own bench line (value = the batched server's
    capacity in queries per second)."""
    r = serving_run(ca, 10_000_000, args.rate, args.duration, args.slots, ["persistent", "batched"],
                    max(8.0, args.duration / 4))
    b = r["batched"]
    out = {
        "metric": "MSV serving capacity (queries/s) on 1 MI355X, batched serving",
        "value": b["capacity"]["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": b["offered"]["submitted"],
        "warmup": 64, "ms_per_step": round(1e3 / b["capacity"]["queries_per_s"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": r["workload"], "cells_per_query_mean": r["cells_per_query_mean"]},
        "serving": r,
    }
    print(json.dumps(out))


def roofline(se, cells, npids):
    """Roofline of the dominant kernel from an eager search's stats: the
    batched launch (msv_batch, one per lane-group size) that ran longest,
    or, when no profile went through the batched kernel (column-segmented
    models, tables beyond LDS), the whole J-free stage over all its cells.  achieved = 0.75 packed
    lane-instructions per cell x the launch's cells / its HIP-event time."""
    nb = int(se.get("n_batch", 0))
    if nb:
        k = max(range(nb), key=lambda j: se["batch_ms"][j])
        ms, lcells, G = se["batch_ms"][k], se["batch_cells"][k], se["batch_lanes"][k]
        kernel = f"msv_batch<{G}> (one persistent launch over the {G}-lane profiles)"
        tot_ms, tot_cells = sum(se["batch_ms"][:nb]), sum(se["batch_cells"][:nb])
    else:                       # the whole J-free stage (its launches run one after another)
        ms, lcells = se["ms_fast"], cells
        kernel = f"msv_seg / msv_fast (the J-free stage's {int(se['n_fast_launches'])} launches together)"
        tot_ms, tot_cells = se["ms_fast"], cells
    ach = INSTR_PER_CELL * lcells / (ms / 1e3)
    cps = lcells / (ms / 1e3)
    roof = {"bound": "valu", "achieved": round(ach / 1e12, 3), "peak": round(PEAK_LANE_INSTR / 1e12, 3),
            "unit": "T lane-instr/s", "frac": round(ach / PEAK_LANE_INSTR, 4), "traffic": None,
            "kernel": kernel, "ms_per_launch": round(ms, 3), "cells_per_launch": int(lcells),
            "cells_per_s": round(cps, 1),
            "frac_3ops_packed16": round(3 * cps / (2 * PEAK_LANE_INSTR), 4),
            "algorithmic": f"3 ops per DP cell = {INSTR_PER_CELL} packed lane-instructions per cell (fused: "
                           f"v_pk_add_f16 with clamp per 2 cells + v_pk_maximum3_f16 per 4 cells); "
                           f"frac_3ops_packed16 counts 3 ops per cell against 2 x the issue peak",
            "all_batched_launches": {"launches": nb, "ms": round(tot_ms, 3), "cells": int(tot_cells),
                                     "frac": round(INSTR_PER_CELL * tot_cells / (tot_ms / 1e3) / PEAK_LANE_INSTR, 4)
                                     if tot_ms > 0 else None},
            "j_stage_ms": round(se.get("ms_jrow", 0.0), 3)}
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-probe", action="store_true", help="run a 256 MiB D2D copy first (counter scaling)")
    ap.add_argument("--batched-only", action="store_true", help="time only the batched mode (profiling runs)")
    ap.add_argument("--rate", type=float, default=500.0, help="serving: offered queries per second")
    ap.add_argument("--duration", type=float, default=60.0, help="c5: seconds of arrivals")
    ap.add_argument("--slots", type=int, default=2, help="serving: searches on the GPU at once")
    ap.add_argument("--no-vit", action="store_true", help="skip the ViterbiFilter stage on the MSV survivors")
    ap.add_argument("--no-serving", action="store_true", help="c3: skip the serving sub-object")
    ap.add_argument("--no-vit-c2", action="store_true", help="c3: skip the ViterbiFilter stage on the c2 survivors")
    ap.add_argument("--serving-duration", type=float, default=10.0, help="c3: seconds of arrivals in the serving phase")
    ap.add_argument("--scan-profiles", type=int, default=10000, help="scan: profiles in the scanned database")
    ap.add_argument("--scan-queries", type=int, default=64, help="scan: query sequences")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; across GPUs the search would shard the "
                                   "database by residue count (one process per GPU, results gathered), "
                                   "which is not launched here", "n_gpus": args.gpus}))
        sys.exit(2)

    import cudahmmsearch_amd as ca
    from cudahmmsearch_amd import workloads
    if args.workload == "c5":
        serving(ca, args)
        return
    if args.workload == "scan":
        scan_workload(ca, args)
        return
    if args.workload == "forms":
        packed_forms(ca, args)
        return
    if args.copy_probe:
        copy_probe()
    t0 = time.perf_counter()
    profs, db, desc = workloads.make_workload(args.workload)
    log(f"workload {args.workload}: {len(profs)} profiles, {db.n} targets, {db.residues} residues "
        f"({time.perf_counter() - t0:.1f} s)")
    ctx = ca.MSVContext(0)
    pids = []
    for p in profs:
        pid = ctx.add_profile(p)
        p.mu = ctx.calibrate(pid, p.lam)
        pids.append(pid)
    did = ctx.add_database(db)
    log("database resident")
    cells = sum(p.M for p in profs) * db.residues

    tp = None
    if not args.batched_only:
        tp, sp = time_mode(ctx, pids, did, ca.MODE_PLAIN, args.steps, args.warmup)
        n_plain = np.zeros((len(pids), db.n), np.uint8)
        u_plain = np.zeros((len(pids), db.n), np.float32)
        ctx.lib.msvg_results(ctx._c, u_plain.ctypes.data, None, n_plain.ctypes.data)
        log(f"plain: {np.median(tp) * 1e3:.1f} ms/step, fast stage {sp['ms_fast']:.1f} ms")
    tb, sb = time_mode(ctx, pids, did, ca.MODE_BATCHED, args.steps, args.warmup)
    identical = None
    if tp is not None:
        u_b = np.zeros_like(u_plain)
        p_b = np.zeros_like(n_plain)
        ctx.lib.msvg_results(ctx._c, u_b.ctypes.data, None, p_b.ctypes.data)
        identical = bool(np.array_equal(u_b.view(np.uint32), u_plain.view(np.uint32)) and np.array_equal(p_b, n_plain))
        del u_b, p_b, u_plain, n_plain
    log(f"batched: {np.median(tb) * 1e3:.1f} ms/step; identical to plain: {identical}")
    # one eager batched run for the per-stage and per-launch breakdown (same kernels, no graph)
    ctx.search_device(pids, did, mode=ca.MODE_BATCHED | ca.FLAG_EAGER)
    se = ctx.stats()

    ms_b = float(np.median(tb) * 1e3)
    ms_p = float(np.median(tp) * 1e3) if tp is not None else None
    gcups_b = cells / (ms_b / 1e3) / 1e9
    gcups_p = cells / (ms_p / 1e3) / 1e9 if ms_p else None
    roof = roofline(se, cells, len(pids))
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("kernel", "").split(" ")[0] == roof["kernel"].split(" ")[0]:
            roof["traffic"] = tr["bytes_per_launch"]      # rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled
            roof["traffic_source"] = os.path.relpath(tpath, ROOT)
    out = {
        "metric": "MSV GCUPS on 1 MI355X, batched vs plain kernel (bit-exact); serving p50/p99 ms",
        "value": round(gcups_b, 1), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_b, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {desc}", "targets": db.n, "residues": db.residues,
                   "profiles": len(profs), "cells_per_step": cells},
        "plain_gcups": round(gcups_p, 1) if gcups_p else None,
        "plain_ms_per_step": round(ms_p, 3) if ms_p else None,
        "batched_over_plain": round(gcups_b / gcups_p, 3) if gcups_p else None,
        "bit_identical_plain_batched": identical,
        "stages_ms": {"fast": round(se["ms_fast"], 3), "slow": round(se["ms_slow"], 3),
                      "finalize": round(se["ms_finalize"], 3)},
        "launches": {"j_free": int(se["n_fast_launches"]), "batched": int(se["n_batch"]),
                     "batched_lanes": se["batch_lanes"][:se["n_batch"]],
                     "batched_ms": [round(x, 3) for x in se["batch_ms"][:se["n_batch"]]]},
        "deferred_pairs": {"j_aware": int(se["n_jrow"]), "byte_msv": int(se["n_slow"])}, "roofline": roof,
    }
    # LDS traffic of the J-free pass: per target row each lane reads its Q
    # registers' emissions (4 bytes each) and G lanes advance one target,
    # i.e. Q * 4 * G bytes per target row, Q = ceil(M / 2G).
    lds_bytes = 0
    for p in profs:
        G, _, seg = ctx.tile_of(p.M, did)
        if not seg:
            Q = -(-p.M // (2 * G))
            lds_bytes += Q * 4 * G * db.residues
    if lds_bytes and se["ms_fast"] > 0:
        peak_lds = ctx.probe_lds()
        ach_lds = lds_bytes / (se["ms_fast"] / 1e3)
        roof["lds"] = {"achieved": round(ach_lds / 1e12, 3), "probe_peak": round(peak_lds / 1e12, 3),
                       "unit": "TB/s", "frac": round(ach_lds / peak_lds, 4), "bytes_per_search": int(lds_bytes),
                       "note": "emission reads from LDS over the whole J-free stage against the device's "
                               "measured ds_read_b128 bandwidth"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(profs, db)
        log(f"cpu baseline: {out['cpu_baseline']}")
    if args.workload == "c1":
        ctx.search_device(pids, did, mode=ca.MODE_BATCHED)
        ub = np.zeros((len(pids), db.n), np.float32)
        ctx.lib.msvg_results(ctx._c, ub.ctypes.data, None, None)
        out["dropin"] = dropin_stage(ctx, pids[0], db, ub[0])
        out["dropin"]["gcups"] = round(profs[0].M * sum(int(db.lens[i]) for i in range(out["dropin"]["calls"]))
                                       * out["dropin"]["calls_per_s"] / out["dropin"]["calls"] / 1e9, 3)
        log(f"drop-in: {out['dropin']}")
    if not args.no_vit and args.workload in ("c2", "c3", "c4"):
        # the batched search above left its results (and the pass bits) on the device
        ctx.search_device(pids, did, mode=ca.MODE_BATCHED)
        out["vit_stage"] = vit_stage(ca, ctx, profs, pids, did, db, cpu=not args.no_cpu)
        log(f"vit stage: {out['vit_stage']}")
    ctx.close()
    del db
    if not args.no_vit and args.workload == "c3" and not args.no_vit_c2:
        out["vit_stage_c2"] = vit_stage_of("c2", ca, workloads, cpu=not args.no_cpu)
        log(f"vit stage (c2 survivors): {out['vit_stage_c2']}")
    if args.workload == "c3" and not args.no_serving:
        out["serving"] = serving_run(ca, 10_000_000, args.rate, args.serving_duration, args.slots,
                                     ["persistent", "batched"], 6.0)
        log(f"serving: {out['serving']}")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
