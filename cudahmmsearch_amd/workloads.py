"""The benchmark workloads, generated from seeds on the host.

Shared by ``bench.py`` and the GPU tests so both score the same inputs.
Profiles come from ``sample_profile`` (p7_hmm_Sample + p7_ProfileConfig +
p7_oprofile_Convert, src/p7_hmm.c:877-918, src/impl_sse/p7_oprofile.c:840),
targets from ``gen_database`` (i.i.d. residues from p7_AminoFrequencies,
src/hmmer.c:163-186, log-normal lengths with an optional Pareto tail) and
planted homologs from ``core_emit`` (p7_CoreEmit, src/emit.c:71-148).

  c1  one profile M=200 (seed 1) vs 20,000 targets, log-normal median 300,
      sigma 0.6, clipped 30..5000 (seed 2), 1% replaced by p7_CoreEmit
      samples (seed 5)
  c2  one profile M=400 (seed 3) vs 2,000,000 targets, log-normal median
      350, sigma 0.7, clipped 20..10000 (seed 4), 0.5% planted (seed 5)
  c3  64 profiles, M uniform in [50, 2500] (numpy seed 1003; profile seeds
      10..73) vs 1,000,000 targets, log-normal median 300, sigma 0.6, clipped
      20..35000, plus a 2% Pareto tail (x_m 1000, alpha 1.5, max 35000), seed 6
  c4  8 profiles, M in {6000, 8000, 10000, 12000} x 2 (seeds 80..87) vs
      500,000 targets, log-normal median 1200, sigma 0.6, clipped 100..35000,
      seed 7
  c5  serving: a resident database of n targets (10,000,000 for the full
      workload; log-normal median 350, sigma 0.7, clipped 20..10000, seed 8)
      and 64 query profiles with M log-normal (median 250, sigma 0.6,
      clipped 30..4000; numpy seed 2005, profile seeds 1000..1063)
"""
import numpy as np

from . import core_emit, gen_database, plant, sample_profile


def c3_lengths():
    """The 64 model lengths of c3."""
    return [int(m) for m in np.random.default_rng(1003).integers(50, 2501, 64)]


def make_workload(name, nthreads=16):
    """(profiles, database, description) of workload c1, c2, c3 or c4."""
    if name == "c1":
        profs = [sample_profile(1, 200)]
        db = gen_database(2, 20_000, 300, 0.6, 30, 5000, nthreads=nthreads)
        hom = core_emit(profs[0].hmm, 5, int(0.01 * db.n))
        db, _ = plant(db, hom, 5, 0.01)
        desc = ("1 sampled profile M=200 (seed 1) vs 20,000 i.i.d. targets, log-normal median 300 sigma 0.6 "
                "clip 30-5000, 1% planted p7_CoreEmit homologs")
    elif name == "c2":
        profs = [sample_profile(3, 400)]
        db = gen_database(4, 2_000_000, 350, 0.7, 20, 10000, nthreads=nthreads)
        hom = core_emit(profs[0].hmm, 5, int(0.005 * db.n))
        db, _ = plant(db, hom, 5, 0.005)
        desc = ("1 sampled profile M=400 (seed 3) vs 2,000,000 i.i.d. targets, log-normal median 350 sigma 0.7 "
                "clip 20-10000, 0.5% planted p7_CoreEmit homologs")
    elif name == "c3":
        profs = [sample_profile(10 + i, M) for i, M in enumerate(c3_lengths())]
        db = gen_database(6, 1_000_000, 300, 0.6, 20, 35000, tail_frac=0.02, tail_xm=1000, tail_alpha=1.5,
                          tail_max=35000, nthreads=nthreads)
        desc = ("64 sampled profiles M~U[50,2500] (seeds 10-73) vs 1,000,000 targets, log-normal median 300 sigma 0.6 "
                "+ 2% Pareto tail (x_m 1000, alpha 1.5, max 35000)")
    elif name == "c4":
        Ms = [6000, 6000, 8000, 8000, 10000, 10000, 12000, 12000]
        profs = [sample_profile(80 + i, M) for i, M in enumerate(Ms)]
        db = gen_database(7, 500_000, 1200, 0.6, 100, 35000, nthreads=nthreads)
        desc = "8 sampled profiles M in {6000,8000,10000,12000}x2 vs 500,000 targets, log-normal median 1200"
    else:
        raise ValueError(f"unknown workload {name}")
    return profs, db, desc


def serving_queries():
    """The 64 query profiles of c5 (uncalibrated: mu is set by the caller)."""
    rs = np.random.default_rng(2005)
    Ms = np.clip(np.round(np.exp(np.log(250.0) + 0.6 * rs.standard_normal(64))), 30, 4000).astype(int)
    return [sample_profile(1000 + i, int(M)) for i, M in enumerate(Ms)]


def serving_database(n=10_000_000, nthreads=16):
    """The c5 resident database (n targets; 10,000,000 in the full workload)."""
    return gen_database(8, n, 350, 0.7, 20, 10000, nthreads=nthreads)
