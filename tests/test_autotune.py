"""The autotuner's bin construction on the CPU: a stand-in context whose
J-free pass time is a known function of (M, lanes), with crossovers at
M = 120 (8 -> 16 lanes), M = 700 (16 -> 32) and M = 1500 (32 -> 64; but 32
lanes stop fitting the batched kernel above M = 1024).  The bins must
follow the winners, their edges must land within the bisection's
resolution of the true crossovers (or of the feasibility limit), and a
lane count inside the tolerance of the best must not split a bin."""
import numpy as np

import cudahmmsearch_amd as ca
from cudahmmsearch_amd import autotune

CROSS = {8: (0, 120), 16: (120, 700), 32: (700, 1500), 64: (1500, 1 << 20)}


def cost(M, lanes):
    lo, hi = CROSS[lanes]
    d = 0.0 if lo <= M < hi else min(abs(M - lo), abs(M - hi)) / max(M, 1) + 0.05
    return M * 1e-3 * (1.0 + d)


class FakeCtx:
    def __init__(self, cost_fn):
        self.cost, self.tile, self.M, self.calls = cost_fn, None, None, 0

    def set_tail_guard(self, on):
        pass

    def add_database(self, db):
        return 1

    def remove_database(self, did):
        pass

    def add_profile(self, prof):
        self.M = prof.M
        return 7

    def calibrate(self, pid, lam):
        return -8.0

    def remove_profile(self, pid):
        self.M = None

    def clear_tiles(self):
        self.tile = None

    def set_tile(self, lo, hi, lanes, waves):
        self.tile = lanes

    def search_device(self, pids, did, mode=0):
        self.calls += 1

    def stats(self):
        return {"ms_fast": self.cost(self.M, self.tile)}


def test_bins_follow_the_measured_crossovers():
    db = ca.gen_database(5, 200, 300, 0.5, 20, 2000)
    ctx = FakeCtx(cost)
    bins = autotune.tune(ctx, db, m_lo=16, m_hi=3000, probes=18, log=lambda *_: None)
    assert [b["lanes"] for b in bins] == [8, 16, 32, 64]
    assert bins[0]["lo"] == 1 and bins[-1]["hi"] == 3000
    for a, b in zip(bins, bins[1:]):
        assert b["lo"] == a["hi"] + 1
    for b, edge in zip(bins[:-1], (120, 700)):
        assert abs(b["hi"] + 1 - edge) <= max(3, edge // 32), (b["hi"], edge)
    assert bins[2]["hi"] == 1024                # 32 lanes win up to where they stop fitting
    for b in bins:
        assert autotune.feasible(b["lo"], b["lanes"]) or b["lo"] == 1


def test_a_near_tie_keeps_the_previous_bin():
    # 16 lanes 1% slower than 8 lanes above M = 300: inside the 2% tolerance,
    # so one bin of 8 lanes up to where 8 lanes stop fitting
    def near(M, lanes):
        base = {8: 1.0, 16: 1.01 if M > 300 else 1.2, 32: 1.3, 64: 1.5}[lanes]
        return M * 1e-3 * base
    db = ca.gen_database(5, 200, 300, 0.5, 20, 2000)
    bins = autotune.tune(FakeCtx(near), db, m_lo=16, m_hi=256, probes=10, log=lambda *_: None)
    assert [b["lanes"] for b in bins] == [8]


def test_a_bin_ends_where_its_choice_stops_fitting():
    # probes up to M = 4096, where no lane count fits: the 64-lane bin ends
    # at its own limit (M = 3584 over 20 residue codes), not at the last probe
    db = ca.gen_database(5, 200, 300, 0.5, 20, 2000)
    bins = autotune.tune(FakeCtx(cost), db, m_lo=16, m_hi=4096, probes=18, log=lambda *_: None)
    assert bins[-1]["lanes"] == 64 and bins[-1]["hi"] == 3584


def test_feasible_matches_the_batched_kernel_limits():
    # registers per lane: 16 at 8 and 32 lanes, 24 at 16, 32 at 64
    assert autotune.feasible(256, 8) and not autotune.feasible(257, 8)
    assert autotune.feasible(768, 16) and not autotune.feasible(769, 16)
    assert autotune.feasible(1024, 32) and not autotune.feasible(1025, 32)
    # 64 lanes: 21 rows x ceil(Q/4) KiB within 160 KiB -> Q <= 28
    assert autotune.feasible(3584, 64) and not autotune.feasible(3585, 64)
    # narrow layouts hold at most 32 residue rows
    assert not autotune.feasible(100, 16, rows=33)


# ViterbiFilter shapes: 16 lanes best below M = 300, 32 lanes to 800, one
# wave to 3000 (but it stops fitting above M = 2432, 19 registers), two waves
# beyond (up to M = 4096, 16 registers), where four waves take over; eight
# waves always slower.
VCROSS = {(16, 1): (0, 300), (32, 1): (300, 800), (64, 1): (800, 3000), (64, 2): (3000, 1 << 20),
          (64, 4): (1 << 20, 1 << 21), (64, 8): (1 << 20, 1 << 21)}


def vcost(M, shape):
    lo, hi = VCROSS[shape]
    d = 0.0 if lo <= M < hi else min(abs(M - lo), abs(M - hi)) / max(M, 1) + 0.05
    return M * 1e-3 * (1.0 + d)


class FakeVitCtx(FakeCtx):
    def __init__(self, cost_fn):
        super().__init__(cost_fn)
        self.vtiles, self.shape, self.targets = [], None, None

    def clear_vit_tiles(self):
        self.vtiles = []

    def set_vit_tile(self, lo, hi, lanes, waves=1):
        self.vtiles.append((lo, hi, lanes, waves))

    def set_viterbi(self, pid, vit):
        assert vit.M == self.M
        self.shape = None
        for lo, hi, lanes, waves in self.vtiles:
            if lo <= self.M <= hi:
                self.shape = (lanes, waves)
        assert self.shape is not None and autotune.vit_feasible(self.M, self.shape)

    def vit_search(self, pid, did, targets):
        self.targets = np.asarray(targets)
        self.calls += 1

    def stats(self):
        return {"ms_fast": self.cost(self.M, self.shape)}


def test_vit_bins_follow_the_measured_crossovers():
    db = ca.gen_database(5, 300, 300, 0.5, 20, 4000)
    ctx = FakeVitCtx(vcost)
    bins = autotune.vit_tune(ctx, db, m_lo=16, m_hi=5000, probes=16, n_targets=100, max_len=1500,
                             log=lambda *_: None)
    assert [(b["lanes"], b["waves"]) for b in bins] == [(16, 1), (32, 1), (64, 1), (64, 2), (64, 4)]
    assert bins[0]["lo"] == 1 and bins[-1]["hi"] == 5000
    for a, b in zip(bins, bins[1:]):
        assert b["lo"] == a["hi"] + 1
    for b, edge in zip(bins[:2], (300, 800)):
        assert abs(b["hi"] + 1 - edge) <= max(3, edge // 32), (b["hi"], edge)
    # edges where the left shape stops fitting (and still wins) are exact
    assert bins[2]["hi"] == 2432 and bins[3]["hi"] == 4096
    # timed on a sample of the short targets, and the tiles cleared afterwards
    assert ctx.targets.size == 100 and np.all(np.asarray(db.lens)[ctx.targets] <= 1500)
    assert ctx.vtiles == []


def test_vit_feasible_matches_the_kernel_limits():
    # 16 registers per lane at 16 and 32 lanes and for 2, 4, 8 waves per target
    assert autotune.vit_feasible(512, (16, 1)) and not autotune.vit_feasible(513, (16, 1))
    assert autotune.vit_feasible(1024, (32, 1)) and not autotune.vit_feasible(1025, (32, 1))
    assert autotune.vit_feasible(4096, (64, 2)) and not autotune.vit_feasible(4097, (64, 2))
    assert autotune.vit_feasible(16384, (64, 8)) and not autotune.vit_feasible(16385, (64, 8))
    # several waves only beyond one wave's limit, and only the fewest that fit
    assert not autotune.vit_feasible(2432, (64, 2)) and autotune.vit_feasible(2433, (64, 2))
    assert not autotune.vit_feasible(3000, (64, 4)) and autotune.vit_feasible(4097, (64, 4))
    assert not autotune.vit_feasible(8192, (64, 8))
    # one wave: tables of Kp + 1 emission and 3 transition rows within 160 KiB
    assert autotune.vit_feasible(2432, (64, 1)) and not autotune.vit_feasible(2433, (64, 1))
    assert autotune.vit_feasible(2560, (64, 1), Kp=18)
