"""The hull sweep's candidate keeping (Top3 in csrc/d3d_device.hpp), restated
on the host and checked against brute force: after a lane's sweep, whenever
its third best score lies below the rescan threshold, its two kept indices
hold every vertex scoring at or above the threshold, and m1, m2 are the two
largest scores (ties kept).  This is what lets a failed certificate evaluate
only the kept candidates in FP64 instead of sweeping the hull again."""
import math

import numpy as np

SENTINEL = 0x7FFFFFFF


def fmaxf(a, b):  # IEEE maxNum, as the device's fmaxf: NaN loses
    if math.isnan(a):
        return b
    if math.isnan(b):
        return a
    return max(a, b)


def fminf(a, b):
    if math.isnan(a):
        return b
    if math.isnan(b):
        return a
    return min(a, b)


class Top3:
    def __init__(self):
        self.m1 = self.m2 = self.m3 = -math.inf
        self.i1 = self.i2 = SENTINEL

    def take(self, sc, j):
        c1, c2 = sc > self.m1, sc > self.m2
        self.m3 = self.m2 if c2 else fmaxf(self.m3, sc)
        self.i2 = self.i1 if c1 else (j if c2 else self.i2)
        self.m2 = fmaxf(self.m2, fminf(sc, self.m1))
        if c1:
            self.m1, self.i1 = sc, j

    def holds(self, thr):
        over = (not (self.m3 < thr)) or ((not (self.m1 < thr)) and self.i1 == SENTINEL) or \
               ((not (self.m2 < thr)) and self.i2 == SENTINEL)
        return not over


def _streams(rng, n_streams=3000):
    for _ in range(n_streams):
        n = int(rng.integers(0, 12))
        # few distinct values, so ties are common
        vals = rng.choice(np.float32([-2.0, -1.0, 0.0, 0.5, 1.0, 1.0, 2.0, 3.0]), n)
        if n and rng.random() < 0.1:
            vals[rng.integers(0, n)] = np.nan
        yield [float(v) for v in vals]


def test_top_scores_and_kept_candidates_cover_the_threshold():
    rng = np.random.default_rng(0)
    checked = 0
    for vals in _streams(rng):
        t = Top3()
        for j, v in enumerate(vals):
            t.take(v, j)
        finite = sorted((v for v in vals if not math.isnan(v)), reverse=True)
        if not any(math.isnan(v) for v in vals):
            # m1, m2, m3: the three largest scores, ties counted
            padded = finite + [-math.inf] * 3
            assert (t.m1, t.m2, t.m3) == tuple(padded[:3])
            if finite:
                assert vals[t.i1] == t.m1 and t.i1 == vals.index(t.m1)
            if len(finite) > 1:
                assert vals[t.i2] == t.m2
        for thr in (-math.inf, -1.5, 0.0, 0.75, 1.0, 2.5, math.nan):
            if math.isnan(thr) or thr == -math.inf:
                assert not t.holds(thr)  # never trusted: the full rescan runs
                continue
            if t.holds(thr):
                need = {j for j, v in enumerate(vals) if not math.isnan(v) and v >= thr}
                kept = {i for i, m in ((t.i1, t.m1), (t.i2, t.m2)) if not (m < thr)}
                assert need <= kept, (vals, thr, need, kept)
                assert SENTINEL not in kept
                checked += 1
    assert checked > 1000
