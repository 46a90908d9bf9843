"""The autotuner's selection logic on recorded timings (no GPU): pick,
coordinate descent over the launch shape, the per-class lane choice, and the
tables it writes."""
import json

import numpy as np

from distance3d_amd import autotune

# per-class GJK times (ms) of the 1,048,576-pair mixed workload on one MI355X,
# one entry per candidate lane count (as autotune.tune records them)
RECORDED_CLASS_MS = {
    0: {1: 1.5495, 2: 1.9667, 4: 2.6473},
    1: {1: 0.92, 2: 0.828, 4: 0.7633, 8: 0.7926},
    2: {2: 1.7709, 4: 1.2982, 8: 1.3045, 16: 1.4362},
    3: {4: 2.0177, 8: 1.8787, 16: 1.8666, 32: 2.026},
    4: {8: 1.6173, 16: 1.4656, 32: 1.5229, 64: 1.5615},
    5: {8: 2.2614, 16: 2.0326, 32: 1.8951, 64: 1.825},
    6: {16: 2.8876, 32: 2.5166, 64: 2.3143},
    7: {16: 0.3067, 32: 0.2487, 64: 0.2065},
}


def test_pick_smallest_and_first_on_ties():
    assert autotune.pick({1: 3.0, 2: 1.0, 4: 2.0}) == 2
    assert autotune.pick({8: 1.0, 4: 1.0, 2: 5.0}) == 8
    assert autotune.pick({"a": 0.5}) == "a"


def test_lanes_from_recorded_class_times():
    lanes = [autotune.pick(RECORDED_CLASS_MS[c]) for c in range(8)]
    assert lanes == [1, 4, 4, 16, 16, 64, 64, 64]
    # every choice is a candidate the tuner would try for its class
    for c, g in enumerate(lanes):
        assert g in autotune.CANDIDATES[c]


def _table_measure(table, calls):
    """A recorded step-time table over the launch shape (separable plus an
    interaction term), as a stand-in for the GPU measurement."""
    def measure(cfg):
        calls.append(dict(cfg))
        return table[(cfg["blocks_per_cu"], cfg["claim_factor"], cfg["epa_blocks_per_cu"])]
    return measure


def test_descend_finds_the_recorded_optimum_of_a_separable_table():
    b = {2: 0.4, 3: 0.0}
    c = {1: 0.3, 2: 0.0, 4: 0.1}
    e = {4: 0.25, 8: 0.05, 12: 0.0}
    table = {(x, y, z): 10.0 + b[x] + c[y] + e[z] for x in b for y in c for z in e}
    calls = []
    cfg, report = autotune.descend(_table_measure(table, calls), autotune.LAUNCH_DEFAULT,
                                   autotune.LAUNCH_CANDIDATES)
    assert cfg == {"blocks_per_cu": 3, "claim_factor": 2, "epa_blocks_per_cu": 12}
    assert report["final"]["time"] == min(table.values())
    # start + 2 + 3 + 3 candidates + final
    assert len(calls) == 1 + 8 + 1
    assert report["sweep0_claim_factor"] == {"1": 10.3 + 0.05, "2": 10.05, "4": 10.15}


def test_descend_keeps_the_start_when_nothing_is_faster():
    calls = []
    measure = lambda cfg: calls.append(cfg) or 1.0
    cfg, _ = autotune.descend(measure, autotune.LAUNCH_DEFAULT, autotune.LAUNCH_CANDIDATES)
    # all equal: the first candidate of each key wins the tie, in candidate order
    assert cfg == {k: v[0] for k, v in autotune.LAUNCH_CANDIDATES.items()}


def test_descend_second_sweep_follows_an_interaction():
    # claim_factor 4 only pays once epa_blocks_per_cu is 12
    def measure(cfg):
        t = 5.0 - (0.5 if cfg["epa_blocks_per_cu"] == 12 else 0.0)
        if cfg["claim_factor"] == 4:
            t += -0.3 if cfg["epa_blocks_per_cu"] == 12 else 0.2
        return t
    one, _ = autotune.descend(measure, autotune.LAUNCH_DEFAULT, autotune.LAUNCH_CANDIDATES)
    two, _ = autotune.descend(measure, autotune.LAUNCH_DEFAULT, autotune.LAUNCH_CANDIDATES,
                              sweeps=2)
    assert one["claim_factor"] != 4 and one["epa_blocks_per_cu"] == 12
    assert two["claim_factor"] == 4 and two["epa_blocks_per_cu"] == 12


def test_tuned_tables_read_back(tmp_path):
    p = tmp_path / "tuning.json"
    assert autotune.tuned_lanes(str(p)) is None and autotune.tuned_launch(str(p)) is None
    p.write_text(json.dumps({"lanes_per_pair": [1, 4, 8, 16, 32, 64, 64, 64],
                             "launch": {"blocks_per_cu": 2, "claim_factor": 4,
                                        "epa_blocks_per_cu": 8},
                             "serving_lanes_per_pair": [4, 8, 64, 32, 64, 32, 64, 64]}))
    # an 8-entry (per size class) table serves pairs of one and of two hulls
    assert autotune.tuned_lanes(str(p)) == [1, 4, 8, 16, 32, 64, 64, 64] * 2
    assert autotune.tuned_lanes(str(p), kind="latency") == [4, 8, 64, 32, 64, 32, 64, 64] * 2
    assert autotune.tuned_launch(str(p)) == {"blocks_per_cu": 2, "claim_factor": 4,
                                             "epa_blocks_per_cu": 8}
    assert autotune.tuned_launch(str(p), kind="latency") is None


def test_work_classes_match_the_device_rule():
    n = np.array([0, 1, 32, 33, 128, 129, 512, 513, 1024, 1025, 2048, 2049, 4096, 4097, 10 ** 6])
    assert list(autotune.work_class(n)) == [0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7]


def test_pair_classes_split_pairs_of_two_hulls():
    from distance3d_amd import _native
    from distance3d_amd.scene import SceneBuilder
    b = SceneBuilder()
    rs = np.random.RandomState(0)
    h40 = b.add_hull(rs.randn(40, 3))
    h600 = b.add_hull(rs.randn(600, 3))
    sph = b.add_sphere(np.zeros(3), 1.0)
    scene = b.build()
    pa = np.array([sph, h40, h40, h600, h40], dtype=np.int32)
    pb = np.array([sph, sph, h40, sph, h600], dtype=np.int32)
    # sizes 0, 40, 80, 600, 640 vertices -> size classes 0, 2, 2, 4, 4
    assert list(autotune.pair_classes(scene, pa, pb)) == [0, 2, 10, 4, 12]
    assert _native.expand_lanes(range(16)) == list(range(16))
    assert _native.expand_lanes([1, 2, 3, 4, 5, 6, 7, 8]) == [1, 2, 3, 4, 5, 6, 7, 8] * 2
    try:
        _native.expand_lanes([1, 2, 3])
    except ValueError:
        pass
    else:
        raise AssertionError("a 3-entry table was accepted")
