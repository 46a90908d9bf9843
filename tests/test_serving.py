"""Serving loop: admission / batching logic on the CPU with a stand-in slot,
and on the GPU the alone-vs-in-the-mix equality and oracle parity of posed
pairs."""
import numpy as np
import pytest

from distance3d_amd import serving, workloads
from oracle import oracle


class _FakeSlot:
    """Stand-in slot: outputs are a deterministic function of each pair and
    complete after ``delay`` polls; records what it was loaded with."""

    log = []

    def __init__(self, server, capacity, delay=2):
        self.server = server
        self.capacity = capacity
        self.delay = delay
        self.batch = []

    def prepare(self):
        pass

    def load(self, requests):
        assert sum(r.n_pairs for r in requests) <= self.capacity
        self.batch = []
        off = 0
        for r in requests:
            self.batch.append((r, off, r.n_pairs))
            off += r.n_pairs
        self.n_used = off
        _FakeSlot.log.append([r.rid for r in requests])

    def launch(self):
        self.polls = 0

    def done(self):
        self.polls += 1
        return self.polls > self.delay

    def results(self, now, arena, start):
        dist = arena[0]
        out = []
        for r, off, k in self.batch:
            s = slice(start + off, start + off + k)
            dist[s] = (r.pair_a * 1000 + r.pair_b).astype(float)
            out.append(serving.Result(r.rid, dist[s], arena[1][s], arena[2][s], arena[3][s],
                                      arena[4][s], arena[5][s], now - r.arrival))
        return out


class _FakeServer(serving.CollisionServer):
    def __init__(self, n_slots, slot_pairs):
        self.slots = [_FakeSlot(self, slot_pairs) for _ in range(n_slots)]


def _reqs(rs, n, max_pairs):
    out = []
    for k in range(n):
        m = rs.randint(1, max_pairs + 1)
        out.append(serving.Request(k, 0.0, rs.randint(0, 50, m), rs.randint(0, 50, m),
                                   np.zeros(m), np.zeros(m)))
    return out


def test_admission_fifo_capacity_and_completeness():
    rs = np.random.RandomState(0)
    reqs = _reqs(rs, 200, 300)
    _FakeSlot.log = []
    srv = _FakeServer(n_slots=3, slot_pairs=700)
    res, stats = srv.serve(reqs, realtime=False)
    assert sorted(res) == list(range(200))
    order = [rid for batch in _FakeSlot.log for rid in batch]
    assert order == sorted(order)  # arrival (FIFO) order, nothing skipped
    assert stats["batches"] == len(_FakeSlot.log)
    for r in reqs:
        assert np.array_equal(res[r.rid].dist, (r.pair_a * 1000 + r.pair_b).astype(float))


def test_oversized_request_rejected():
    srv = _FakeServer(n_slots=1, slot_pairs=10)
    big = serving.Request(0, 0.0, np.zeros(11), np.zeros(11), np.zeros(11), np.zeros(11))
    with pytest.raises(ValueError):
        srv.serve([big], realtime=False)


@pytest.mark.gpu
def test_alone_vs_mix_and_oracle(gpu):
    scene, poses, reqs = workloads.serving_workload(seed=11, n_colliders=96, n_requests=120,
                                                    max_pairs=512)
    dev = scene.to_device()
    srv = serving.CollisionServer(dev, poses, n_slots=2, slot_pairs=4096)
    mix, stats = srv.serve(reqs, realtime=False)
    assert stats["requests"] == len(reqs)
    for r in reqs[:12]:
        alone, _ = srv.serve([serving.Request(r.rid, 0.0, r.pair_a, r.pair_b, r.pose_a,
                                              r.pose_b)], realtime=False)
        assert alone[r.rid].same_as(mix[r.rid])
    # oracle parity of the posed pairs (GJK, then EPA on intersecting pairs)
    for r in reqs[:40]:
        host = workloads.posed_instances(scene, poses, r.pair_a, r.pair_b, r.pose_a, r.pose_b)
        k = r.n_pairs
        a, b = np.arange(0, 2 * k, 2), np.arange(1, 2 * k, 2)
        g = oracle.gjk_distance(host, a, b, want_simplex=True)
        out = mix[r.rid]
        assert np.array_equal(out.dist == 0.0, g["dist"] == 0.0)
        sep = g["dist"] > 1e-9
        assert np.all(np.abs(out.dist[sep] - g["dist"][sep]) <= 1e-5 * g["dist"][sep])
        hit = np.nonzero(g["dist"] == 0.0)[0]
        if len(hit):
            e = oracle.epa(host, a[hit], b[hit], g["simplex"][hit], g["n_points"][hit])
            assert np.array_equal(out.epa_status[hit], e["status"])
            assert np.all(np.abs(out.depth[hit] - e["depth"]) <= 1e-6 * np.maximum(1, e["depth"]))
        assert np.all(out.epa_status[g["dist"] > 0.0] == -1)


def test_slot_input_layout_is_contiguous_per_bucket():
    """A loaded batch occupies the first 4*bucket staging entries as
    [pair_a | pair_b | pose_a | pose_b] with dummy padding, so one contiguous
    copy of that range carries the whole batch."""
    import torch
    slot = object.__new__(serving.GpuSlot)
    slot.capacity = 4096
    slot.server = type("S", (), {"dummy": (7, 8, 9, 10)})()
    slot.h_in = torch.full((4 * 4096,), -1, dtype=torch.int32)
    r1 = serving.Request(0, 0.0, [1, 2, 3], [4, 5, 6], [11, 12, 13], [14, 15, 16])
    r2 = serving.Request(1, 0.0, [20], [21], [22], [23])
    slot.load([r1, r2])
    assert slot.bucket == 64 and slot.batch[1][1:] == (3, 1)
    h = slot.h_in.numpy()[:4 * 64].reshape(4, 64)
    assert list(h[:, :4].ravel()) == [1, 2, 3, 20, 4, 5, 6, 21, 11, 12, 13, 22, 14, 15, 16, 23]
    assert np.array_equal(h[:, 4:], np.array([7, 8, 9, 10])[:, None].repeat(60, 1))
    big = serving.Request(2, 0.0, np.arange(300), np.arange(300), np.zeros(300), np.zeros(300))
    slot.load([big])
    assert slot.bucket == 1024
    h = slot.h_in.numpy()[:4 * 1024].reshape(4, 1024)
    assert np.array_equal(h[0, :300], np.arange(300)) and np.all(h[1, 300:] == 8)


def test_slot_size_from_a_memory_budget():
    """Slots sized from a byte budget: the largest multiple of 256 pairs whose
    slots fit, monotone in the budget, and 0 when nothing fits."""
    S = serving.CollisionServer
    one = S.slot_bytes(256)
    assert S.slot_pairs_for(4 * one - 1, n_slots=4) == 0
    assert S.slot_pairs_for(4 * one, n_slots=4) == 256
    prev = 0
    for budget in (1 << 20, 1 << 24, 1 << 28, 1 << 32, 1 << 36):
        p = S.slot_pairs_for(budget, n_slots=4)
        assert p % 256 == 0 and p >= prev
        assert 4 * S.slot_bytes(p) <= budget
        if p < (1 << 24):
            assert 4 * S.slot_bytes(p + 256) > budget
        prev = p
    assert S.slot_pairs_for(1 << 40, n_slots=4) == 1 << 24
    for bad in (0.0, -0.5, 1.5, float("nan")):
        with pytest.raises(ValueError):
            S.with_memory_fraction(None, None, fraction=bad)


@pytest.mark.gpu
def test_server_sized_from_free_memory(gpu):
    """A server whose slots take a fraction of the free device memory admits
    and serves like one sized by hand: the same results pair for pair."""
    import torch
    scene, poses, reqs = workloads.serving_workload(seed=12, n_colliders=64, n_requests=40,
                                                    max_pairs=512)
    dev = scene.to_device()
    free, _ = torch.cuda.mem_get_info(dev.device)
    want = serving.CollisionServer.slot_bytes(8192) * 2
    srv = serving.CollisionServer.with_memory_fraction(dev, poses, fraction=want / free,
                                                       n_slots=2)
    assert 8192 - 256 <= srv.slots[0].capacity <= 8192
    ref = serving.CollisionServer(dev, poses, n_slots=2, slot_pairs=4096)
    got, _ = srv.serve(reqs, realtime=False)
    exp, _ = ref.serve(reqs, realtime=False)
    for r in reqs:
        assert got[r.rid].same_as(exp[r.rid])
