"""Continuous-batching collision server over a device-resident scene.

The scene (collider table and hull vertex pools) and a table of poses stay in
HBM.  A request is a list of pairs (collider a, collider b, pose row for a,
pose row for b); for every pair the server returns the GJK distance, closest
points and status, plus EPA depth, normal and contact points when the posed
colliders intersect.

Scheduling:
* ``n_slots`` batch slots, each with its own stream and preallocated device
  buffers for ``slot_pairs`` pairs; together they are the memory budget
  (:meth:`CollisionServer.slot_bytes` bytes per slot).
* Whenever a slot is free and requests are waiting, the slot is filled in
  arrival order with every waiting request that still fits and launched
  immediately -- requests that arrive while other batches run are admitted as
  soon as a slot frees up, and batches on different slots overlap on the GPU.
* A batch is padded to the next size bucket and runs as a captured HIP graph
  of that bucket (copy-in, pose instancing, GJK, EPA, copy-out), so a launch
  costs one graph replay.
* Results of a pair do not depend on the other pairs of its batch: every
  pair is computed independently, in FP64, by the same kernels.
* Each slot's stream needs a hardware queue of its own: HIP places streams
  round-robin on ``GPU_MAX_HW_QUEUES`` queues (4 by default), and two slots on
  one queue run their batches one after the other.  Set it to at least
  ``n_slots + 2`` in the environment before the process first uses the GPU
  (the server warns when it finds a smaller value).
"""
import collections
import gc
import os
import time
import warnings

import numpy as np

from . import kernels


class Request:
    """Pairs to evaluate: collider ids ``pair_a``/``pair_b`` and pose rows
    ``pose_a``/``pose_b`` (indices into the server's pose table)."""

    __slots__ = ("rid", "arrival", "pair_a", "pair_b", "pose_a", "pose_b")

    def __init__(self, rid, arrival, pair_a, pair_b, pose_a, pose_b):
        self.rid = rid
        self.arrival = float(arrival)
        self.pair_a = np.ascontiguousarray(pair_a, dtype=np.int32)
        self.pair_b = np.ascontiguousarray(pair_b, dtype=np.int32)
        self.pose_a = np.ascontiguousarray(pose_a, dtype=np.int32)
        self.pose_b = np.ascontiguousarray(pose_b, dtype=np.int32)

    @property
    def n_pairs(self):
        return len(self.pair_a)


class Result:
    """Per-pair outputs of one request (host arrays)."""

    __slots__ = ("rid", "dist", "cp_a", "cp_b", "status", "epa", "epa_status", "latency")

    def __init__(self, rid, dist, cp_a, cp_b, status, epa, epa_status, latency):
        self.rid = rid
        self.dist = dist
        self.cp_a = cp_a
        self.cp_b = cp_b
        self.status = status
        self.epa = epa
        self.epa_status = epa_status
        self.latency = latency

    @property
    def depth(self):
        return self.epa[:, 0]

    @property
    def normal(self):
        return self.epa[:, 1:4]

    def same_as(self, other):
        """Bitwise equality of every output array."""
        return all(np.array_equal(getattr(self, k), getattr(other, k), equal_nan=True)
                   for k in ("dist", "cp_a", "cp_b", "status", "epa", "epa_status"))


BUCKETS = (64, 256, 1024, 4096, 16384, 65536)


class GpuSlot:
    """Device buffers, pinned staging, stream and captured graphs for one
    in-flight batch of up to ``capacity`` pairs."""

    def __init__(self, server, capacity, use_graphs=True):
        import torch
        self.server = server
        self.capacity = capacity
        dev = server.scene.device
        self.stream = torch.cuda.Stream(device=dev)
        self.event = torch.cuda.Event()
        # inputs of a batch padded to bucket b sit contiguously in the first
        # 4b entries as [pair_a | pair_b | pose_a | pose_b], so the captured
        # copy-in is one contiguous pinned -> device range (a strided host
        # slice would be staged through a temporary that a graph captures
        # once and never refreshes)
        self.d_in = torch.empty(4 * capacity, dtype=torch.int32, device=dev)
        self.h_in = torch.empty(4 * capacity, dtype=torch.int32, pin_memory=True)
        self.inst = (torch.empty((2 * capacity, 16), dtype=torch.float64, device=dev),
                     torch.empty((2 * capacity, 4), dtype=torch.int32, device=dev),
                     torch.empty(capacity, dtype=torch.int32, device=dev),
                     torch.empty(capacity, dtype=torch.int32, device=dev))
        self.out = kernels.GjkEpaResult.empty(capacity, dev)
        self.ws = torch.empty(kernels.gjk_epa_workspace_bytes(capacity), dtype=torch.uint8,
                              device=dev)
        self.h_dist = torch.empty(capacity, dtype=torch.float64, pin_memory=True)
        self.h_cp = torch.empty((2, capacity, 3), dtype=torch.float64, pin_memory=True)
        self.h_status = torch.empty((2, capacity), dtype=torch.int32, pin_memory=True)
        self.h_epa = torch.empty((capacity, 13), dtype=torch.float64, pin_memory=True)
        self.use_graphs = use_graphs
        self.graphs = {}
        self.batch = []
        self.n_used = 0
        self.bucket = 0

    def _pipeline(self, n):
        srv = self.server
        self.d_in[:4 * n].copy_(self.h_in[:4 * n], non_blocking=True)
        d = self.d_in
        f64, i32, ia, ib = self.inst
        view, pa, pb = kernels.instance_pairs(
            srv.scene, d[0:n], d[n:2 * n], srv.poses, srv.poses,
            d[2 * n:3 * n], d[3 * n:4 * n], srv.n_poses,
            out=(f64[:2 * n], i32[:2 * n], ia[:n], ib[:n]))
        o = self.out
        res = kernels.GjkEpaResult(o.dist[:n], o.cp_a[:n], o.cp_b[:n], o.status[:n],
                                   o.epa_out[:n], o.epa_status[:n])
        kernels.gjk_epa_batch(view, pa, pb, out=res, workspace=self.ws,
                              lanes_per_pair=srv.lanes_per_pair, launch=srv.launch)
        self.h_dist[:n].copy_(o.dist[:n], non_blocking=True)
        self.h_cp[0, :n].copy_(o.cp_a[:n], non_blocking=True)
        self.h_cp[1, :n].copy_(o.cp_b[:n], non_blocking=True)
        self.h_status[0, :n].copy_(o.status[:n], non_blocking=True)
        self.h_status[1, :n].copy_(o.epa_status[:n], non_blocking=True)
        self.h_epa[:n].copy_(o.epa_out[:n], non_blocking=True)

    def prepare(self):
        """Warm up and capture one graph per bucket size."""
        import torch
        for b in BUCKETS:
            if b > self.capacity:
                break
            self._pad(b, 0)
            with torch.cuda.stream(self.stream):
                self._pipeline(b)
            self.stream.synchronize()
            if self.use_graphs:
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.stream):
                    self._pipeline(b)
                self.graphs[b] = g
        self.stream.synchronize()

    def _pad(self, bucket, start):
        """Fill entries [start, bucket) of the bucket layout with the dummy pair."""
        if bucket > start:
            h = self.h_in.numpy()
            for c, v in enumerate(self.server.dummy):
                h[c * bucket + start:(c + 1) * bucket] = v

    def load(self, requests):
        total = sum(r.n_pairs for r in requests)
        self.bucket = min(next((b for b in BUCKETS if b >= total), self.capacity),
                          self.capacity)
        b = self.bucket
        h = self.h_in.numpy()
        off = 0
        self.batch = []
        for r in requests:
            k = r.n_pairs
            h[off:off + k] = r.pair_a
            h[b + off:b + off + k] = r.pair_b
            h[2 * b + off:2 * b + off + k] = r.pose_a
            h[3 * b + off:3 * b + off + k] = r.pose_b
            self.batch.append((r, off, k))
            off += k
        self.n_used = off
        self._pad(b, off)

    def launch(self):
        import torch
        with torch.cuda.stream(self.stream):
            g = self.graphs.get(self.bucket)
            if g is not None:
                g.replay()
            else:
                self._pipeline(self.bucket)
            self.event.record(self.stream)

    def done(self):
        return self.event.query()

    def results(self, now, arena, start):
        """Copies the batch's outputs into ``arena`` rows [start, start + n) (the
        batch's requests are consecutive there) and returns per-request
        Results that are views of the arena."""
        n = self.n_used
        dist, cpa, cpb, st, epa, est = arena
        dist[start:start + n] = self.h_dist.numpy()[:n]
        cpa[start:start + n] = self.h_cp.numpy()[0, :n]
        cpb[start:start + n] = self.h_cp.numpy()[1, :n]
        st[start:start + n] = self.h_status.numpy()[0, :n]
        est[start:start + n] = self.h_status.numpy()[1, :n]
        epa[start:start + n] = self.h_epa.numpy()[:n]
        out = []
        for r, off, k in self.batch:
            s = slice(start + off, start + off + k)
            out.append(Result(r.rid, dist[s], cpa[s], cpb[s], st[s], epa[s], est[s],
                              now - r.arrival))
        return out


class CollisionServer:
    """Continuous batching over a device-resident scene and pose table."""

    def __init__(self, scene, poses, n_slots=4, slot_pairs=16384, lanes_per_pair=None,
                 use_graphs=True, slot_factory=None, launch=None):
        """``scene``: DeviceScene; ``poses``: (P, 12) float64 pose rows (R row
        major, t), kept on the device with two extra rows used to pad
        batches (a pair placed 1e4 apart, clipped in one GJK iteration).
        ``lanes_per_pair``: per work class; None takes the autotuner's
        latency table.  ``launch``: launch shape (see kernels.gjk_epa_batch);
        None takes the autotuner's latency table, else the library default."""
        import torch
        self.scene = scene
        P = np.asarray(poses, dtype=np.float64).reshape(-1, 12)
        far = np.zeros((2, 12))
        far[:, [0, 4, 8]] = 1.0
        far[1, 9] = 1.0e4
        self.n_poses = len(P) + 2
        self.poses = torch.from_numpy(np.vstack([P, far])).to(scene.device)
        i32 = scene.host.shape_i32
        cost = np.where(i32[:, 0] == 0, i32[:, 1], 0)
        k = int(np.argmin(cost))
        self.dummy = (k, k, len(P), len(P) + 1)
        if lanes_per_pair is None:
            # the autotuner's latency table (lanes per pair that minimise the
            # time of one small batch), else the library defaults
            from . import autotune
            lanes_per_pair = autotune.tuned_lanes(kind="latency")
        if launch is None:
            from . import autotune
            launch = autotune.tuned_launch(kind="latency")
        self.lanes_per_pair = lanes_per_pair
        self.launch = launch
        queues = os.environ.get("GPU_MAX_HW_QUEUES")
        if slot_factory is None and (queues is None or int(queues) < n_slots + 2):
            warnings.warn("GPU_MAX_HW_QUEUES=%s: fewer hardware queues than the %d slot streams "
                          "need (n_slots + 2); slots will share queues and serialize"
                          % (queues or "4 (default)", n_slots))
        factory = slot_factory or (lambda srv, cap: GpuSlot(srv, cap, use_graphs))
        self.slots = [factory(self, slot_pairs) for _ in range(n_slots)]
        for s in self.slots:
            s.prepare()

    @staticmethod
    def slot_bytes(slot_pairs):
        """Device bytes one slot holds (inputs, instances, outputs, workspace)."""
        per_pair = 4 * 4 + 2 * (16 * 8 + 4 * 4) + 2 * 4 + (8 + 48 + 4 + 13 * 8 + 4)
        return slot_pairs * per_pair + kernels.gjk_epa_workspace_bytes(slot_pairs)

    @classmethod
    def slot_pairs_for(cls, budget_bytes, n_slots=4, granule=256, max_pairs=1 << 24):
        """The largest slot size (pairs, a multiple of ``granule``) whose
        ``n_slots`` slots fit in ``budget_bytes``; 0 when not even one granule
        fits."""
        lo, hi = 0, max_pairs // granule
        while lo < hi:
            mid = (lo + hi + 1) // 2
            if n_slots * cls.slot_bytes(mid * granule) <= budget_bytes:
                lo = mid
            else:
                hi = mid - 1
        return lo * granule

    @classmethod
    def with_memory_fraction(cls, scene, poses, fraction=0.5, n_slots=4, **kw):
        """A server whose slots take ``fraction`` of the device memory free now:
        the memory budget that decides which requests a slot admits, sized
        from the device instead of a fixed pair count."""
        if not 0.0 < fraction <= 1.0:
            raise ValueError("fraction must be in (0, 1], got %r" % (fraction,))
        import torch
        free, _ = torch.cuda.mem_get_info(scene.device)
        pairs = cls.slot_pairs_for(int(free * fraction), n_slots)
        if pairs == 0:
            raise MemoryError("%.0f%% of the free device memory (%d bytes) holds no slot"
                              % (100 * fraction, free))
        return cls(scene, poses, n_slots=n_slots, slot_pairs=pairs, **kw)

    def serve(self, requests, realtime=True, clock=time.perf_counter):
        """Serve requests sorted by ``arrival`` (seconds from the start).
        Returns (results by request id, stats)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            if r.n_pairs > self.slots[0].capacity:
                raise ValueError("request %r has %d pairs, more than a slot holds (%d)"
                                 % (r.rid, r.n_pairs, self.slots[0].capacity))
        # one result arena for the whole stream, rows in arrival order, touched
        # before the clock starts so that no page fault lands in a batch
        total = int(sum(r.n_pairs for r in reqs))
        arena = (np.zeros(total), np.zeros((total, 3)), np.zeros((total, 3)),
                 np.zeros(total, np.int32), np.zeros((total, 13)), np.zeros(total, np.int32))
        row = 0
        first_row = {}
        for r in reqs:
            first_row[r.rid] = row
            row += r.n_pairs
        # a full garbage collection while batches are in flight stalls every
        # slot at once (tens of ms with the stream's result objects alive), so
        # the collector is paused for the duration of the stream
        gc_was_enabled = gc.isenabled()
        gc.disable()
        try:
            return self._serve(reqs, arena, first_row, realtime, clock)
        finally:
            if gc_was_enabled:
                gc.enable()

    def _serve(self, reqs, arena, first_row, realtime, clock):
        pending = collections.deque()
        free = collections.deque(self.slots)
        busy = collections.deque()
        results = {}
        i = 0
        n = len(reqs)
        t0 = clock()
        batches = 0
        batch_pairs = []
        while len(results) < n:
            now = clock() - t0 if realtime else float("inf")
            while i < n and reqs[i].arrival <= now:
                pending.append(reqs[i])
                i += 1
            for s in [b for b in busy if b.done()]:
                busy.remove(s)
                start = first_row[s.batch[0][0].rid]
                for res in s.results(clock() - t0 if realtime else 0.0, arena, start):
                    results[res.rid] = res
                free.append(s)
            if free and pending:
                s = free.popleft()
                cap = s.capacity
                take = []
                used = 0
                while pending and used + pending[0].n_pairs <= cap:
                    r = pending.popleft()
                    take.append(r)
                    used += r.n_pairs
                s.load(take)
                s.launch()
                busy.append(s)
                batches += 1
                batch_pairs.append(used)
        makespan = clock() - t0
        lat = np.array([results[r.rid].latency for r in reqs])
        sizes = np.array([r.n_pairs for r in reqs])
        big = sizes >= 2048
        stats = {
            "requests": n,
            "pairs": int(sizes.sum()),
            "makespan_s": makespan,
            "throughput_pairs_per_s": float(sizes.sum() / makespan) if makespan > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "p99_ms_ge2048_pairs": float(np.percentile(lat[big], 99) * 1e3) if big.any() else None,
            "batches": batches,
            "mean_batch_pairs": float(np.mean(batch_pairs)) if batch_pairs else 0.0,
        }
        return results, stats
