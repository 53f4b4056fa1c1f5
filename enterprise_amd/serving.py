"""Continuous-batching BFS serving over HBM-resident graphs.

No counterpart in the reference (one graph, one root at a time,
bfs_gpu_opt.cuh:370). Each registered graph is uploaded once. It gets a set of
*slots*: a slot is a second handle on the same resident CSR
(ent_graph_share), with its own HIP stream, multi-source workspace and
output buffer. The number of slots per graph is fixed at start-up by a
device-memory budget: each slot reserves max_batch depth rows plus the
multi-source workspace (per-level source masks and seen masks). So admission
is by depth/status-array memory.

Queries queue per graph. Whenever a slot is free, its worker takes up to
max_batch queued queries (FIFO) and runs them as one multi-source search
(ent_bfs_batch_device), or, when at most plain_max were waiting, as
single-source searches one after another (ent_bfs_device): a multi-source
level sweeps every vertex, which costs more than a few direction-optimizing
single-source searches when the batch is that small. A batch deeper than the
slot's reserved levels still runs as one batch: the library writes the kept
levels to the rows and reuses their slots (ent_batch_stats.flushes), so one
deep query does not fail or split the batch it is in. With small_batch, a
graph's first slot takes batches of up to max_batch and the others up to
small_batch, and a smaller slot takes work only while the whole queue fits
in it. Measured on R-MAT 26 + R-MAT 20: single-source searches alone
saturate at about 520 queries/s, four 64-query slots at about 1,150, and one
512-query slot with two 64-query slots (8 kept levels) at about 3,300;
plain_max = 12 gave the lowest latencies below saturation. Queries that arrive while other batches run are
admitted in the next free slot. Batches on different slots run concurrently
on the device. Each query's result is its depth row in the slot's device
buffer. The row's device checksum (ent_device_row_hash) is returned, and
optionally a host copy. Latency runs from submit() to the end of the batch
that served the query.
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time

import numpy as np

from . import _lib as L
from .bfs import DeviceGraph


class QueryResult:
    __slots__ = ("qid", "graph", "root", "t_submit", "t_start", "t_done", "row_hash", "depth",
                 "batch_id", "batch_size", "error", "_event")

    def __init__(self, qid, graph, root):
        self.qid, self.graph, self.root = qid, graph, root
        self.t_submit = time.perf_counter()
        self.t_start = self.t_done = None
        self.row_hash = None
        self.depth = None
        self.batch_id = None
        self.batch_size = 0
        self.error = None
        self._event = threading.Event()

    def wait(self, timeout=None):
        if not self._event.wait(timeout):
            raise TimeoutError(f"query {self.qid} not finished")
        if self.error:
            raise self.error
        return self

    @property
    def latency_ms(self):
        return None if self.t_done is None else (self.t_done - self.t_submit) * 1e3

    def done(self):
        return self._event.is_set()


class Rejected(RuntimeError):
    """Raised by submit() when the graph's queue is full (admission refused)."""


class _Slot:
    def __init__(self, owner: DeviceGraph, max_batch, max_levels, share=True):
        self.dg = DeviceGraph(share=owner) if share else owner
        self.shared = share
        self.capacity = max_batch  # most queries per batch on this slot
        self.dg.batch_reserve(max_batch, max_levels)
        self.out = self.dg.device_alloc(max_batch * owner.n * 4)

    def close(self):
        self.dg.device_free(self.out)
        if self.shared:
            self.dg.free()


class _GraphLane:
    def __init__(self, name, dg, slots, max_batch, max_queue, copy_rows):
        self.name, self.dg = name, dg
        self.slots = slots
        self.max_batch = max_batch
        self.max_queue = max_queue
        self.copy_rows = copy_rows
        self.queue = collections.deque()
        self.cv = threading.Condition()
        self.stats = {"batches": 0, "plain_batches": 0, "queries": 0, "reached_edges": 0, "alg_bytes": 0,
                      "busy_s": 0.0, "batch_sizes": [], "flushes": 0}


class BFSServer:
    """Continuous-batching server for BFS queries over resident graphs."""

    def __init__(self, mem_budget_bytes=None, max_batch=64, max_levels=32, max_queue=100000,
                 max_slots=4, plain_max=12, small_batch=None):
        if not 1 <= max_batch <= 512:
            raise ValueError("max_batch must be in [1, 512]")
        if small_batch is not None and not 1 <= small_batch <= max_batch:
            raise ValueError("small_batch must be in [1, max_batch]")
        self.mem_budget = mem_budget_bytes
        self.max_batch = max_batch
        # With small_batch, each graph's first slot takes batches of up to
        # max_batch (throughput under load) and the others up to small_batch,
        # so that more batches run side by side for the same memory.
        self.small_batch = small_batch
        self.max_levels = max_levels
        self.max_queue = max_queue
        self.max_slots = max_slots
        self.plain_max = plain_max
        self.lanes = {}
        self.threads = []
        self._stop = False
        self._qid = 0
        self._batch_id = 0
        self._lock = threading.Lock()

    # -- set-up ---------------------------------------------------------------
    def slot_bytes(self, dg: DeviceGraph, batch=None):
        batch = self.max_batch if batch is None else batch
        ws = L.lib().ent_batch_workspace_bytes(dg.handle, batch, self.max_levels)
        info = dg.info()
        # Output rows + multi-source workspace + the shared handle's own
        # single-source workspace (depth, bitmaps, marks, queues).
        per_handle = info["n"] * 4 + info["n"] + 3 * ((info["n"] + 63) // 64) * 8 + 2 * 64 * 4096 * 4
        return batch * info["n"] * 4 + int(ws) + per_handle

    def add_graph(self, name, dg: DeviceGraph, share_budget=1.0, copy_rows=False, slots=None):
        """Registers a resident graph. The number of slots follows from the
        memory budget (share_budget of it for this graph), capped by
        max_slots; at least one slot is required. With small_batch, the first
        slot holds max_batch rows and the others small_batch."""
        per = self.slot_bytes(dg)
        per_small = self.slot_bytes(dg, self.small_batch) if self.small_batch else per
        if slots is None:
            if self.mem_budget is None:
                slots = self.max_slots
            else:
                avail = self.mem_budget * share_budget
                slots = 0 if avail < per else int(1 + min(self.max_slots - 1, (avail - per) // per_small))
        if slots < 1:
            raise L.EntError(L.ENT_ENOMEM, f"graph {name}: one slot needs {per} bytes, budget "
                                           f"{self.mem_budget * share_budget:.0f}")
        caps = [self.max_batch] + [self.small_batch or self.max_batch] * (slots - 1)
        made = []
        try:
            for cap in caps:
                made.append(_Slot(dg, cap, self.max_levels, share=True))
        except Exception:
            for s in made:
                s.close()
            raise
        lane = _GraphLane(name, dg, made, self.max_batch, self.max_queue, copy_rows)
        self.lanes[name] = lane
        return slots

    def start(self):
        self._stop = False
        for lane in self.lanes.values():
            for slot in lane.slots:
                t = threading.Thread(target=self._worker, args=(lane, slot), daemon=True)
                t.start()
                self.threads.append(t)
        return self

    def stop(self, drain=True, timeout=60.0):
        if drain:
            t_end = time.perf_counter() + timeout
            for lane in self.lanes.values():
                with lane.cv:
                    while lane.queue and time.perf_counter() < t_end:
                        lane.cv.wait(0.01)
        self._stop = True
        for lane in self.lanes.values():
            with lane.cv:
                lane.cv.notify_all()
        for t in self.threads:
            t.join(timeout)
        self.threads = []

    def close(self):
        self.stop(drain=False)
        for lane in self.lanes.values():
            for s in lane.slots:
                s.close()
        self.lanes = {}

    # -- requests -------------------------------------------------------------
    def submit(self, graph, root):
        lane = self.lanes[graph]
        if not 0 <= int(root) < lane.dg.n:
            # Refused here, so that one bad root cannot fail the batch it
            # would have joined.
            raise ValueError(f"graph {graph}: root {root} outside [0, {lane.dg.n})")
        with self._lock:
            self._qid += 1
            q = QueryResult(self._qid, graph, int(root))
        with lane.cv:
            if len(lane.queue) >= lane.max_queue:
                raise Rejected(f"graph {graph}: {len(lane.queue)} queries queued")
            lane.queue.append(q)
            lane.cv.notify()
        return q

    def cancel_pending(self, graph):
        """Fails every query of `graph` that has not started yet."""
        lane = self.lanes[graph]
        with lane.cv:
            pend = list(lane.queue)
            lane.queue.clear()
        for q in pend:
            q.error = L.EntError(L.ENT_ECANCELED, "cancelled before it started")
            q._event.set()
        return len(pend)

    # -- worker ---------------------------------------------------------------
    def _worker(self, lane: _GraphLane, slot: _Slot):
        n = lane.dg.n
        while True:
            with lane.cv:
                # A smaller slot takes work only while the whole queue fits in
                # it: under load the queue goes to the largest slot in big
                # batches (several batches side by side slow each other down).
                while not self._stop and (not lane.queue or (slot.capacity < lane.max_batch
                                                              and len(lane.queue) > slot.capacity)):
                    lane.cv.wait(0.05)
                if self._stop and not lane.queue:
                    return
                batch = [lane.queue.popleft() for _ in range(min(slot.capacity, len(lane.queue)))]
            with self._lock:
                self._batch_id += 1
                bid = self._batch_id
            t0 = time.perf_counter()
            for q in batch:
                q.t_start = t0
                q.batch_id = bid
                q.batch_size = len(batch)
            roots = np.array([q.root for q in batch], dtype=np.int32)
            try:
                if len(batch) <= self.plain_max:
                    st = {"reached_edges": 0, "alg_bytes": 0}
                    for i, r in enumerate(roots):
                        one = slot.dg.bfs_device(int(r), slot.out + i * n * 4)
                        st["reached_edges"] += one["reached_edges"]
                        st["alg_bytes"] += one["alg_bytes"]
                else:
                    st = slot.dg.bfs_batch_device(roots, slot.out)
                hashes = slot.dg.row_hash(slot.out, len(batch), n)
                rows = None
                if lane.copy_rows:
                    rows = np.empty((len(batch), n), dtype=np.int32)
                    L.check(L.lib().ent_copy_to_host(slot.dg.handle, C.c_void_p(slot.out),
                                                     rows.ctypes.data_as(L.i32p), len(batch) * n))
                t1 = time.perf_counter()
                for i, q in enumerate(batch):
                    q.row_hash = int(hashes[i])
                    if rows is not None:
                        q.depth = rows[i]
                    q.t_done = t1
                with lane.cv:
                    s = lane.stats
                    s["batches"] += 1
                    s["plain_batches"] += int(len(batch) <= self.plain_max)
                    s["queries"] += len(batch)
                    s["reached_edges"] += st["reached_edges"]
                    s["alg_bytes"] += st["alg_bytes"]
                    s["busy_s"] += t1 - t0
                    s["batch_sizes"].append(len(batch))
                    s["flushes"] += int(st.get("flushes", 0))
            except Exception as e:  # the batch fails, the slot stays usable
                t1 = time.perf_counter()
                for q in batch:
                    q.error = e
                    q.t_done = t1
            for q in batch:
                q._event.set()


def poisson_arrivals(rate, duration_s, seed):
    """Seeded Poisson arrival times in [0, duration_s)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration_s:
            return np.array(out)
        out.append(t)


def run_load(server: BFSServer, rate, duration_s, seed, roots_by_graph, mix):
    """Submits a Poisson stream (rate q/s for duration_s) choosing graph g with
    probability mix[g] and a root uniformly from roots_by_graph[g]; waits for
    every query; returns the list of QueryResult."""
    rng = np.random.default_rng(seed + 1)
    times = poisson_arrivals(rate, duration_s, seed)
    names = list(mix)
    probs = np.array([mix[k] for k in names], dtype=float)
    probs /= probs.sum()
    picks = rng.choice(len(names), size=len(times), p=probs)
    res = []
    t_start = time.perf_counter()
    for t, gi in zip(times, picks):
        now = time.perf_counter() - t_start
        if t > now:
            time.sleep(t - now)
        g = names[gi]
        roots = roots_by_graph[g]
        res.append(server.submit(g, int(roots[rng.integers(0, len(roots))])))
    for q in res:
        q.wait(timeout=600)
    return res
