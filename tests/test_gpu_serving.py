"""Continuous-batching server: every query's result is the same alone and in
the mix, admission by queue length and by device memory, cancellation, and a
failed batch leaving its slot usable."""
import sys

import numpy as np
import pytest

import graphs
from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd import _lib as L
from enterprise_amd.bfs import row_hash_host
from enterprise_amd.serving import BFSServer, Rejected, run_load

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def two_graphs():
    a = HostGraph.rmat(14, 16, seed=41)
    b = HostGraph.rmat(12, 8, seed=42)
    da, db = DeviceGraph(a), DeviceGraph(b)
    yield {"big": (a, da), "small": (b, db)}
    da.free()
    db.free()


def test_alone_equals_mix(oracle, two_graphs):
    srv = BFSServer(max_batch=64, max_levels=40, max_slots=3)
    for name, (hg, dg) in two_graphs.items():
        srv.add_graph(name, dg, copy_rows=True)
    srv.start()
    try:
        roots = {k: hg.pick_roots(200, 7) for k, (hg, _) in two_graphs.items()}
        _say("[gpu] serving burst + Poisson mix")
        burst = [srv.submit("big" if i % 5 else "small", roots["big" if i % 5 else "small"][i % 200])
                 for i in range(300)]
        mixed = run_load(srv, 3000, 0.2, 11, roots, {"big": 0.8, "small": 0.2})
        for q in burst + mixed:
            q.wait(120)
        alone = {}
        for q in burst + mixed:
            hg, dg = two_graphs[q.graph]
            key = (q.graph, q.root)
            if key not in alone:
                alone[key] = dg.bfs(q.root)[0]
            assert np.array_equal(q.depth, alone[key]), (q.graph, q.root, q.batch_size)
            assert q.row_hash == int(row_hash_host(alone[key]))
        sizes = srv.lanes["big"].stats["batch_sizes"]
        assert max(sizes) > srv.plain_max  # the multi-source path served some batches
        # Sampled queries against the CPU oracle, among them queries served by
        # multi-source batches.
        multi = [q for q in burst + mixed if q.batch_size > srv.plain_max]
        sample = multi[:3] + [q for q in burst + mixed if q.graph == "small"][:3]
        assert len(sample) == 6
        for q in sample:
            hg, _ = two_graphs[q.graph]
            assert np.array_equal(q.depth, oracle.bfs(hg.beg_pos, hg.adj, q.root)), (q.graph, q.root)
    finally:
        srv.close()


def test_small_batches_use_single_source_path(two_graphs):
    """A lone query runs as a single-source search and equals the result
    alone; a burst larger than plain_max runs as one multi-source batch."""
    hg, dg = two_graphs["big"]
    srv = BFSServer(max_batch=64, max_levels=40, max_slots=1, plain_max=2)
    srv.add_graph("big", dg, copy_rows=True)
    roots = hg.pick_roots(40, 8)
    try:
        qs = [srv.submit("big", int(r)) for r in roots[:2]]
        srv.start()
        for q in qs:
            q.wait(60)
        st = srv.lanes["big"].stats
        assert st["plain_batches"] == 1 and st["batch_sizes"] == [2]
        srv.stop()
        qs += [srv.submit("big", int(r)) for r in roots[2:]]
        srv.start()
        for q in qs:
            q.wait(60)
            assert np.array_equal(q.depth, dg.bfs(q.root)[0]), q.root
        assert st["batch_sizes"][1] == 38 and st["plain_batches"] == 1
    finally:
        srv.close()


def test_queue_admission_and_cancel(two_graphs):
    hg, dg = two_graphs["small"]
    srv = BFSServer(max_batch=8, max_queue=10, max_slots=1)
    srv.add_graph("small", dg)
    try:
        qs = [srv.submit("small", 1) for _ in range(10)]   # not started: stays queued
        with pytest.raises(Rejected):
            srv.submit("small", 1)
        assert srv.cancel_pending("small") == 10
        for q in qs:
            with pytest.raises(L.EntError):
                q.wait(1)
        srv.start()
        q = srv.submit("small", int(hg.pick_roots(1, 3)[0])).wait(60)
        assert q.row_hash is not None
    finally:
        srv.close()


def test_memory_budget_admission(two_graphs):
    hg, dg = two_graphs["big"]
    srv = BFSServer(max_batch=64, max_levels=32)
    per = srv.slot_bytes(dg)
    srv.mem_budget = per // 2
    with pytest.raises(L.EntError) as e:
        srv.add_graph("big", dg)
    assert e.value.code == L.ENT_ENOMEM
    srv.mem_budget = int(per * 2.5)
    assert srv.add_graph("big", dg) == 2       # two slots fit the budget
    srv.close()


def test_mixed_slot_capacities(oracle, two_graphs):
    """small_batch: the first slot takes batches of up to max_batch, the
    others up to small_batch; the budget is charged for each size. The
    smaller slots take work only while the whole queue fits in them, so a
    burst queued before start goes to the large slot in batches beyond
    small_batch; a trickle afterwards is served too. Every row equals the
    single-source result (sampled rows equal the oracle)."""
    hg, dg = two_graphs["big"]
    probe = BFSServer(max_batch=128, max_levels=12, small_batch=16)
    big, small = probe.slot_bytes(dg), probe.slot_bytes(dg, 16)
    assert small < big
    srv = BFSServer(mem_budget_bytes=big + 2 * small + small // 2, max_batch=128, max_levels=12,
                    max_slots=4, plain_max=4, small_batch=16)
    assert srv.add_graph("g", dg, copy_rows=True) == 3
    assert [s.capacity for s in srv.lanes["g"].slots] == [128, 16, 16]
    try:
        roots = hg.pick_roots(300, 17)
        qs = [srv.submit("g", int(r)) for r in roots]   # queued before start
        _say("[gpu] serving: mixed slot capacities")
        srv.start()
        for q in qs:
            q.wait(120)
        sizes = srv.lanes["g"].stats["batch_sizes"]
        assert sum(sizes) == len(qs) and max(sizes) > 16 and max(sizes) <= 128, sizes
        for r in roots[:20]:                      # a trickle: one query at a time
            q = srv.submit("g", int(r))
            q.wait(60)
            qs.append(q)
        for i, q in enumerate(qs):
            assert q.error is None
            want, _ = dg.bfs(q.root)
            assert np.array_equal(q.depth, want), q.root
            if i % 60 == 0:
                assert np.array_equal(q.depth, oracle.bfs(hg.beg_pos, hg.adj, q.root)), q.root
    finally:
        srv.close()


def test_deep_query_in_batch(oracle):
    """One query deeper than the slot's reserved levels (the end of a
    400-vertex path) served in one 64-query multi-source batch with 63
    shallow ones (leaves of a star): the batch flushes its kept levels to the
    rows as it goes, and every query gets its correct depths."""
    n1, n2 = 400, 3000
    e = [(i, i + 1) for i in range(n1 - 1)] + [(n1, n1 + j) for j in range(1, n2)]
    g = graphs.edges_graph(n1 + n2, e)
    with DeviceGraph(g) as dg:
        srv = BFSServer(max_batch=64, max_levels=24, max_slots=1, plain_max=0)
        srv.add_graph("g", dg, copy_rows=True)
        try:
            roots = [n1 + 1 + i for i in range(63)]
            roots.insert(17, 0)
            qs = [srv.submit("g", r) for r in roots]   # queued before start: one batch
            _say("[gpu] serving: deep query in a 64-query batch")
            srv.start()
            for q in qs:
                q.wait(120)
            st = srv.lanes["g"].stats
            assert st["batch_sizes"] == [64] and st["flushes"] > 0, st
            for q in qs:
                want = oracle.bfs(g.beg_pos, g.adj, q.root)
                assert np.array_equal(q.depth, want), q.root
                assert q.row_hash == int(row_hash_host(want))
        finally:
            srv.close()


def test_failed_batch_keeps_slot_usable(monkeypatch):
    """A batch whose search fails passes the error to each of its queries,
    and the same slot then serves the next batch correctly; a root outside
    the graph is refused at submit and joins no batch."""
    h = HostGraph.rmat(10, 8, seed=5)
    with DeviceGraph(h) as dh:
        r = int(h.pick_roots(1, 2)[0])
        srv = BFSServer(max_batch=4, max_levels=32, max_slots=1, plain_max=0)
        srv.add_graph("r", dh, copy_rows=True)
        slot_dg = srv.lanes["r"].slots[0].dg
        real = slot_dg.bfs_batch_device
        calls = {"n": 0}

        def failing_once(roots, out):
            calls["n"] += 1
            if calls["n"] == 1:
                raise L.EntError(L.ENT_EDEVICE, "injected failure")
            return real(roots, out)

        monkeypatch.setattr(slot_dg, "bfs_batch_device", failing_once)
        with pytest.raises(ValueError):
            srv.submit("r", h.n)
        srv.start()
        try:
            with pytest.raises(L.EntError):
                srv.submit("r", r).wait(60)
            q = srv.submit("r", r).wait(60)
            assert np.array_equal(q.depth, dh.bfs(r)[0])
            assert calls["n"] == 2
        finally:
            srv.close()
