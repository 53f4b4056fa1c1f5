"""GPU parity of the single-source BFS (through the C ABI) against the CPU
oracle: bit-exact int32 depth arrays and exact reached-vertex / reached-edge
counts, for every direction policy and launch form."""
import sys

import numpy as np
import pytest

import graphs
from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd import _lib as L

pytestmark = pytest.mark.gpu

DIRS = {"auto": L.DIR_AUTO, "reference": L.DIR_REFERENCE, "topdown": L.DIR_TOPDOWN,
        "bottomup": L.DIR_BOTTOMUP}
LAUNCHES = {"stream": L.LAUNCH_STREAM, "levels": L.LAUNCH_LEVELS, "graph": L.LAUNCH_GRAPH,
            "persistent": L.LAUNCH_PERSISTENT}

# Thresholds small enough that a scale-16 R-MAT fills every class, including
# grid-class pieces.
TINY_CLASSES = dict(t_wave=4, t_block=24, t_grid=96, piece=64)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(oracle, g, dg, root, tag):
    want = oracle.bfs(g.beg_pos, g.adj, root)
    _say(f"[gpu] {tag} root={root}")
    got, st = dg.bfs(root)
    if not np.array_equal(got, want):
        bad = np.flatnonzero(got != want)
        raise AssertionError(f"{tag} root {root}: {bad.size} depth mismatches, first v={bad[0]} "
                             f"gpu={got[bad[0]]} oracle={want[bad[0]]}")
    edges, verts = oracle.report(g.beg_pos, want)
    assert st["reached_edges"] == edges, (tag, st, edges)
    assert st["reached_vertices"] == verts, (tag, st, verts)
    assert st["levels"] == int(want.max()) + 1, (tag, st)
    return st


def test_wave_primitives():
    from enterprise_amd.bfs import selftest_wave
    selftest_wave(0)


@pytest.fixture(scope="module")
def rmat16():
    g = HostGraph.rmat(16, 16, seed=16)
    dg = DeviceGraph(g)
    yield g, dg
    dg.free()


@pytest.mark.parametrize("launch", list(LAUNCHES))
@pytest.mark.parametrize("name", list(graphs.CLOSED_FORMS))
def test_closed_forms(oracle, name, launch):
    g, depth_fn = graphs.CLOSED_FORMS[name]()
    roots = sorted({0, g.n // 2, g.n - 1})
    with DeviceGraph(g) as dg:
        for dname, d in DIRS.items():
            dg.configure(direction=d, launch=LAUNCHES[launch])
            for r in roots:
                _check(oracle, g, dg, r, f"{name}/{dname}/{launch}")
                got, _ = dg.bfs(r)
                assert np.array_equal(got, np.asarray(depth_fn(r), dtype=np.int32)), (name, r)


@pytest.mark.parametrize("launch", list(LAUNCHES))
def test_rmat16_config1(oracle, rmat16, launch):
    g, dg = rmat16
    root = int(g.pick_roots(1, 1001)[0])
    for dname, d in DIRS.items():
        dg.configure(direction=d, launch=LAUNCHES[launch])
        _check(oracle, g, dg, root, f"rmat16/{dname}/{launch}")


@pytest.mark.parametrize("launch", list(LAUNCHES))
def test_rmat16_all_classes(oracle, rmat16, launch):
    g, dg = rmat16
    base = dg.config
    try:
        for dname, d in DIRS.items():
            dg.configure(direction=d, launch=LAUNCHES[launch], **TINY_CLASSES)
            for r in g.pick_roots(3, 77):
                _check(oracle, g, dg, int(r), f"rmat16-tiny/{dname}/{launch}")
    finally:
        dg.set_config(base)


def test_mega_hub_grid_class(oracle):
    g, _ = graphs.star(70000)
    with DeviceGraph(g) as dg:
        for launch in LAUNCHES.values():
            for d in DIRS.values():
                dg.configure(direction=d, launch=launch)
                for r in (0, 1, 69999):
                    _check(oracle, g, dg, r, f"star70000/{d}/{launch}")


@pytest.mark.parametrize("bs", [64, 128, 512, 1024])
def test_block_sizes(oracle, rmat16, bs):
    g, dg = rmat16
    base = dg.config
    try:
        for launch in LAUNCHES.values():
            dg.configure(block_size=bs, launch=launch, direction=L.DIR_AUTO)
            for r in g.pick_roots(2, 5):
                _check(oracle, g, dg, int(r), f"rmat16/bs{bs}/{launch}")
    finally:
        dg.set_config(base)


def test_isolated_root(oracle):
    g = graphs.edges_graph(10, [(1, 2), (2, 3)])
    with DeviceGraph(g) as dg:
        for launch in LAUNCHES.values():
            dg.configure(launch=launch)
            got, st = dg.bfs(0)
            assert got[0] == 0 and (got[1:] == -1).all()
            assert st["levels"] == 1 and st["reached_edges"] == 0


def test_many_roots_schedule_independent(oracle, rmat16):
    """Forced top-down, bottom-up, reference and auto schedules give the same
    depths for many roots (the depths do not depend on the schedule)."""
    g, dg = rmat16
    roots = g.pick_roots(12, 2024)
    base = dg.config
    try:
        for r in roots:
            want = oracle.bfs(g.beg_pos, g.adj, int(r))
            for d in DIRS.values():
                dg.configure(direction=d, launch=L.LAUNCH_STREAM)
                got, _ = dg.bfs(int(r))
                assert np.array_equal(got, want)
    finally:
        dg.set_config(base)


HUBS = [(6, 0), (10, 1), (13, 1), (13, 0)]


@pytest.mark.parametrize("launch", list(LAUNCHES))
@pytest.mark.parametrize("hub", HUBS, ids=[f"h{h}td{t}" for h, t in HUBS])
def test_hub_cache(oracle, rmat16, launch, hub):
    """LDS hub table: frontier-hub membership in bottom-up and deferred hub
    marks in top-down leave depths and counts bit-exact, with the default and
    the tiny class thresholds, under every direction policy."""
    g, dg = rmat16
    base = dg.config
    h, td = hub
    try:
        for classes in ({}, TINY_CLASSES):
            for dname, d in DIRS.items():
                dg.configure(direction=d, launch=LAUNCHES[launch], hub_cache=h, hub_td=td, **classes)
                for r in g.pick_roots(2, 91):
                    _check(oracle, g, dg, int(r), f"rmat16/hub{h}td{td}/{dname}/{launch}/{len(classes)}")
    finally:
        dg.set_config(base)


def test_hub_cache_star_and_closed_forms(oracle):
    """Hub roots and hub-only frontiers: a 70,000-leaf star (the centre is the
    only hub), and closed forms whose degrees fall below t_wave (empty table)."""
    g, _ = graphs.star(70000)
    with DeviceGraph(g) as dg:
        for launch in LAUNCHES.values():
            for d in DIRS.values():
                dg.configure(direction=d, launch=launch, hub_cache=8, hub_td=1)
                for r in (0, 1, 69999):
                    _check(oracle, g, dg, r, f"star70000/hub/{d}/{launch}")
    for name in ("hypercube10", "grid37x23", "disjoint", "complete40", "bintree11"):
        g, depth_fn = graphs.CLOSED_FORMS[name]()
        with DeviceGraph(g) as dg:
            for d in DIRS.values():
                dg.configure(direction=d, hub_cache=7, hub_td=1, t_wave=2)
                for r in sorted({0, g.n // 2, g.n - 1}):
                    st = _check(oracle, g, dg, r, f"{name}/hub/{d}")
                    got, _ = dg.bfs(r)
                    assert np.array_equal(got, np.asarray(depth_fn(r), dtype=np.int32)), (name, r, st)


def test_hub_cache_rejects_bad_sizes(oracle, rmat16):
    g, dg = rmat16
    base = dg.config
    try:
        for bad in (1, 5, 14):
            with pytest.raises(L.EntError):
                dg.configure(hub_cache=bad)
        with pytest.raises(L.EntError):
            dg.configure(hub_cache=10, hub_td=2)
        # The largest table with deferred marks fits the largest block.
        dg.configure(hub_cache=13, hub_td=1, block_size=1024)
        _check(oracle, g, dg, int(g.pick_roots(1, 3)[0]), "rmat16/hub13td1/bs1024")
    finally:
        dg.set_config(base)


def test_row_order_upload(oracle):
    """Default upload reorders each row by descending floor(log2(neighbour
    degree)), stably; ENT_UPLOAD_KEEP_ORDER keeps the caller's rows. Both give
    the oracle's depths and counts."""
    g = HostGraph.rmat(14, 16, seed=61)
    deg = np.diff(g.beg_pos)
    cls = np.where(deg > 1, np.floor(np.log2(np.maximum(deg, 1))), 0).astype(np.int64)
    with DeviceGraph(g) as dg, DeviceGraph(g, keep_order=True) as dk:
        b, a = dg.download()
        assert np.array_equal(b, g.beg_pos)
        bk, ak = dk.download()
        assert np.array_equal(ak, g.adj)
        for v in range(g.n):
            s, e = int(b[v]), int(b[v + 1])
            row, orig = a[s:e], g.adj[s:e]
            want = orig[np.argsort(-cls[orig], kind="stable")]
            assert np.array_equal(row, want), v
        for r in g.pick_roots(4, 17):
            for d in DIRS.values():
                dg.configure(direction=d)
                dk.configure(direction=d)
                _check(oracle, g, dg, int(r), f"ordered/{d}")
                _check(oracle, g, dk, int(r), f"kept/{d}")


def test_examined_counts(oracle, rmat16):
    """The counting kernel variant leaves depths unchanged. Forced top-down
    loads every adjacency entry of every reached vertex exactly once, so its
    examined edges equal the reached-edge count M_c; the early exits of the
    other schedules examine fewer."""
    g, dg = rmat16
    base = dg.config
    try:
        for r in g.pick_roots(3, 404):
            r = int(r)
            want = oracle.bfs(g.beg_pos, g.adj, r)
            for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT):
                dg.configure(direction=L.DIR_TOPDOWN, count_examined=1, launch=launch)
                st = _check(oracle, g, dg, r, f"count/td/{launch}")
                assert st["examined_edges"] == st["reached_edges"], st
                assert st["examined_bytes"] >= 8 * st["examined_edges"] + 8 * st["reached_vertices"]
            dg.configure(direction=L.DIR_AUTO, count_examined=1, launch=L.LAUNCH_STREAM)
            st = _check(oracle, g, dg, r, "count/auto")
            assert 0 < st["examined_edges"] < st["reached_edges"], st
            dg.configure(count_examined=0)
            got, st0 = dg.bfs(r)
            assert np.array_equal(got, want) and st0["examined_edges"] == -1
    finally:
        dg.set_config(base)


def test_allocation_failure_leaves_handle_usable(oracle, rmat16):
    """An allocation the device cannot satisfy fails with ENT_ENOMEM and a
    message; the handle keeps working (searches, multi-source workspace)."""
    g, dg = rmat16
    with pytest.raises(L.EntError) as e:
        dg.device_alloc(1 << 52)  # 4 PiB
    assert e.value.code == L.ENT_ENOMEM and "hipMalloc" in str(e.value)
    r = int(g.pick_roots(1, 8)[0])
    _check(oracle, g, dg, r, "after-enomem")
    out, _ = dg.bfs_batch(g.pick_roots(8, 9))
    assert out.shape == (8, g.n)


def test_level_trace(oracle, rmat16):
    """The counting variant records every launch of the search: levels never
    decrease, and each record's frontier vertex and degree counts equal the
    oracle's for its level. Without counting there are no records."""
    g, dg = rmat16
    base = dg.config
    deg = np.diff(g.beg_pos)
    try:
        # The hipGraph form replays 6 launches at a time, so a search spans
        # several replays whose launch arguments restart at 0.
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT, L.LAUNCH_GRAPH):
            for d in (L.DIR_AUTO, L.DIR_TOPDOWN):
                dg.configure(count_examined=1, launch=launch, direction=d, graph_levels=6)
                r = int(g.pick_roots(1, 55)[0])
                st = _check(oracle, g, dg, r, f"trace/{launch}/{d}")
                recs = dg.level_trace()
                lv = [(x["level"], x["phase"]) for x in recs]
                # Every launch but the last, which finds the frontier empty.
                assert len(recs) == st["phases"] - 1, (lv, st)
                assert [x["launch"] for x in recs] == list(range(len(recs)))
                lv = [x["level"] for x in recs]
                assert lv == sorted(lv) and lv[0] == 0 and lv[-1] == st["levels"] - 1, lv
                want = oracle.bfs(g.beg_pos, g.adj, r)
                for x in recs:
                    at = want == x["level"]
                    assert x["frontier_vertices"] == int(at.sum()), x
                    assert x["frontier_edges"] == int(deg[at].sum()), x
                if d == L.DIR_TOPDOWN:
                    assert "bottom-up" not in {x["phase"] for x in recs}
        dg.configure(count_examined=0)
        dg.bfs(int(g.pick_roots(1, 56)[0]), copy=False)
        assert dg.level_trace() == []
    finally:
        dg.set_config(base)


def test_results_final_on_return(oracle, rmat16):
    """ent_bfs_device returns once the search is finished, while launches
    queued behind it may still run: the depth array read at once through
    another stream (a synchronous hipMemcpy) is already the final one, and the
    device times are consistent."""
    import ctypes as C
    g, dg = rmat16
    base = dg.config
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    try:
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_LEVELS):
            dg.configure(launch=launch)
            for r in g.pick_roots(6, 31):
                r = int(r)
                st = dg.bfs_device(r)
                got = np.empty(g.n, dtype=np.int32)
                assert hip.hipMemcpy(got.ctypes.data, dg.depth_device_ptr(), 4 * g.n, 2) == 0
                assert np.array_equal(got, oracle.bfs(g.beg_pos, g.adj, r)), (launch, r)
                ev = dg.last_device_ms()
                assert 0 < st["device_ms"] <= ev * 1.2 + 0.05, (st["device_ms"], ev)
    finally:
        dg.set_config(base)


def test_back_to_back_searches(oracle, rmat16):
    """Searches on one handle that follow each other at once, every one with
    another launch form, direction and root: the host's wait on a search must
    never take the previous search's final state (its mirrored state words
    are overwritten while the host reads them) for the current one."""
    g, dg = rmat16
    base = dg.config
    roots = [int(r) for r in g.pick_roots(48, 77)]
    forms = list(LAUNCHES.items())
    dirs = list(DIRS.items())
    try:
        for k, r in enumerate(roots):
            lname, launch = forms[k % len(forms)]
            dname, d = dirs[(k // len(forms) + k) % len(dirs)]
            dg.configure(direction=d, launch=launch)
            _check(oracle, g, dg, r, f"rmat16/back-to-back {k}/{dname}/{lname}")
    finally:
        dg.set_config(base)


@pytest.mark.parametrize("n", [3, 5, 1021, 1022, 1023, 1024, 1037])
def test_unaligned_depth_arrays(oracle, n):
    """ent_bfs_device into a caller's depth array at every 4-byte offset from
    a 16-byte boundary, on graphs whose vertex count is not a multiple of 4
    or 16: the depths are exact and the words either side stay untouched."""
    import ctypes as C
    rng = np.random.default_rng(n)
    src = rng.integers(0, n, size=4 * n)
    dst = rng.integers(0, n, size=4 * n)
    g = HostGraph.from_edges(n, src, dst)
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    guard = np.frombuffer(b"\x7f" * 4, dtype=np.int32)[0]
    with DeviceGraph(g) as dg:
        base = dg.device_alloc(4 * (n + 8))
        try:
            for k in range(4):
                assert hip.hipMemset(base, 0x7F, 4 * (n + 8)) == 0
                root = (7 * k + 1) % n
                _say(f"[gpu] unaligned depth n={n} offset={4 * k} root={root}")
                dg.bfs_device(root, base + 4 * k)
                got = dg.copy_depth(base, n + 8)
                assert np.array_equal(got[k:k + n], oracle.bfs(g.beg_pos, g.adj, root)), (n, k, root)
                assert (got[:k] == guard).all() and (got[k + n:] == guard).all(), (n, k)
        finally:
            dg.device_free(base)


@pytest.mark.parametrize("launch", list(LAUNCHES))
def test_topdown_directed_sinks(oracle, launch):
    """Forced top-down is exact on a directed adjacency: the 1,000 vertices
    without out-edges are reached through their in-edges (they are not
    taken as isolated), in every launch form, with tiny class thresholds."""
    beg, adj = graphs.directed_with_sinks(3000, 1000, 3, 77)
    with DeviceGraph(beg_pos=beg, adj=adj) as dg:
        dg.configure(direction=L.DIR_TOPDOWN, launch=LAUNCHES[launch], **TINY_CLASSES)
        for r in (0, 1234, 2999):
            _say(f"[gpu] directed/topdown {launch} root {r}")
            got, st = dg.bfs(r)
            want = oracle.bfs(beg, adj, r)
            assert np.array_equal(got, want), (launch, r)
            assert np.any(want[3000:] > 0)
            reached = want >= 0
            assert st["reached_vertices"] == int(reached.sum())
            assert st["reached_edges"] == int(np.diff(beg)[reached].sum())
