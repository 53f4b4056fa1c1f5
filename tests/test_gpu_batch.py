"""GPU parity of the multi-source BFS: every output row equals the
single-source GPU result and the CPU oracle for its root."""
import sys

import numpy as np
import pytest

import graphs
from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd import _lib as L
from enterprise_amd.bfs import row_hash_host

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat16():
    g = HostGraph.rmat(16, 16, seed=16)
    dg = DeviceGraph(g)
    yield g, dg
    dg.free()


def _check_batch(oracle, g, dg, roots, tag, oracle_rows=4):
    _say(f"[gpu] batch {tag} k={len(roots)}")
    out, st = dg.bfs_batch(roots)
    assert out.shape == (len(roots), g.n)
    for i, r in enumerate(roots):
        want, _ = dg.bfs(int(r))
        if not np.array_equal(out[i], want):
            bad = np.flatnonzero(out[i] != want)
            raise AssertionError(f"{tag}: source {i} root {r}: {bad.size} mismatches, first v={bad[0]} "
                                 f"batch={out[i][bad[0]]} single={want[bad[0]]}")
    for i in range(min(oracle_rows, len(roots))):
        assert np.array_equal(out[i], oracle.bfs(g.beg_pos, g.adj, int(roots[i])))
    deg = g.degrees()
    reached = out >= 0
    assert st["reached_vertices"] == int(reached.sum())
    assert st["reached_edges"] == int((reached * deg[None, :]).sum())
    assert st["levels"] == int(out.max()) + 1
    return st


@pytest.mark.parametrize("k", [1, 7, 64, 65, 130, 512])
def test_batch_sizes(oracle, rmat16, k):
    g, dg = rmat16
    roots = g.pick_roots(k, 100 + k) if k <= 64 else np.r_[g.pick_roots(64, 5),
                                                          g.pick_roots(k - 64, 6)]
    _check_batch(oracle, g, dg, roots, f"rmat16/k{k}")


def test_batch_uneven_and_repeated_roots(oracle, rmat16):
    g, dg = rmat16
    roots = g.pick_roots_mixed(64, 32, 1004)
    roots = np.r_[roots, roots[:5]]           # repeats are allowed
    _check_batch(oracle, g, dg, roots, "rmat16/mixed+repeats")


@pytest.mark.parametrize("direction", ["auto", "topdown", "bottomup", "reference"])
def test_batch_directions(oracle, rmat16, direction):
    g, dg = rmat16
    base = dg.config
    d = {"auto": L.DIR_AUTO, "topdown": L.DIR_TOPDOWN, "bottomup": L.DIR_BOTTOMUP,
         "reference": L.DIR_REFERENCE}[direction]
    try:
        dg.configure(direction=d, t_wave=4, t_block=24, t_grid=96, piece=64)
        _check_batch(oracle, g, dg, g.pick_roots(96, 9), f"rmat16/{direction}/tiny-classes")
    finally:
        dg.set_config(base)


@pytest.mark.parametrize("name", ["cycle257", "star5000", "grid37x23", "disjoint", "hypercube10",
                                  "bintree11"])
def test_batch_closed_forms(oracle, name):
    g, depth_fn = graphs.CLOSED_FORMS[name]()
    roots = np.unique(np.r_[0, g.n // 2, g.n - 1, np.arange(0, g.n, max(1, g.n // 40))])[:64]
    with DeviceGraph(g) as dg:
        dg.batch_reserve(len(roots), 200)
        out, _ = dg.bfs_batch(roots)
        for i, r in enumerate(roots):
            assert np.array_equal(out[i], np.asarray(depth_fn(int(r)), dtype=np.int32)), (name, r)


def expected_flushes(levels, kept):
    """Flush passes of a search of `levels` joint levels keeping `kept`
    levels: level l starts with a flush of base .. l when l - base > kept."""
    base = flushes = 0
    for lv in range(1, levels):
        if lv - base > kept:
            flushes += 1
            base = lv + 1
    return flushes


@pytest.mark.parametrize("launch", ["stream", "persistent"])
def test_batch_deeper_than_kept_levels(oracle, launch):
    """A search deeper than the kept levels runs as one batch: flush passes
    write the kept levels to the rows (whole rows the first time, then only
    the depths each pass found) and their slots are reused."""
    form = {"stream": L.LAUNCH_STREAM, "persistent": L.LAUNCH_PERSISTENT}[launch]
    p300, pdepth = graphs.path(300)
    with DeviceGraph(p300) as dp:
        dp.configure(launch=form)
        dp.batch_reserve(4, max_levels=10)
        _say(f"[gpu] batch deeper than kept levels: path300 {launch}")
        out, st = dp.bfs_batch([0, 1, 2, 299])
        for i, r in enumerate([0, 1, 2, 299]):
            assert np.array_equal(out[i], np.asarray(pdepth(r), dtype=np.int32)), r
        assert st["levels"] == 300 and st["flushes"] == expected_flushes(300, 10) == 25  # at levels 11, 23, .., 299
        out, st = dp.bfs_batch([150])            # a second search on the same ring
        assert np.array_equal(out[0], np.asarray(pdepth(150), dtype=np.int32))
        assert st["flushes"] == expected_flushes(151, 10)
        dp.batch_reserve(4, max_levels=250)
        out, st = dp.bfs_batch([100, 150])      # 200 levels kept
        assert out[0][299] == 199 and out[1][0] == 150 and st["flushes"] == 0
    # One deep source among 63 shallow ones: a 400-vertex path next to a
    # 2,000-leaf star.
    n1, n2 = 400, 2000
    e = [(i, i + 1) for i in range(n1 - 1)] + [(n1, n1 + j) for j in range(1, n2)]
    g = graphs.edges_graph(n1 + n2, e)
    roots = np.r_[np.arange(n1 + 1, n1 + 64), 0].astype(np.int32)
    with DeviceGraph(g) as dg:
        dg.configure(launch=form)
        dg.batch_reserve(64, max_levels=40)
        _say(f"[gpu] batch beyond capacity: path+star {launch}")
        out, st = dg.bfs_batch(roots)
        assert st["flushes"] == expected_flushes(n1, 40), st
        for i, r in enumerate(roots):
            assert np.array_equal(out[i], oracle.bfs(g.beg_pos, g.adj, int(r))), (i, r)
        deg = g.degrees()
        reached = out >= 0
        assert st["reached_vertices"] == int(reached.sum())
        assert st["reached_edges"] == int((reached * deg[None, :]).sum())
        assert st["levels"] == n1


def test_batch_unaligned_output(rmat16):
    """Rows written to an output pointer 4 bytes past a 16-byte boundary (the
    scalar-store path) equal those written to an aligned one."""
    g, dg = rmat16
    roots = g.pick_roots(32, 77)
    ptr = dg.device_alloc(32 * g.n * 4 + 64)
    try:
        _say("[gpu] batch output alignment")
        dg.bfs_batch_device(roots, ptr + 16)
        want = dg.row_hash(ptr + 16, 32, g.n)
        for off in (4, 8, 12):
            dg.bfs_batch_device(roots, ptr + off)
            assert np.array_equal(dg.row_hash(ptr + off, 32, g.n), want), off
        d, _ = dg.bfs(int(roots[3]))
        assert want[3] == row_hash_host(d)
    finally:
        dg.device_free(ptr)


def test_row_hash_many_rows(rmat16):
    """Checksums of more rows than one launch sums (512 per launch)."""
    g, dg = rmat16
    roots = g.pick_roots(32, 78)
    ptr = dg.device_alloc(32 * g.n * 4)
    try:
        dg.bfs_batch_device(roots, ptr)
        host = dg.copy_depth(ptr, 32 * g.n)
        rows, cols = 1100, (32 * g.n) // 1100
        got = dg.row_hash(ptr, rows, cols)
        assert np.array_equal(got, row_hash_host(host[:rows * cols].reshape(rows, cols)))
    finally:
        dg.device_free(ptr)


def test_batch_device_hash(rmat16):
    g, dg = rmat16
    roots = g.pick_roots(32, 55)
    ptr = dg.device_alloc(32 * g.n * 4)
    try:
        dg.bfs_batch_device(roots, ptr)
        h = dg.row_hash(ptr, 32, g.n)
        for i, r in enumerate(roots[:8]):
            d, _ = dg.bfs(int(r))
            assert h[i] == row_hash_host(d)
    finally:
        dg.device_free(ptr)


@pytest.mark.parametrize("k", [1, 65, 130, 512])
def test_batch_persistent_form(oracle, rmat16, k):
    """The whole multi-source search in one cooperative launch gives the same
    rows and counts as the per-phase launches."""
    g, dg = rmat16
    roots = g.pick_roots(k, 300 + k) if k <= 64 else np.r_[g.pick_roots(64, 7),
                                                          g.pick_roots(k - 64, 8)]
    base = dg.config
    try:
        dg.configure(launch=L.LAUNCH_STREAM)
        want, st_s = dg.bfs_batch(roots)
        dg.configure(launch=L.LAUNCH_PERSISTENT)
        st = _check_batch(oracle, g, dg, roots, f"rmat16/persistent/k{k}")
        got, _ = dg.bfs_batch(roots)
        assert np.array_equal(got, want)
        for key in ("levels", "reached_edges", "reached_vertices", "alg_bytes"):
            assert st[key] == st_s[key], key
    finally:
        dg.set_config(base)


@pytest.mark.parametrize("name", ["cycle257", "star5000", "grid37x23", "disjoint"])
def test_batch_persistent_closed_forms(oracle, name):
    g, depth_fn = graphs.CLOSED_FORMS[name]()
    roots = np.unique(np.r_[0, g.n // 2, g.n - 1, np.arange(0, g.n, max(1, g.n // 40))])[:64]
    with DeviceGraph(g) as dg:
        dg.configure(launch=L.LAUNCH_PERSISTENT)
        dg.batch_reserve(len(roots), 200)
        out, _ = dg.bfs_batch(roots)
        for i, r in enumerate(roots):
            assert np.array_equal(out[i], np.asarray(depth_fn(int(r)), dtype=np.int32)), (name, r)


def test_batch_split_levels(oracle, rmat16):
    """Half the roots from the top-1% degree vertices, half of degree 1..4:
    the hub words are pulled while the lagging words are pushed at the same
    level (both level bits set), and every row still equals the single-source
    result, in both launch forms and with forced directions."""
    g, dg = rmat16
    roots = g.pick_roots_mixed(512, 256, 2024)
    base = dg.config
    try:
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT):
            dg.configure(launch=launch, direction=L.DIR_AUTO)
            st = _check_batch(oracle, g, dg, roots, f"rmat16/split/{launch}")
            assert st["bu_level_mask"] & st["td_level_mask"], st
        for d in (L.DIR_TOPDOWN, L.DIR_BOTTOMUP, L.DIR_REFERENCE):
            dg.configure(launch=L.LAUNCH_STREAM, direction=d)
            st = _check_batch(oracle, g, dg, roots, f"rmat16/split/dir{d}")
            assert not (st["bu_level_mask"] & st["td_level_mask"]), st
    finally:
        dg.set_config(base)


def test_batch_stale_level_rows(oracle, rmat16):
    """Level masks are defined only where the level's joint frontier holds
    the vertex, and a level clears the next-but-one level's masks only for
    vertices the next push can reach. Batches that leave different masks
    behind run back to back on one workspace (hub roots, then low-degree
    roots, then the hub roots again, in both launch forms), and every row
    still equals the single-source result."""
    g, dg = rmat16
    hubs = g.pick_roots_mixed(256, 256, 3031)
    low = g.pick_roots_mixed(192, 0, 3032)
    base = dg.config
    try:
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT):
            dg.configure(launch=launch)
            for j, roots in enumerate((hubs, low, hubs)):
                _check_batch(oracle, g, dg, roots, f"rmat16/stale/{launch}/{j}")
    finally:
        dg.set_config(base)


def test_batch_directed_topdown(oracle):
    """A forced top-down batch is exact on a directed adjacency: vertices
    without out-edges are still reached through their in-edges (their level
    masks are cleared like every other unfinished vertex's). Two batches in
    a row on one workspace, against the oracle and the single-source path."""
    beg, adj = graphs.directed_with_sinks(3000, 1000, 3, 77)
    n = beg.size - 1
    rng = np.random.default_rng(78)
    with DeviceGraph(beg_pos=beg, adj=adj) as dg:
        dg.configure(direction=L.DIR_TOPDOWN)
        for j, k in enumerate((130, 64)):
            roots = rng.choice(3000, size=k, replace=False).astype(np.int32)
            _say(f"[gpu] batch directed/topdown k={k}")
            out, st = dg.bfs_batch(roots)
            for i, r in enumerate(roots):
                want = oracle.bfs(beg, adj, int(r))
                assert np.array_equal(out[i], want), (j, i, int(r))
                single, _ = dg.bfs(int(r))
                assert np.array_equal(single, want), (j, i, int(r))
            assert np.any(out[:, 3000:] > 0)                 # sinks reached
            assert st["levels"] == int(out.max()) + 1
            assert n == 4000


@pytest.mark.parametrize("launch", ["stream", "persistent"])
def test_batch_grid_many_flushes(oracle, launch):
    """A 64-source batch on a 200 x 40 grid (about 240 levels) keeping 16
    levels: more than a dozen flush passes, every row against the oracle and
    the reached counts."""
    form = {"stream": L.LAUNCH_STREAM, "persistent": L.LAUNCH_PERSISTENT}[launch]
    g = HostGraph.grid(200, 40, delete_ppm=20000, seed=240, keep_lcc=True)
    roots = g.pick_roots(64, 241)
    with DeviceGraph(g) as dg:
        dg.configure(launch=form)
        dg.batch_reserve(64, max_levels=16)
        _say(f"[gpu] batch grid200x40 {launch}")
        out, st = dg.bfs_batch(roots)
        for i, r in enumerate(roots):
            assert np.array_equal(out[i], oracle.bfs(g.beg_pos, g.adj, int(r))), (i, int(r))
        assert st["levels"] == int(out.max()) + 1 and st["levels"] > 150
        assert st["flushes"] == expected_flushes(st["levels"], 16) > 10
        deg = g.degrees()
        reached = out >= 0
        assert st["reached_vertices"] == int(reached.sum())
        assert st["reached_edges"] == int((reached * deg[None, :]).sum())


@pytest.mark.parametrize("launch", ["stream", "persistent"])
def test_batch_flushes_wide_masks(oracle, launch):
    """Flushes with eight mask words per vertex: 300 sources (W = 8) on a
    100 x 30 grid keeping 8 levels; every row against the oracle."""
    form = {"stream": L.LAUNCH_STREAM, "persistent": L.LAUNCH_PERSISTENT}[launch]
    g = HostGraph.grid(100, 30, delete_ppm=20000, seed=300, keep_lcc=True)
    roots = g.pick_roots(300, 301)
    with DeviceGraph(g) as dg:
        dg.configure(launch=form)
        dg.batch_reserve(300, max_levels=8)
        _say(f"[gpu] batch grid100x30 W=8 {launch}")
        out, st = dg.bfs_batch(roots)
        assert st["words"] == 8
        for i, r in enumerate(roots):
            assert np.array_equal(out[i], oracle.bfs(g.beg_pos, g.adj, int(r))), (i, int(r))
        assert st["flushes"] == expected_flushes(st["levels"], 8) > 5


def _rows_and_stats(dg, roots, **cfg):
    base = dg.config
    try:
        dg.configure(**cfg)
        return dg.bfs_batch(roots)
    finally:
        dg.set_config(base)


@pytest.mark.parametrize("launch", ["stream", "persistent"])
def test_batch_fused_levels(oracle, rmat16, launch):
    """Fused levels (push and settle in one launch, claims by the returning
    OR into the next level's masks) give the rows of unfused levels. One
    workspace per source count (W = 1, 2, 4) runs batch_fused = 0, 1 (every
    all-push level fused), 8, 0, 1 in a row, so searches that leave the
    level ring unlisted are followed by fused ones; every row equals the
    unfused row, the first rows the oracle, and the counts agree."""
    form = {"stream": L.LAUNCH_STREAM, "persistent": L.LAUNCH_PERSISTENT}[launch]
    g, dg = rmat16
    for k in (64, 128, 256):
        dg.batch_release()  # a workspace of the batch's own width
        roots = g.pick_roots_mixed(k, k // 4, 4000 + k)
        ref = None
        for fused in (0, 1, 8, 0, 1):
            _say(f"[gpu] batch fused={fused} k={k} {launch}")
            out, st = _rows_and_stats(dg, roots, launch=form, batch_fused=fused)
            assert st["words"] == k // 64
            if ref is None:
                ref = (out, st)
                for i in range(4):
                    assert np.array_equal(out[i], oracle.bfs(g.beg_pos, g.adj, int(roots[i])))
                continue
            assert np.array_equal(out, ref[0]), (k, fused)
            for key in ("levels", "reached_vertices", "reached_edges", "alg_bytes"):
                assert st[key] == ref[1][key], (k, fused, key)
        dg.batch_release()


@pytest.mark.parametrize("launch", ["stream", "persistent"])
def test_batch_fused_grid(oracle, launch):
    """A high-diameter batch with fused levels takes one launch per such
    level instead of two, with flushes in between (200 x 40 grid, 64
    sources, 16 levels kept); rows against the oracle."""
    form = {"stream": L.LAUNCH_STREAM, "persistent": L.LAUNCH_PERSISTENT}[launch]
    g = HostGraph.grid(200, 40, delete_ppm=20000, seed=240, keep_lcc=True)
    roots = g.pick_roots(64, 242)
    with DeviceGraph(g) as dg:
        dg.batch_reserve(64, max_levels=16)
        _say(f"[gpu] batch fused grid200x40 {launch}")
        plain, st0 = _rows_and_stats(dg, roots, launch=form, batch_fused=0)
        fused, st1 = _rows_and_stats(dg, roots, launch=form, batch_fused=1)
        for i, r in enumerate(roots):
            assert np.array_equal(fused[i], oracle.bfs(g.beg_pos, g.adj, int(r))), (i, int(r))
        assert np.array_equal(plain, fused)
        assert st1["levels"] == st0["levels"] > 150
        assert st1["flushes"] == st0["flushes"] > 10
        # Levels whose joint frontier's degree sum exceeds n stay unfused.
        assert st1["phases"] < 0.75 * st0["phases"], (st0["phases"], st1["phases"])
        assert st1["alg_bytes"] == st0["alg_bytes"]


def test_batch_fused_directed_topdown(oracle):
    """Fused forced top-down levels on a directed adjacency: vertices without
    out-edges are claimed through their in-edges like any other."""
    beg, adj = graphs.directed_with_sinks(3000, 1000, 3, 79)
    rng = np.random.default_rng(80)
    with DeviceGraph(beg_pos=beg, adj=adj) as dg:
        dg.configure(direction=L.DIR_TOPDOWN, batch_fused=1)
        for k in (200, 64):
            roots = rng.choice(3000, size=k, replace=False).astype(np.int32)
            _say(f"[gpu] batch fused directed/topdown k={k}")
            out, st = dg.bfs_batch(roots)
            for i, r in enumerate(roots):
                assert np.array_equal(out[i], oracle.bfs(beg, adj, int(r))), (k, i, int(r))
            assert np.any(out[:, 3000:] > 0)


def test_batch_pull_in_level_kernel(oracle, rmat16):
    """After a search that never pulled, the next search on the workspace
    queues one kernel per launch, and a pull it meets runs in that kernel
    (W <= 4). A forced top-down batch, then two direction-optimizing batches
    (the first pulls inside k_blevel, the second with the pull kernel
    queued again), for W = 1 and W = 4: every row against the single-source
    path and the first rows against the oracle."""
    g, dg = rmat16
    base = dg.config
    try:
        for k in (64, 256):
            dg.batch_release()
            roots = g.pick_roots_mixed(k, k // 2, 5000 + k)
            dg.configure(direction=L.DIR_TOPDOWN)
            st = _check_batch(oracle, g, dg, roots, f"rmat16/k{k}/topdown")
            assert st["bu_level_mask"] == 0
            dg.configure(direction=L.DIR_AUTO)
            for j in range(2):
                st = _check_batch(oracle, g, dg, roots, f"rmat16/k{k}/auto{j}")
                assert st["bu_level_mask"] != 0 and st["words"] == k // 64
    finally:
        dg.set_config(base)


def test_batch_narrow_on_wide_workspace(oracle, rmat16):
    """A workspace reserved for 512 sources runs each batch at the width its
    own source count needs: 64, 200, 512 and 64 sources again, in both launch
    forms, each row against the single-source path and the first rows
    against the oracle; the fused 64-source search after the 512-source one
    starts from rings in another layout."""
    g, dg = rmat16
    base = dg.config
    try:
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT):
            dg.configure(launch=launch)
            dg.batch_release()
            dg.batch_reserve(512, 12)
            for j, k in enumerate((64, 200, 512, 64)):
                roots = g.pick_roots_mixed(k, k // 2, 6000 + 10 * j + k)
                st = _check_batch(oracle, g, dg, roots, f"rmat16/wide/{launch}/{k}")
                assert st["words"] == (1 if k <= 64 else 4 if k <= 256 else 8)
    finally:
        dg.set_config(base)
        dg.batch_release()


def test_batch_root_order(oracle, rmat16):
    """Sources are placed in the 64-source words by root degree, so the same
    roots in another order give the same rows, permuted, and the same
    counts: 512 mixed roots (256 hubs, 256 of degree 1..4) in the picker's
    order, shuffled, and reversed, in both launch forms, every row against
    the single-source path."""
    g, dg = rmat16
    roots = g.pick_roots_mixed(512, 256, 7100)
    rng = np.random.default_rng(7101)
    base = dg.config
    try:
        for launch in (L.LAUNCH_STREAM, L.LAUNCH_PERSISTENT):
            dg.configure(launch=launch)
            ref, ref_st = dg.bfs_batch(roots)
            for name, perm in (("shuffled", rng.permutation(512)), ("reversed", np.arange(511, -1, -1))):
                _say(f"[gpu] batch root order {name} {launch}")
                out, st = dg.bfs_batch(roots[perm])
                assert np.array_equal(out, ref[perm]), name
                for key in ("levels", "reached_vertices", "reached_edges", "alg_bytes"):
                    assert st[key] == ref_st[key], (name, key)
            for i in (0, 255, 256, 511):
                want, _ = dg.bfs(int(roots[i]))
                assert np.array_equal(ref[i], want), i
            assert np.array_equal(ref[0], oracle.bfs(g.beg_pos, g.adj, int(roots[0])))
    finally:
        dg.set_config(base)
