"""Parity at the benchmark sizes: R-MAT scale 22 (64-root workload) against
the oracle bit for bit, the 512-root multi-source search on R-MAT scale 23
against the single-source path on every row (device row checksums) and the
oracle on sampled rows, and the 4096 x 4096 grid against the oracle."""
import sys

import numpy as np
import pytest

from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd.bfs import row_hash_host

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)
    sys.stdout.flush()


def test_rmat22_roots_bit_exact(oracle):
    g = HostGraph.rmat(22, 16, seed=22)
    roots = g.pick_roots(64, 1002)
    with DeviceGraph(g) as dg:
        for i in (0, 17, 63):
            r = int(roots[i])
            _say(f"[gpu] rmat22 root {r}")
            got, st = dg.bfs(r)
            want = oracle.bfs(g.beg_pos, g.adj, r)
            assert np.array_equal(got, want), r
            assert oracle.validate(g.beg_pos, g.adj, r, got) == 0
            edges, verts = oracle.report(g.beg_pos, want)
            assert (st["reached_edges"], st["reached_vertices"]) == (edges, verts)
        d, p, _ = dg.bfs_parents(int(roots[5]))
        assert oracle.validate_parents(g.beg_pos, g.adj, int(roots[5]), d, p) == 0
        # Every benchmark root: the device all-vertex check (zero counts
        # exactly for the BFS from the root).
        _say("[gpu] rmat22 validate 64 roots on the device")
        for r in roots:
            dg.bfs_device(int(r))
            assert dg.validate_device(int(r)) == [0, 0, 0, 0], int(r)


def test_rmat23_batch_512_rows(oracle):
    g = HostGraph.rmat(23, 16, seed=23)
    roots = g.pick_roots_mixed(512, 256, 1004)
    with DeviceGraph(g) as dg:
        k = len(roots)
        out = dg.device_alloc(k * g.n * 4)
        try:
            dg.batch_reserve(k, 48)
            _say("[gpu] rmat23 batch 512")
            st = dg.bfs_batch_device(roots, out)
            h = dg.row_hash(out, k, g.n)
            plain = np.empty(k, dtype=np.uint64)
            for i, r in enumerate(roots):
                dg.bfs_device(int(r))
                plain[i] = dg.row_hash(dg.depth_device_ptr(), 1, g.n)[0]
            assert np.array_equal(h, plain)
            assert st["bu_level_mask"] & st["td_level_mask"]   # some level split its words
            for i in (0, 255, 256, 511):
                want = oracle.bfs(g.beg_pos, g.adj, int(roots[i]))
                assert int(h[i]) == int(row_hash_host(want)), i
            _say("[gpu] rmat23 validate 512 rows on the device")
            for i, r in enumerate(roots):
                assert dg.validate_device(int(r), out + 4 * g.n * i) == [0, 0, 0, 0], i
        finally:
            dg.device_free(out)


def test_grid4096_root(oracle):
    g = HostGraph.grid(4096, 4096, delete_ppm=20000, seed=4096, keep_lcc=True)
    r = int(g.pick_roots(16, 1003)[3])
    with DeviceGraph(g) as dg:
        _say(f"[gpu] grid4096 root {r}")
        got, st = dg.bfs(r)
        want = oracle.bfs(g.beg_pos, g.adj, r)
        assert np.array_equal(got, want)
        assert st["levels"] == int(want.max()) + 1 and st["levels"] > 4000
        for r2 in g.pick_roots(16, 1003):
            dg.bfs_device(int(r2))
            assert dg.validate_device(int(r2)) == [0, 0, 0, 0], int(r2)
