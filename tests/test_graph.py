"""Graph substrate: generator determinism and CSR invariants, the reference's
binary file format (checked against the reference's own loader when it has
been built), and the root pickers."""
import ctypes
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from enterprise_amd import HostGraph
from enterprise_amd import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_LOADER = os.path.join(ROOT, "oracle", "_ref", "ref_loader")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_rmat_hashes_and_invariants():
    with open(os.path.join(HERE, "golden", "hashes.json")) as f:
        h = json.load(f)
    for scale, seed in [(10, 3), (16, 16)]:
        g = HostGraph.rmat(scale, 16, seed=seed)
        g.check()
        e = h[f"rmat{scale}_seed{seed}"]
        assert (g.n, g.m) == (e["n"], e["m"])
        assert sha(g.beg_pos) == e["beg_pos"] and sha(g.adj) == e["adj"]


def test_rmat_thread_count_independent():
    a = HostGraph.rmat(12, 16, seed=5)
    L.lib().ent_set_host_threads(1)
    try:
        b = HostGraph.rmat(12, 16, seed=5)
    finally:
        L.lib().ent_set_host_threads(0)
    assert np.array_equal(a.beg_pos, b.beg_pos) and np.array_equal(a.adj, b.adj)


def test_rmat_degree_skew():
    g = HostGraph.rmat(14, 16, seed=1, relabel=False)
    deg = g.degrees()
    # Without relabelling, vertex 0 collects the a-quadrant mass: the highest
    # expected degree of the Kronecker recursion.
    assert deg[0] == deg.max()
    assert deg.max() > 50 * np.median(deg[deg > 0])


def test_grid_generator():
    g = HostGraph.grid(30, 20, delete_ppm=0, keep_lcc=False)
    g.check()
    assert g.n == 600 and g.m == 2 * (29 * 20 + 30 * 19)
    g2 = HostGraph.grid(64, 64, delete_ppm=20000, seed=4096, keep_lcc=True)
    g2.check()
    with open(os.path.join(HERE, "golden", "hashes.json")) as f:
        e = json.load(f)["grid64_lcc"]
    assert (g2.n, g2.m) == (e["n"], e["m"]) and sha(g2.adj) == e["adj"]
    # Largest component only: every vertex reachable from vertex 0.
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    a = sp.csr_matrix((np.ones(g2.m), g2.adj, g2.beg_pos), shape=(g2.n, g2.n))
    assert connected_components(a)[0] == 1


def test_from_edges_dedup_loops():
    g = HostGraph.from_edges(5, [0, 0, 1, 2, 3, 3], [1, 1, 0, 2, 4, 4])
    g.check()
    assert list(g.beg_pos) == [0, 1, 2, 2, 3, 4]
    assert list(g.adj) == [1, 0, 4, 3]


def test_csr_check_rejects():
    bad = HostGraph.from_arrays([0, 1, 1], [1])  # edge 0->1 without 1->0
    with pytest.raises(L.EntError):
        bad.check(symmetric=True)
    bad.check(symmetric=False)


def test_bin_roundtrip(tmp_path):
    g = HostGraph.rmat(9, 8, seed=2)
    b, c = tmp_path / "g_beg_pos.bin", tmp_path / "g_csr.bin"
    g.write_bin(b, c)
    assert os.path.getsize(b) == 8 * (g.n + 1) and os.path.getsize(c) == 8 * g.m
    assert np.array_equal(np.fromfile(b, dtype=np.int64), g.beg_pos)
    assert np.array_equal(np.fromfile(c, dtype=np.int64), g.adj.astype(np.int64))
    h = HostGraph.read_bin(b, c)
    assert np.array_equal(h.beg_pos, g.beg_pos) and np.array_equal(h.adj, g.adj)


def test_bin_reader_rejects(tmp_path):
    b, c = tmp_path / "x_beg_pos.bin", tmp_path / "x_csr.bin"
    np.array([0, 2, 3], dtype=np.int64).tofile(b)
    np.array([1, 1, 5], dtype=np.int64).tofile(c)   # id 5 >= n = 2
    with pytest.raises(L.EntError):
        HostGraph.read_bin(b, c)
    np.array([0, 2, 4], dtype=np.int64).tofile(b)   # offsets end past the file
    np.array([1, 1, 0], dtype=np.int64).tofile(c)
    with pytest.raises(L.EntError):
        HostGraph.read_bin(b, c)
    with pytest.raises(L.EntError):
        HostGraph.read_bin(tmp_path / "missing", c)


@pytest.mark.skipif(not os.path.exists(REF_LOADER), reason="reference loader not built")
def test_bin_format_matches_reference_loader(tmp_path):
    """The reference's own graph.hpp loader reads our files to the same arrays
    (narrowed to int32, as main.cu instantiates it)."""
    g = HostGraph.rmat(10, 16, seed=3)
    b, c, out = tmp_path / "r_beg_pos.bin", tmp_path / "r_csr.bin", tmp_path / "out.bin"
    g.write_bin(b, c)
    subprocess.run([REF_LOADER, str(b), str(c), str(out)], check=True, capture_output=True)
    raw = np.fromfile(out, dtype=np.int32)
    n, m = int(raw[0]), int(raw[1])
    assert (n, m) == (g.n, g.m)
    assert np.array_equal(raw[2:3 + n], g.beg_pos.astype(np.int32))
    assert np.array_equal(raw[3 + n:], g.adj)


def test_pick_roots_rules():
    g = HostGraph.rmat(10, 4, seed=7)
    deg = g.degrees()
    r = g.pick_roots(64, 1002)
    assert len(set(r.tolist())) == 64 and (deg[r] > 0).all()
    assert np.array_equal(r, g.pick_roots(64, 1002))
    assert not np.array_equal(r, g.pick_roots(64, 1003))


def test_pick_roots_glibc_matches_libc_rand():
    """ent_pick_roots_glibc reproduces main.cu:55-78 under glibc's default seed:
    recomputed here with the C library's own rand()."""
    g = HostGraph.rmat(11, 4, seed=11)
    deg = g.degrees()
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    want = []
    while len(want) < 64:
        v = libc.rand() % g.n
        if deg[v] != 0 and v not in want:
            want.append(v)
    assert g.pick_roots_glibc(64).tolist() == want


def test_pick_roots_mixed():
    g = HostGraph.rmat(14, 16, seed=23)
    deg = g.degrees()
    r = g.pick_roots_mixed(64, 32, 1004)
    assert len(set(r.tolist())) == 64
    cut = np.sort(deg)[::-1][(g.n + 99) // 100 - 1]
    assert (deg[r[:32]] >= cut).all()
    assert ((deg[r[32:]] >= 1) & (deg[r[32:]] <= 4)).all()
