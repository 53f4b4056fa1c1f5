"""Pins the CPU oracle before it is trusted: scipy BFS on random graphs, closed
forms, committed golden vectors, invariants, and planted mistakes that the
checkers must reject."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import shortest_path

import graphs
from enterprise_amd import HostGraph

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def scipy_depth(g, root):
    a = sp.csr_matrix((np.ones(g.m), g.adj, g.beg_pos), shape=(g.n, g.n))
    d = shortest_path(a, unweighted=True, indices=root)
    return np.where(np.isinf(d), -1, d).astype(np.int32)


@pytest.mark.parametrize("name", list(graphs.CLOSED_FORMS))
def test_closed_forms(oracle, name):
    g, depth_fn = graphs.CLOSED_FORMS[name]()
    for r in sorted({0, g.n // 3, g.n - 1}):
        d = oracle.bfs(g.beg_pos, g.adj, r)
        assert np.array_equal(d, np.asarray(depth_fn(r), dtype=np.int32)), (name, r)
        assert oracle.validate(g.beg_pos, g.adj, r, d) == 0
        assert np.array_equal(oracle.bfs_do(g.beg_pos, g.adj, r, threads=3), d)


@pytest.mark.parametrize("seed", range(6))
def test_random_graphs_vs_scipy(oracle, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 400))
    ne = int(rng.integers(0, 4 * n))
    g = HostGraph.from_edges(n, rng.integers(0, n, ne), rng.integers(0, n, ne))
    for r in rng.integers(0, n, 4):
        want = scipy_depth(g, int(r))
        d = oracle.bfs(g.beg_pos, g.adj, int(r))
        assert np.array_equal(d, want)
        assert oracle.validate(g.beg_pos, g.adj, int(r), d) == 0
        assert np.array_equal(oracle.bfs_do(g.beg_pos, g.adj, int(r), threads=4), want)
        e, v = oracle.report(g.beg_pos, d)
        assert v == int((want >= 0).sum())
        assert e == int(np.diff(g.beg_pos)[want >= 0].sum())


def test_directed_vs_scipy(oracle):
    """The oracle's top-down BFS follows out-edges only, as scipy's directed
    shortest paths do: on a directed adjacency with 1,000 vertices that have
    in-edges only, every depth matches (the GPU forced top-down tests check
    against it there)."""
    beg, adj = graphs.directed_with_sinks(3000, 1000, 3, 77)
    n = beg.size - 1
    a = sp.csr_matrix((np.ones(adj.size), adj, beg), shape=(n, n))
    for r in (0, 1234, 2999):
        d = shortest_path(a, directed=True, unweighted=True, indices=r)
        want = np.where(np.isinf(d), -1, d).astype(np.int32)
        got = oracle.bfs(beg, adj, r)
        assert np.array_equal(got, want), r
        assert np.any(want[3000:] > 0) and np.all(np.diff(beg)[3000:] == 0)


def test_golden_vectors(oracle):
    data = np.load(os.path.join(GOLDEN, "small_graphs.npz"))
    names = sorted({k.split("__")[0] for k in data.files})
    assert len(names) >= 5
    for name in names:
        b, a = data[f"{name}__beg_pos"], data[f"{name}__adj"]
        for r, want in zip(data[f"{name}__roots"], data[f"{name}__depth"]):
            assert np.array_equal(oracle.bfs(b, a, int(r)), want), (name, r)


def test_rmat16_config1_golden(oracle):
    import hashlib
    import json
    with open(os.path.join(GOLDEN, "hashes.json")) as f:
        h = json.load(f)["config1"]
    g = HostGraph.rmat(16, 16, seed=16)
    root = int(g.pick_roots(1, 1001)[0])
    assert root == h["root"]
    d = oracle.bfs(g.beg_pos, g.adj, root)
    assert hashlib.sha256(d.tobytes()).hexdigest() == h["depth"]
    assert int(d.max()) + 1 == h["levels"] and int((d >= 0).sum()) == h["reached"]


@pytest.fixture(scope="module")
def rmat10():
    g = HostGraph.rmat(10, 16, seed=3)
    return g


def test_validate_rejects_planted_mistakes(oracle, rmat10):
    g = rmat10
    r = int(g.pick_roots(1, 9)[0])
    d = oracle.bfs(g.beg_pos, g.adj, r)
    assert oracle.validate(g.beg_pos, g.adj, r, d) == 0
    far = int(np.flatnonzero(d == d.max())[0])
    for mutate, codes in [
        (lambda x: x.__setitem__(far, x[far] + 1), {-1, -6}),      # one level too deep
        (lambda x: x.__setitem__(far, x[far] - 1), {-1, -6}),      # one level too shallow
        (lambda x: x.__setitem__(r, 1), {-5}),                      # root not at 0
        (lambda x: x.__setitem__(far, -1), {-1}),                   # reached marked unreached
        (lambda x: x.__setitem__(far, -3), {-7}),                   # garbage value
    ]:
        bad = d.copy()
        mutate(bad)
        assert oracle.validate(g.beg_pos, g.adj, r, bad) in codes
    # An isolated vertex marked reached has no parent.
    iso = np.flatnonzero(np.diff(g.beg_pos) == 0)
    if iso.size:
        bad = d.copy()
        bad[iso[0]] = 2
        assert oracle.validate(g.beg_pos, g.adj, r, bad) == -6
    # Two vertices at different depths swapped.
    a1 = int(np.flatnonzero(d == 1)[0])
    a3 = int(np.flatnonzero(d == 3)[0])
    bad = d.copy()
    bad[a1], bad[a3] = bad[a3], bad[a1]
    assert oracle.validate(g.beg_pos, g.adj, r, bad) != 0


def test_validate_spot_matches_reference_rules(oracle):
    # Path 0-1-2-3-4, root 0.
    g, _ = graphs.path(5)
    d = oracle.bfs(g.beg_pos, g.adj, 0)
    for b in range(5):
        assert oracle.validate_spot(g.beg_pos, g.adj, d, b) == 0
    bad = d.copy()
    bad[3] = 5          # neighbour of 2 and 4 jumps by more than one
    assert oracle.validate_spot(g.beg_pos, g.adj, bad, 2) == -1
    # -2: a shallower neighbour c of b has a neighbour deeper than b.
    # Star centre 0 with leaves 1..3, root 1: depth 0->1, leaves 2, 3 -> 2.
    g, _ = graphs.star(4)
    d = oracle.bfs(g.beg_pos, g.adj, 1)
    bad = d.copy()
    bad[2] = 1          # leaf 2 claims depth 1 (same as the centre)
    # base = leaf 3 (depth 2): its neighbour 0 (depth 1) is shallower and has
    # neighbour 1.. fine; base = 2 (depth 1): neighbour 0 has depth 1 -> equal
    # branch: 0's neighbours 1 (0), 3 (2) are within +-1 of base.
    assert oracle.validate_spot(g.beg_pos, g.adj, bad, 2) == 0
    bad2 = d.copy()
    bad2[3] = 3         # leaf 3 too deep: base 0 (depth 1) sees neighbour at 3
    assert oracle.validate_spot(g.beg_pos, g.adj, bad2, 0) == -1
    # -3: a deeper neighbour c of b has a neighbour shallower than b.
    # Base 1 (depth 1) has neighbours [2, 3]; 2 is deeper and touches 0
    # (depth 0 < 1). Checked in CSR order, as validate.h:64-99 does.
    g = graphs.edges_graph(4, [(1, 2), (0, 2), (1, 3)])
    d = np.array([0, 1, 2, 2], dtype=np.int32)
    assert oracle.validate_spot(g.beg_pos, g.adj, d, 1) == -3
    # Same shape with the shallower neighbour first gives -2 instead.
    g = graphs.edges_graph(4, [(0, 1), (1, 2), (2, 3), (0, 2)])
    d = np.array([0, 1, 2, 3], dtype=np.int32)
    assert oracle.validate_spot(g.beg_pos, g.adj, d, 1) == -2
    # -4: an equal-depth neighbour c of b has a neighbour two levels away.
    g = graphs.edges_graph(4, [(0, 1), (1, 2), (2, 3)])
    d = np.array([1, 1, 3, 0], dtype=np.int32)
    assert oracle.validate_spot(g.beg_pos, g.adj, d, 0) == -4


def test_parents(oracle, rmat10):
    g = rmat10
    for r in g.pick_roots(3, 4):
        d, p = oracle.bfs_parents(g.beg_pos, g.adj, int(r))
        assert np.array_equal(d, oracle.bfs(g.beg_pos, g.adj, int(r)))
        assert oracle.validate_parents(g.beg_pos, g.adj, int(r), d, p) == 0
        v = int(np.flatnonzero(d == 2)[0])
        bad = p.copy()
        bad[v] = int(r)            # root is not a neighbour at distance 2
        assert oracle.validate_parents(g.beg_pos, g.adj, int(r), d, bad) in (-10, -11)
        bad = p.copy()
        bad[int(r)] = v
        assert oracle.validate_parents(g.beg_pos, g.adj, int(r), d, bad) == -9


def test_report_closed_form(oracle):
    g, _ = graphs.star(100)
    d = oracle.bfs(g.beg_pos, g.adj, 5)
    e, v = oracle.report(g.beg_pos, d)
    assert v == 100 and e == 2 * 99
