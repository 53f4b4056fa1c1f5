"""GPU parent output: for every search the parents form a BFS tree that
passes the Graph500 parent checks of the oracle (parent[root] = root, each
parent a neighbour one level closer, -1 for unreached vertices), with depths
bit-exact against the oracle; depths that are not a BFS from the given root
are rejected."""
import sys

import numpy as np
import pytest

import graphs
from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd import _lib as L

pytestmark = pytest.mark.gpu


def _check(oracle, g, dg, root, tag):
    print(f"[gpu] parents {tag} root={root}", flush=True)
    sys.stdout.flush()
    d, p, st = dg.bfs_parents(root)
    assert np.array_equal(d, oracle.bfs(g.beg_pos, g.adj, root)), tag
    rc = oracle.validate_parents(g.beg_pos, g.adj, root, d, p)
    assert rc == 0, (tag, root, rc)
    assert p[root] == root
    assert (p[d < 0] == -1).all()
    return p


@pytest.mark.parametrize("name", list(graphs.CLOSED_FORMS))
def test_parents_closed_forms(oracle, name):
    g, _ = graphs.CLOSED_FORMS[name]()
    with DeviceGraph(g) as dg:
        for r in sorted({0, g.n // 2, g.n - 1}):
            _check(oracle, g, dg, r, name)


@pytest.mark.parametrize("keep_order", [False, True])
def test_parents_rmat16(oracle, keep_order):
    g = HostGraph.rmat(16, 16, seed=16)
    with DeviceGraph(g, keep_order=keep_order) as dg:
        for r in g.pick_roots(6, 31):
            _check(oracle, g, dg, int(r), f"rmat16/keep{keep_order}")
        dg.configure(direction=L.DIR_BOTTOMUP)
        _check(oracle, g, dg, int(g.pick_roots(1, 32)[0]), "rmat16/bu")


def test_parents_star_hub_and_leaf(oracle):
    g, _ = graphs.star(70000)
    with DeviceGraph(g) as dg:
        p = _check(oracle, g, dg, 0, "star/hub")
        assert (p[1:] == 0).all()
        p = _check(oracle, g, dg, 5, "star/leaf")
        assert p[0] == 5 and (np.delete(p, [0, 5]) == 0).all()


def test_parents_reject_foreign_depths():
    g = HostGraph.rmat(12, 8, seed=3)
    a, b = (int(x) for x in g.pick_roots(2, 9))
    with DeviceGraph(g) as dg:
        buf = dg.device_alloc(8 * g.n)
        try:
            dg.bfs_device(a, buf)                     # depths of a search from a
            with pytest.raises(L.EntError) as e:      # ... are no BFS from b
                dg.parents_device(b, buf, buf + 4 * g.n)
            assert e.value.code == L.ENT_EINVAL
            dg.parents_device(a, buf, buf + 4 * g.n)  # and are one from a
        finally:
            dg.device_free(buf)
