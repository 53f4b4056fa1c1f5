"""The device all-vertex validator (ent_bfs_validate_device) against the
oracle's validator: zero counts exactly on the oracle's accepted depth arrays,
and a count in the class of the oracle's first failure for each planted
mistake (mistakes are planted on the CPU, in arrays copied to the device)."""
import ctypes as C

import numpy as np
import pytest

import graphs
from enterprise_amd import DeviceGraph, HostGraph

pytestmark = pytest.mark.gpu

# Oracle return codes -> device count index.
CLASS = {-5: 0, -7: 1, -1: 2, -6: 3}


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _planted(depth, root, g, rng):
    """Depth arrays with one mistake each (name, array)."""
    n = depth.size
    reached = np.flatnonzero(depth > 0)
    out = []
    d = depth.copy(); d[root] = 1; out.append(("root_not_zero", d))
    if reached.size:
        v = int(rng.choice(reached))
        d = depth.copy(); d[v] += 2; out.append(("plus_two", d))
        d = depth.copy(); d[v] = -1; out.append(("reached_marked_unreached", d))
        d = depth.copy(); d[v] = -5; out.append(("below_minus_one", d))
        d = depth.copy(); d[v] = n + 3; out.append(("at_least_n", d))
        d = depth.copy(); d[v] -= 1; out.append(("minus_one", d))
        far = reached[depth[reached] == depth.max()]
        near = reached[depth[reached] == 1]
        if far.size and near.size and depth.max() > 2:
            a, b = int(far[0]), int(near[0])
            d = depth.copy(); d[a], d[b] = d[b], d[a]; out.append(("swap_levels", d))
    unreached = np.flatnonzero(depth < 0)
    if unreached.size:
        d = depth.copy(); d[int(unreached[0])] = 1; out.append(("unreached_marked_reached", d))
    return out


@pytest.mark.parametrize("name", ["rmat16", "grid", "disjoint", "star", "path"])
def test_validator_matches_oracle(oracle, name):
    if name == "rmat16":
        g = HostGraph.rmat(16, 16, seed=16)
    else:
        g = {"grid": graphs.CLOSED_FORMS["grid37x23"], "disjoint": graphs.CLOSED_FORMS["disjoint"],
             "star": graphs.CLOSED_FORMS["star5000"], "path": graphs.CLOSED_FORMS["path300"]}[name]()[0]
    rng = np.random.default_rng(3)
    hip = _hip()
    with DeviceGraph(g) as dg:
        buf = dg.device_alloc(4 * g.n)
        try:
            for r in [int(x) for x in g.pick_roots(3, 5)]:
                depth, _ = dg.bfs(r)
                assert dg.validate_device(r) == [0, 0, 0, 0]
                assert oracle.validate(g.beg_pos, g.adj, r, depth) == 0
                # The depths of another root fail for this one.
                cases = _planted(depth, r, g, rng)
                other = int(g.pick_roots(1, 99)[0])
                if other != r:
                    cases.append(("other_root", oracle.bfs(g.beg_pos, g.adj, other)))
                for tag, d in cases:
                    d = np.ascontiguousarray(d, dtype=np.int32)
                    assert hip.hipMemcpy(buf, d.ctypes.data, 4 * g.n, 1) == 0
                    got = dg.validate_device(r, buf)
                    want = oracle.validate(g.beg_pos, g.adj, r, d)
                    assert (want == 0) == (sum(got) == 0), (name, r, tag, want, got)
                    if want:
                        assert got[CLASS[want]] > 0, (name, r, tag, want, got)
        finally:
            dg.device_free(buf)
