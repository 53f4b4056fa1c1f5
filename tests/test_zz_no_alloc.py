"""The no-allocation row checksum entry point against the handle's one:
two calls in a row (the second with fewer rows and columns), the row limit,
and argument errors."""
import ctypes as C
import sys

import numpy as np
import pytest

from enterprise_amd import DeviceGraph, HostGraph
from enterprise_amd import _lib as L
from enterprise_amd.bfs import row_hash_host

pytestmark = pytest.mark.gpu

MAX_ROWS = 4096  # ENT_NOALLOC_HASH_ROWS


def _noalloc(ptr, rows, cols, device=0):
    out = np.zeros(rows, dtype=np.uint64)
    rc = L.lib().ent_device_row_hash_noalloc(int(device), C.c_void_p(ptr), int(rows), int(cols),
                                             out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, out


def test_row_hash_noalloc_matches_handle():
    g = HostGraph.rmat(16, 16, seed=16)
    roots = g.pick_roots(40, 9)
    with DeviceGraph(g) as dg:
        ptr = dg.device_alloc(40 * g.n * 4)
        try:
            dg.bfs_batch_device(roots, ptr)
            print("[gpu] no-alloc row hash: 40 rows", flush=True)
            sys.stdout.flush()
            want = dg.row_hash(ptr, 40, g.n)
            rc, got = _noalloc(ptr, 40, g.n)
            assert rc == L.ENT_OK and np.array_equal(got, want)
            d, _ = dg.bfs(int(roots[0]))
            assert got[0] == row_hash_host(d)
            # Second call, smaller input: fewer rows and a shorter row length
            # (the scratch still holds the first call's larger sums).
            cols2 = g.n // 2 + 3
            want2 = dg.row_hash(ptr + 8, 7, cols2)
            rc, got2 = _noalloc(ptr + 8, 7, cols2)
            assert rc == L.ENT_OK and np.array_equal(got2, want2)
            host = dg.copy_depth(ptr + 8, 7 * cols2)
            assert np.array_equal(got2, row_hash_host(host.reshape(7, cols2)))
            # The largest row count, then one more.
            cols3 = (40 * g.n - 2) // MAX_ROWS
            rc, got3 = _noalloc(ptr, MAX_ROWS, cols3)
            assert rc == L.ENT_OK and np.array_equal(got3, dg.row_hash(ptr, MAX_ROWS, cols3))
            rc, _ = _noalloc(ptr, MAX_ROWS + 1, 1)
            assert rc == L.ENT_EINVAL and b"scratch" in L.lib().ent_last_error()
        finally:
            dg.device_free(ptr)
    rc, _ = _noalloc(None, 1, 1)
    assert rc == L.ENT_EINVAL
    rc, _ = _noalloc(8, 1, 1, device=1 << 20)
    assert rc == L.ENT_EINVAL
