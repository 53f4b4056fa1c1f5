"""The drop-in for the reference's main.cu and bfs_gpu_coalescing_mem: the
ent_main CLI on the reference's binary CSR files and on a generated graph,
and the compatibility entry point called through the C ABI."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from enterprise_amd import HostGraph
from enterprise_amd import _lib as L

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "enterprise_amd", "bin", "ent_main")


def _run(args):
    if not os.path.exists(EXE):
        pytest.fail(f"{EXE} is missing: build it with __graft_entry__.build()")
    r = subprocess.run([EXE] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=120)
    out = r.stdout.decode()
    assert r.returncode == 0, (r.returncode, out[-2000:], r.stderr.decode()[-2000:])
    return out


def _check_report(out, g):
    tests = re.findall(r"Test (\d+)\nStarted from: (\d+)", out)
    assert len(tests) == 64 and [int(t) for t, _ in tests] == list(range(1, 65))
    deg = np.diff(g.beg_pos)
    assert all(deg[int(r)] > 0 for _, r in tests)          # non-isolated roots (main.cu:55-78)
    avg = re.findall(r"Final Average TEPS \(Billion\): ([0-9.eE+-]+)", out)
    assert len(avg) == 1 and float(avg[0]) > 0
    # Every counted root (more than 2 levels) is validated on the device.
    counted = len(re.findall(r"Current TEPS", out))
    assert counted > 0 and out.count("BFS result validation: CORRECT") == counted
    assert "WRONG" not in out


def test_cli_reads_reference_files(tmp_path):
    g = HostGraph.rmat(14, 16, seed=14)
    b, c = tmp_path / "g_beg_pos.bin", tmp_path / "g_csr.bin"
    g.write_bin(b, c)
    out = _run([str(b), str(c)])
    assert f"{g.n} verts, {g.m} edges" in out
    _check_report(out, g)


def test_cli_generated_graph():
    out = _run(["--rmat", "12", "16", "12"])
    g = HostGraph.rmat(12, 16, seed=12)
    _check_report(out, g)


def test_compat_entry_point_returns_zero():
    g = HostGraph.rmat(11, 8, seed=4)
    roots = g.pick_roots(64, 77).astype(np.int32)
    b = np.ascontiguousarray(g.beg_pos, dtype=np.int64)
    a = np.ascontiguousarray(g.adj, dtype=np.int32)
    rc = L.lib().ent_bfs_gpu_coalescing_mem(roots.ctypes.data_as(L.i32p), 64, b.ctypes.data_as(L.i64p),
                                             a.ctypes.data_as(L.i32p), C.c_int64(g.n), C.c_int64(g.m), 0)
    assert rc == 0
