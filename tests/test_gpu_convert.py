"""GPU edge-list to CSR construction (include/ent_convert.h) against the
oracle: every flag combination bit for bit (offsets, adjacency, weights),
the raw R-MAT tuples against the host generator's CSR, the ent_convert CLI
against the reference converter's committed outputs, and BFS on a handle
built from tuples."""
import os
import subprocess
import sys

import numpy as np
import pytest

from enterprise_amd import DeviceGraph, EntError, HostGraph
from enterprise_amd import _lib as L
from enterprise_amd import convert

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLD = os.path.join(HERE, "golden", "convert")
CLI = os.path.join(ROOT, "enterprise_amd", "bin", "ent_convert")


def _say(msg):
    print(msg, flush=True)


def _tuples(seed, n, ne):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, ne).astype(np.int32)
    dst = rng.integers(0, n, ne).astype(np.int32)
    dup = rng.integers(0, max(ne, 1), ne // 6)
    if ne:
        src, dst = np.concatenate([src, src[dup]]), np.concatenate([dst, dst[dup]])
    return src, dst


@pytest.mark.parametrize("flags", list(range(16)))
@pytest.mark.parametrize("n,ne,seed", [(1, 0, 0), (1, 5, 1), (7, 30, 2), (300, 2000, 3), (5000, 40000, 4)])
def test_build_matches_oracle(oracle, flags, n, ne, seed):
    src, dst = _tuples(seed, n, ne)
    w = np.random.default_rng(seed + 99).integers(1, 64, src.size).astype(np.int32)
    _say(f"[gpu] csr_build n={n} ne={src.size} flags={flags}")
    g, wo, st = convert.csr_build(n, src, dst, flags, weight=w)
    b, a, ww = oracle.csr_build(n, src, dst, flags, weight=w)
    assert np.array_equal(g.beg_pos, b), (flags, n)
    assert np.array_equal(g.adj, a), (flags, n)
    assert np.array_equal(wo, ww), (flags, n)
    assert st["m"] == a.size and st["records"] == src.size * (2 if flags & L.CSR_REVERSE else 1)
    # Without weights the sorted forms sort keys alone: same CSR.
    g2, _, _ = convert.csr_build(n, src, dst, flags)
    assert np.array_equal(g2.beg_pos, b) and np.array_equal(g2.adj, a)


def test_build_large_ids_and_isolated_runs(oracle):
    """Ids at both ends of n = 2^30 + 12345 (runs of a billion empty rows,
    filled by the gap kernel) and a 62-bit sort key."""
    n = (1 << 30) + 12345
    rng = np.random.default_rng(8)
    src = (n - 1 - rng.integers(0, 1000, 3000)).astype(np.int32)
    dst = (n - 1 - rng.integers(0, 1000, 3000)).astype(np.int32)
    src[:10] = np.arange(10)
    dst[10:20] = np.arange(500, 510)
    _say("[gpu] csr_build n=2^30+12345")
    for flags in (L.CSR_SIMPLE, L.CSR_REVERSE):
        g, _, st = convert.csr_build(n, src, dst, flags)
        assert st["key_bits"] == (62 if flags & L.CSR_DEDUP else 31)
        # Against the oracle on relabelled ids (its offsets would need 8 GiB).
        ids = np.unique(np.concatenate([src, dst]))
        rs, rd = np.searchsorted(ids, src).astype(np.int32), np.searchsorted(ids, dst).astype(np.int32)
        wb, wa, _ = oracle.csr_build(ids.size, rs, rd, flags)
        b = g.beg_pos
        assert np.array_equal(b[ids], wb[:-1]) and b[-1] == wb[-1] == g.m
        assert np.array_equal(g.adj, ids[wa].astype(np.int32))
        # Offsets of the empty rows between the touched ones.
        assert np.array_equal(b[np.arange(0, n + 1, 1 << 20)], wb[np.searchsorted(ids, np.arange(0, n + 1, 1 << 20))])


def test_bad_tuples_rejected():
    with pytest.raises(EntError) as e:
        convert.csr_build(10, np.array([1, 10], np.int32), np.array([2, 3], np.int32), 0)
    assert e.value.code == L.ENT_EINVAL
    with pytest.raises(EntError) as e:
        convert.csr_build(10, np.array([1], np.int32), np.array([2], np.int32), 64)
    assert e.value.code == L.ENT_EINVAL


@pytest.mark.parametrize("scale", [16, 20])
def test_rmat_tuples_build_generator_graph(scale):
    """Size-independent property at scale 20: the device build of the raw
    tuples equals the host generator's CSR."""
    src, dst = convert.rmat_tuples(scale, 16, seed=scale)
    _say(f"[gpu] rmat tuples scale {scale}")
    g, _, st = convert.csr_build(1 << scale, src, dst, L.CSR_SIMPLE)
    h = HostGraph.rmat(scale, 16, seed=scale)
    assert np.array_equal(g.beg_pos, h.beg_pos) and np.array_equal(g.adj, h.adj)
    assert st["device_ms"] > 0


def test_build_device_arrays(oracle):
    """ent_csr_build_device on device-resident arrays."""
    src, dst = _tuples(12, 4000, 30000)
    h = HostGraph.rmat(8, 4, seed=1)
    with DeviceGraph(h) as dg:
        ne = src.size
        ps, pd = dg.device_alloc(4 * ne), dg.device_alloc(4 * ne)
        pb, pa = dg.device_alloc(8 * 4001), dg.device_alloc(4 * 2 * ne)
        try:
            import ctypes as C
            hip = C.CDLL("libamdhip64.so")
            hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            assert hip.hipMemcpy(ps, src.ctypes.data, 4 * ne, 1) == 0
            assert hip.hipMemcpy(pd, dst.ctypes.data, 4 * ne, 1) == 0
            m, st = convert.csr_build_device(4000, ne, ps, pd, L.CSR_SIMPLE, pb, pa, 2 * ne)
            b = np.empty(4001, np.int64)
            a = np.empty(m, np.int32)
            assert hip.hipMemcpy(b.ctypes.data, pb, 8 * 4001, 2) == 0
            assert hip.hipMemcpy(a.ctypes.data, pa, 4 * m, 2) == 0
        finally:
            for p in (ps, pd, pb, pa):
                dg.device_free(p)
    wb, wa, _ = oracle.csr_build(4000, src, dst, L.CSR_SIMPLE)
    assert np.array_equal(b, wb) and np.array_equal(a, wa) and st["m"] == m


@pytest.mark.parametrize("reverse", [0, 1])
def test_cli_matches_reference_outputs(tmp_path, reverse):
    text = open(os.path.join(GOLD, "tuples.txt")).read()
    p = tmp_path / "t.txt"
    p.write_text(text)
    _say(f"[gpu] ent_convert reverse={reverse}")
    r = subprocess.run([CLI, str(p), str(reverse), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    ref = np.load(os.path.join(GOLD, f"ref_rev{reverse}.npz"))
    got = {k: np.fromfile(f"{p}_{k}.bin", dtype=np.int64) for k in ("beg_pos", "csr", "deg", "head", "weight")}
    for k in ("beg_pos", "csr", "deg", "weight"):
        assert np.array_equal(got[k], ref[k]), k
    rows = np.repeat(np.arange(got["beg_pos"].size - 1), np.diff(got["beg_pos"]))
    assert np.array_equal(got["head"], rows)
    # The files load as a graph through the reference's binary format.
    g = HostGraph.read_bin(f"{p}_beg_pos.bin", f"{p}_csr.bin")
    assert g.m == ref["csr"].size


def test_cli_simple_graph_options(oracle, tmp_path):
    text = "1 2\n2 2\n2 3\n1 2\n3 1\n"
    p = tmp_path / "s.txt"
    p.write_text(text)
    r = subprocess.run([CLI, str(p), "1", "0", "--no-loops", "--dedup", "--out", str(tmp_path / "o")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    g = HostGraph.read_bin(tmp_path / "o_beg_pos.bin", tmp_path / "o_csr.bin")
    b, a, _ = oracle.csr_build(3, [0, 1, 1, 0, 2], [1, 1, 2, 1, 0], L.CSR_SIMPLE)
    assert np.array_equal(g.beg_pos, b) and np.array_equal(g.adj, a)


def test_upload_edges_bfs(oracle):
    """A BFS handle built from tuples gives the oracle's depths on the host
    generator's graph."""
    src, dst = convert.rmat_tuples(14, 16, seed=14)
    h = HostGraph.rmat(14, 16, seed=14)
    with DeviceGraph.from_edges(1 << 14, src, dst) as dg:
        assert dg.m == h.m
        for r in h.pick_roots(4, 77):
            got, _ = dg.bfs(int(r))
            assert np.array_equal(got, oracle.bfs(h.beg_pos, h.adj, int(r))), r
