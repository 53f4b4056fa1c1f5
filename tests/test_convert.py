"""Edge-list to CSR conversion (include/ent_convert.h) on the CPU side: the
oracle's restatement of the reference converter pinned against the converter
itself (compiled from its source into oracle/_ref/text_to_bin) and its
committed outputs (tests/golden/convert), sorted and deduplicated rows against
scipy, and the library's host-side text reader, tuple generator and file
writer against the oracle."""
import ctypes
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np
import pytest
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(HERE, "golden"))

from enterprise_amd import EntError, HostGraph  # noqa: E402
from enterprise_amd import _lib as L  # noqa: E402
from enterprise_amd import convert  # noqa: E402
import make_convert_golden as mg  # noqa: E402

GOLD = os.path.join(HERE, "golden", "convert")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "text_to_bin")
need_ref = pytest.mark.skipif(not os.path.exists(REF_BIN), reason="oracle/_ref/text_to_bin not built")


def glibc_weights(k):
    """rand() % 63 + 1 from glibc's default seed, as the converter draws them."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    return np.array([libc.rand() % 63 + 1 for _ in range(k)], dtype=np.int32)


def oracle_convert(oracle, path, reverse, skip):
    toks, vmin, vmax = oracle.tuples_parse(path, skip)
    pairs = toks.reshape(-1, 2) - vmin
    n = int(vmax - vmin + 1)
    w = glibc_weights(pairs.shape[0])
    b, a, wo = oracle.csr_build(n, pairs[:, 0], pairs[:, 1], L.CSR_REVERSE if reverse else 0, weight=w)
    return n, b, a, wo


def rows_of(beg_pos):
    return np.repeat(np.arange(beg_pos.size - 1), np.diff(beg_pos))


@pytest.mark.parametrize("reverse", [0, 1])
def test_oracle_matches_golden_reference(oracle, reverse):
    ref = np.load(os.path.join(GOLD, f"ref_rev{reverse}.npz"))
    n, b, a, w = oracle_convert(oracle, os.path.join(GOLD, "tuples.txt"), reverse, 2)
    assert np.array_equal(ref["beg_pos"], b)
    assert np.array_equal(ref["csr"], a)
    assert np.array_equal(ref["deg"], np.diff(b))
    assert np.array_equal(ref["weight"], w)
    rows = rows_of(b)
    if reverse:  # the converter writes head only for forward entries, 0 elsewhere
        assert np.all((ref["head"] == rows) | (ref["head"] == 0))
        assert np.count_nonzero(ref["head"] == rows) >= a.size // 2
    else:
        assert np.array_equal(ref["head"], rows)


@need_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
@pytest.mark.parametrize("reverse", [0, 1])
@pytest.mark.parametrize("skip", [0, 2])
def test_oracle_matches_reference_converter(oracle, tmp_path, seed, reverse, skip):
    text = mg.tuples_text(seed=seed, n_tuples=40 + 30 * seed, n=11 * seed, base=seed)
    if skip == 0:
        text = "\n".join(text.split("\n")[2:])
    ref = mg.run_reference(text, reverse, skip)
    path = tmp_path / "t.txt"
    path.write_text(text)
    n, b, a, w = oracle_convert(oracle, path, reverse, skip)
    assert np.array_equal(ref["beg_pos"], b)
    assert np.array_equal(ref["csr"], a)
    assert np.array_equal(ref["deg"], np.diff(b))
    assert np.array_equal(ref["weight"], w)


@need_ref
def test_reversed_self_loop_differs_by_one_slot(oracle, tmp_path):
    """With reverse = 1 the converter writes both copies of a self-loop into
    one slot and leaves the next one unwritten (0 in a new file); the oracle
    (and the library) store the loop twice. Nothing else differs."""
    text = "1 2\n3 3\n2 3\n3 1\n"
    ref = mg.run_reference(text, 1, 0)
    path = tmp_path / "t.txt"
    path.write_text(text)
    n, b, a, w = oracle_convert(oracle, path, 1, 0)
    assert np.array_equal(ref["beg_pos"], b)
    diff = np.flatnonzero(ref["csr"] != a)
    # Row of id 3 (vertex 2): the loop twice, then 1 (from "2 3") and 0 (from "3 1").
    assert a[b[2]:b[3]].tolist() == [2, 2, 1, 0]
    assert diff.tolist() == [int(b[2]) + 1] and ref["csr"][diff[0]] == 0


def _random_tuples(seed, n, ne, loops=True):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, ne).astype(np.int32)
    dst = rng.integers(0, n, ne).astype(np.int32)
    if not loops:
        dst = np.where(dst == src, (dst + 1) % n, dst).astype(np.int32)
    dup = rng.integers(0, ne, ne // 8)
    return np.concatenate([src, src[dup]]), np.concatenate([dst, dst[dup]])


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_dedup_matches_scipy(oracle, seed):
    n = 50 + 40 * seed
    src, dst = _random_tuples(seed, n, 6 * n)
    b, a, _ = oracle.csr_build(n, src, dst, L.CSR_SIMPLE)
    m = sp.coo_matrix((np.ones(src.size), (src, dst)), shape=(n, n))
    m = (m + m.T).tocsr()
    m.setdiag(0)
    m.eliminate_zeros()
    m.sort_indices()
    assert np.array_equal(b, m.indptr.astype(np.int64))
    assert np.array_equal(a, m.indices.astype(np.int32))
    # Without the reversed copies: the directed simple graph.
    b1, a1, _ = oracle.csr_build(n, src, dst, L.CSR_NO_LOOPS | L.CSR_DEDUP)
    d = sp.coo_matrix((np.ones(src.size), (src, dst)), shape=(n, n)).tocsr()
    d.setdiag(0)
    d.eliminate_zeros()
    d.sort_indices()
    assert np.array_equal(b1, d.indptr.astype(np.int64)) and np.array_equal(a1, d.indices.astype(np.int32))


def test_oracle_input_order_and_sort_brute_force(oracle):
    """Tiny cases against a direct Python statement of each flag."""
    rng = np.random.default_rng(5)
    for trial in range(40):
        n = int(rng.integers(1, 9))
        ne = int(rng.integers(0, 14))
        src = rng.integers(0, n, ne).astype(np.int32)
        dst = rng.integers(0, n, ne).astype(np.int32)
        w = rng.integers(1, 64, ne).astype(np.int32)
        for flags in range(16):
            rows = [[] for _ in range(n)]
            for i in range(ne):
                u, v = int(src[i]), int(dst[i])
                if flags & L.CSR_NO_LOOPS and u == v:
                    continue
                rows[u].append((v, 2 * i, int(w[i])))
                if flags & L.CSR_REVERSE:
                    rows[v].append((u, 2 * i + 1, int(w[i])))
            want_b, want_a, want_w = [0], [], []
            for r in rows:
                if flags & (L.CSR_SORT | L.CSR_DEDUP):
                    r = sorted(r)
                if flags & L.CSR_DEDUP:
                    r = [x for k, x in enumerate(r) if k == 0 or x[0] != r[k - 1][0]]
                want_a += [x[0] for x in r]
                want_w += [x[2] for x in r]
                want_b.append(len(want_a))
            b, a, wo = oracle.csr_build(n, src, dst, flags, weight=w)
            assert b.tolist() == want_b and a.tolist() == want_a and wo.tolist() == want_w, (trial, flags)


def test_tuples_read_matches_oracle(oracle, tmp_path):
    cases = {
        "golden": (open(os.path.join(GOLD, "tuples.txt")).read(), 2),
        "crlf": ("5 7\r\n7 9\r\n9 5\r\n", 0),
        "no_trailing_newline": ("1 2\n2 3\n3 4", 0),
        "tabs_and_runs": ("10\t11\n11   12\t\n\n12 10\n", 0),
        "leading_separator": ("\n4 5\n6 4\n", 0),  # the scan's first token starts at offset 0
        "skip_with_leading_newline": ("\n#h\n1 2\n2 3\n", 1),
        "negative_ids": ("-3 4\n4 -1\n", 0),
    }
    for name, (text, skip) in cases.items():
        p = tmp_path / f"{name}.txt"
        p.write_text(text)
        toks, vmin, vmax = oracle.tuples_parse(p, skip)
        if toks.size % 2:
            with pytest.raises(EntError) as e:
                convert.read_tuples(p, skip)
            assert e.value.code == L.ENT_EINVAL, name
            continue
        n, vm, src, dst = convert.read_tuples(p, skip)
        assert vm == vmin and n == vmax - vmin + 1, name
        assert np.array_equal(src, toks[0::2] - vmin) and np.array_equal(dst, toks[1::2] - vmin), name


def test_tuples_read_large_file_threads(oracle, tmp_path):
    """A file large enough for the multi-threaded scan (many 64 KiB chunks)."""
    rng = np.random.default_rng(11)
    u = rng.integers(0, 1 << 20, 200_000)
    v = rng.integers(0, 1 << 20, 200_000)
    p = tmp_path / "big.txt"
    p.write_text("".join(f"{a} {b}\n" for a, b in zip(u, v)))
    toks, vmin, vmax = oracle.tuples_parse(p, 0)
    n, vm, src, dst = convert.read_tuples(p, 0)
    assert (n, vm) == (vmax - vmin + 1, vmin)
    assert np.array_equal(src, toks[0::2] - vmin) and np.array_equal(dst, toks[1::2] - vmin)


def test_tuples_read_errors(tmp_path):
    p = tmp_path / "odd.txt"
    p.write_text("1 2\n3\n")
    with pytest.raises(EntError) as e:
        convert.read_tuples(p, 0)
    assert e.value.code == L.ENT_EINVAL
    q = tmp_path / "short.txt"
    q.write_text("1 2\n")
    with pytest.raises(EntError) as e:
        convert.read_tuples(q, 5)
    assert e.value.code == L.ENT_EINVAL
    with pytest.raises(EntError) as e:
        convert.read_tuples(tmp_path / "missing.txt", 0)
    assert e.value.code == L.ENT_EIO


@pytest.mark.parametrize("scale,relabel", [(8, True), (11, True), (10, False)])
def test_rmat_tuples_build_the_generator_graph(oracle, scale, relabel):
    """The exported raw tuples, merged by the oracle, give exactly the CSR of
    the host generator with the same arguments."""
    src, dst = convert.rmat_tuples(scale, 16, seed=scale + 100, relabel=relabel)
    b, a, _ = oracle.csr_build(1 << scale, src, dst, L.CSR_SIMPLE)
    g = HostGraph.rmat(scale, 16, seed=scale + 100, relabel=relabel)
    assert np.array_equal(b, g.beg_pos) and np.array_equal(a, g.adj)
    assert np.count_nonzero(src == dst) > 0  # loops are part of the raw tuples


@pytest.mark.parametrize("reverse", [0, 1])
def test_write_ref_files_matches_golden(oracle, tmp_path, reverse):
    """The library's writer, given the oracle's CSR and weights, produces the
    converter's files (head: the row of every entry)."""
    ref = np.load(os.path.join(GOLD, f"ref_rev{reverse}.npz"))
    n, b, a, w = oracle_convert(oracle, os.path.join(GOLD, "tuples.txt"), reverse, 2)
    g = HostGraph.from_arrays(b, a)
    prefix = str(tmp_path / "out")
    convert.write_ref_files(g, prefix, weight=w)
    got = {k: np.fromfile(f"{prefix}_{k}.bin", dtype=np.int64) for k in ("beg_pos", "csr", "deg", "head", "weight")}
    for k in ("beg_pos", "csr", "deg", "weight"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["head"], rows_of(b))
    if not reverse:
        assert np.array_equal(got["head"], ref["head"])
