"""The C ABI: the library loads without a GPU, exports every function the
public headers declare, and rejects bad arguments with error codes (no
device work is started here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from enterprise_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        with open(os.path.join(ROOT, "include", h)) as f:
            text = f.read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ent_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing


def test_signature_table_covers_headers():
    covered = {n for n, _, _ in L.SIGNATURES}
    assert declared_functions() <= covered


def test_default_config():
    from enterprise_amd import default_config
    c = default_config()
    assert c.block_size in (64, 128, 256, 512, 1024)
    assert 1 <= c.t_wave <= c.t_block <= c.t_grid and c.piece <= c.t_grid
    assert c.hub_cache == 0 or 6 <= c.hub_cache <= 13
    assert c.hub_td in (0, 1)


def test_upload_rejects_malformed_csr():
    lib = L.lib()
    h = C.c_void_p()
    beg = np.array([0, 2, 1], dtype=np.int64)     # decreasing offsets
    adj = np.array([1, 0], dtype=np.int32)
    rc = lib.ent_graph_upload(beg.ctypes.data_as(L.i64p), adj.ctypes.data_as(L.i32p), 2, 2, 0,
                              C.byref(h))
    assert rc == L.ENT_EINVAL
    beg = np.array([0, 1, 2], dtype=np.int64)
    adj = np.array([1, 7], dtype=np.int32)         # id out of range
    rc = lib.ent_graph_upload(beg.ctypes.data_as(L.i64p), adj.ctypes.data_as(L.i32p), 2, 2, 0,
                              C.byref(h))
    assert rc == L.ENT_EINVAL and b"malformed" in lib.ent_last_error()
    beg = np.array([0, 1, 3], dtype=np.int64)      # offsets end past m
    rc = lib.ent_graph_upload(beg.ctypes.data_as(L.i64p), adj.ctypes.data_as(L.i32p), 2, 2, 0,
                              C.byref(h))
    assert rc == L.ENT_EINVAL
    rc = lib.ent_graph_upload(beg.ctypes.data_as(L.i64p), adj.ctypes.data_as(L.i32p), 2, 3, -1,
                              C.byref(h))
    assert rc == L.ENT_EINVAL


def test_null_handles_rejected():
    lib = L.lib()
    st = L.Stats()
    assert lib.ent_bfs(None, 0, None, C.byref(st)) == L.ENT_EINVAL
    assert lib.ent_graph_free(None) == L.ENT_OK
    assert lib.ent_set_config(None, None) == L.ENT_EINVAL
