"""Edge lists to CSR on the GPU and the reference's text tuple format
(include/ent_convert.h), the replacement of the reference's offline converter
tuple_text_to_binary_csr/tuple_text_to_bin.cpp:56-315."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import BuildStats, Tuples, check, i32p, lib
from .graph import HostGraph

REVERSE, NO_LOOPS, SORT, DEDUP, SIMPLE = L.CSR_REVERSE, L.CSR_NO_LOOPS, L.CSR_SORT, L.CSR_DEDUP, L.CSR_SIMPLE


def read_tuples(path, skip_lines=0):
    """(n, vmin, src int32[ne], dst int32[ne]) of a text tuple file, read as
    the reference converter reads it (ids stored minus vmin)."""
    t = Tuples()
    check(lib().ent_tuples_read(str(path).encode(), int(skip_lines), C.byref(t)))
    try:
        ne = int(t.ne)
        src = np.ctypeslib.as_array(t.src, shape=(ne,)).copy() if ne else np.zeros(0, np.int32)
        dst = np.ctypeslib.as_array(t.dst, shape=(ne,)).copy() if ne else np.zeros(0, np.int32)
        return int(t.n), int(t.vmin), src, dst
    finally:
        lib().ent_tuples_free(C.byref(t))


def rmat_tuples(scale, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, relabel=True):
    """The raw tuples behind HostGraph.rmat with the same arguments."""
    ne = (1 << int(scale)) * int(edgefactor)
    src = np.empty(ne, dtype=np.int32)
    dst = np.empty(ne, dtype=np.int32)
    check(lib().ent_gen_rmat_tuples(int(scale), int(edgefactor), float(a), float(b), float(c), int(seed),
                                    1 if relabel else 0, src.ctypes.data_as(i32p), dst.ctypes.data_as(i32p)))
    return src, dst


def csr_build(n, src, dst, flags, weight=None, device=0):
    """(HostGraph, weights int32[m] or None, stats) built on GPU `device`."""
    s = np.ascontiguousarray(src, dtype=np.int32)
    d = np.ascontiguousarray(dst, dtype=np.int32)
    if s.shape != d.shape:
        raise ValueError("src and dst differ in length")
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.int32)
    g = HostGraph()
    wp = i32p()
    st = BuildStats()
    check(lib().ent_csr_build(int(device), int(n), int(s.size), s.ctypes.data_as(i32p), d.ctypes.data_as(i32p),
                              w.ctypes.data_as(i32p) if w is not None else None, int(flags), C.byref(g._csr),
                              C.byref(wp) if w is not None else None, C.byref(st)))
    g._owned = True
    wout = None
    if w is not None:
        wout = np.ctypeslib.as_array(wp, shape=(g.m,)).copy() if g.m else np.zeros(0, np.int32)
        C.CDLL(None).free(wp)
    return g, wout, st.as_dict()


def csr_build_device(n, ne, src_dev, dst_dev, flags, beg_pos_dev, adj_dev, adj_capacity, device=0):
    """Device arrays in and out; returns (m, stats)."""
    m = C.c_int64()
    st = BuildStats()
    check(lib().ent_csr_build_device(int(device), int(n), int(ne), C.c_void_p(src_dev), C.c_void_p(dst_dev),
                                     int(flags), C.c_void_p(beg_pos_dev), C.c_void_p(adj_dev),
                                     int(adj_capacity), C.byref(m), C.byref(st)))
    return m.value, st.as_dict()


def write_ref_files(graph, prefix, weight=None):
    """The reference converter's output files for `prefix`."""
    w = None if weight is None else np.ascontiguousarray(weight, dtype=np.int32)
    check(lib().ent_csr_write_ref_files(C.byref(graph._csr), w.ctypes.data_as(i32p) if w is not None else None,
                                        str(prefix).encode()))
