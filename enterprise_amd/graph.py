"""Host CSR graphs for the BFS drop-in: generators, the reference's binary CSR
files, and root pickers. Arrays are owned by libent_bfs.so and exposed as
numpy views (int64 beg_pos, int32 adj)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CSR, check, i32p, i64p, lib


class HostGraph:
    """A CSR graph in host memory (int64 offsets, int32 adjacency)."""

    def __init__(self):
        self._csr = CSR()
        self._owned = False

    @classmethod
    def _filled(cls, fill):
        g = cls()
        check(fill(C.byref(g._csr)))
        g._owned = True
        return g

    # -- constructors -----------------------------------------------------
    @classmethod
    def rmat(cls, scale, edgefactor=16, a=0.57, b=0.19, c=0.19, seed=1, relabel=True):
        """Seeded R-MAT, symmetrized, loops and duplicates removed."""
        return cls._filled(lambda p: lib().ent_gen_rmat(
            int(scale), int(edgefactor), float(a), float(b), float(c), int(seed),
            1 if relabel else 0, p))

    @classmethod
    def grid(cls, width, height, delete_ppm=20000, seed=4096, keep_lcc=True):
        """4-neighbour grid with seeded edge deletion (delete_ppm per million)."""
        return cls._filled(lambda p: lib().ent_gen_grid(
            int(width), int(height), int(delete_ppm), int(seed), 1 if keep_lcc else 0, p))

    @classmethod
    def from_edges(cls, n, src, dst):
        """Symmetrized, deduplicated, loop-free CSR from an edge list."""
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        return cls._filled(lambda p: lib().ent_csr_from_edges(
            int(n), int(s.size), s.ctypes.data_as(i32p), d.ctypes.data_as(i32p), p))

    @classmethod
    def from_arrays(cls, beg_pos, adj):
        """Copies an existing CSR (no symmetrization)."""
        b = np.ascontiguousarray(beg_pos, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        return cls._filled(lambda p: lib().ent_csr_from_arrays(
            int(b.size - 1), int(a.size), b.ctypes.data_as(i64p), a.ctypes.data_as(i32p), p))

    @classmethod
    def read_bin(cls, beg_file, csr_file):
        """Reads the reference's <name>_beg_pos.bin / <name>_csr.bin (int64)."""
        return cls._filled(lambda p: lib().ent_csr_read_bin(
            str(beg_file).encode(), str(csr_file).encode(), p))

    def write_bin(self, beg_file, csr_file):
        check(lib().ent_csr_write_bin(C.byref(self._csr), str(beg_file).encode(),
                                      str(csr_file).encode()))

    # -- views ------------------------------------------------------------
    @property
    def n(self):
        return int(self._csr.n)

    @property
    def m(self):
        return int(self._csr.m)

    @property
    def beg_pos(self):
        return np.ctypeslib.as_array(self._csr.beg_pos, shape=(self.n + 1,))

    @property
    def adj(self):
        if self.m == 0:
            return np.zeros(0, dtype=np.int32)
        return np.ctypeslib.as_array(self._csr.adj, shape=(self.m,))

    def degrees(self):
        return np.diff(self.beg_pos)

    def check(self, sorted_rows=True, no_loops=True, symmetric=True):
        flags = (1 if sorted_rows else 0) | (2 if no_loops else 0) | (4 if symmetric else 0)
        check(lib().ent_csr_check(C.byref(self._csr), flags))

    # -- roots ------------------------------------------------------------
    def pick_roots(self, k, seed):
        out = np.zeros(k, dtype=np.int32)
        check(lib().ent_pick_roots(C.byref(self._csr), int(k), int(seed), out.ctypes.data_as(i32p)))
        return out

    def pick_roots_glibc(self, k=64):
        out = np.zeros(k, dtype=np.int32)
        check(lib().ent_pick_roots_glibc(C.byref(self._csr), int(k), out.ctypes.data_as(i32p)))
        return out

    def pick_roots_mixed(self, k, k_hub, seed):
        out = np.zeros(k, dtype=np.int32)
        check(lib().ent_pick_roots_mixed(C.byref(self._csr), int(k), int(k_hub), int(seed),
                                         out.ctypes.data_as(i32p)))
        return out

    def free(self):
        if self._owned:
            lib().ent_csr_free(C.byref(self._csr))
            self._owned = False

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
