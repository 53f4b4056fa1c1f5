"""ctypes access to the CPU oracle (oracle/libent_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and the CPU-baseline
leg of bench.py as the checker, never as the measured or shipped path."""
import ctypes as C
import os

import numpy as np

ORACLE_SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libent_oracle.so")

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run __graft_entry__.build()")
        L = C.CDLL(ORACLE_SO)
        L.ent_oracle_bfs.restype = C.c_int32
        L.ent_oracle_bfs.argtypes = [C.c_int64, i64p, i32p, C.c_int32, i32p]
        L.ent_oracle_bfs_parents.restype = C.c_int32
        L.ent_oracle_bfs_parents.argtypes = [C.c_int64, i64p, i32p, C.c_int32, i32p, i32p]
        L.ent_oracle_bfs_do.restype = C.c_int32
        L.ent_oracle_bfs_do.argtypes = [C.c_int64, i64p, i32p, C.c_int32, i32p, C.c_int32]
        L.ent_oracle_validate.restype = C.c_int32
        L.ent_oracle_validate.argtypes = [C.c_int64, i64p, i32p, C.c_int32, i32p]
        L.ent_oracle_validate_spot.restype = C.c_int32
        L.ent_oracle_validate_spot.argtypes = [C.c_int64, i64p, i32p, i32p, C.c_int32]
        L.ent_oracle_validate_parents.restype = C.c_int32
        L.ent_oracle_validate_parents.argtypes = [C.c_int64, i64p, i32p, C.c_int32, i32p, i32p]
        L.ent_oracle_report.restype = None
        L.ent_oracle_report.argtypes = [C.c_int64, i64p, i32p, i64p, i64p]
        L.ent_oracle_tuples_parse.restype = C.c_int64
        L.ent_oracle_tuples_parse.argtypes = [C.c_char_p, C.c_int64, i64p, C.c_int64, i64p, i64p]
        L.ent_oracle_csr_build.restype = C.c_int64
        L.ent_oracle_csr_build.argtypes = [C.c_int64, C.c_int64, i32p, i32p, i32p, C.c_int32, i64p, i32p,
                                           i32p]
        self.L = L

    @staticmethod
    def _csr(beg_pos, adj):
        b = np.ascontiguousarray(beg_pos, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        if a.size == 0:
            a = np.zeros(1, dtype=np.int32)
        return b, a

    def bfs(self, beg_pos, adj, root):
        b, a = self._csr(beg_pos, adj)
        d = np.empty(b.size - 1, dtype=np.int32)
        lv = self.L.ent_oracle_bfs(b.size - 1, b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                   int(root), d.ctypes.data_as(i32p))
        assert lv >= 1
        return d

    def bfs_parents(self, beg_pos, adj, root):
        b, a = self._csr(beg_pos, adj)
        d = np.empty(b.size - 1, dtype=np.int32)
        p = np.empty(b.size - 1, dtype=np.int32)
        self.L.ent_oracle_bfs_parents(b.size - 1, b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                      int(root), d.ctypes.data_as(i32p), p.ctypes.data_as(i32p))
        return d, p

    def bfs_do(self, beg_pos, adj, root, threads=4):
        b, a = self._csr(beg_pos, adj)
        d = np.empty(b.size - 1, dtype=np.int32)
        self.L.ent_oracle_bfs_do(b.size - 1, b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                 int(root), d.ctypes.data_as(i32p), int(threads))
        return d

    def validate(self, beg_pos, adj, root, depth):
        b, a = self._csr(beg_pos, adj)
        d = np.ascontiguousarray(depth, dtype=np.int32)
        return self.L.ent_oracle_validate(b.size - 1, b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                          int(root), d.ctypes.data_as(i32p))

    def validate_spot(self, beg_pos, adj, depth, base_v):
        b, a = self._csr(beg_pos, adj)
        d = np.ascontiguousarray(depth, dtype=np.int32)
        return self.L.ent_oracle_validate_spot(b.size - 1, b.ctypes.data_as(i64p),
                                               a.ctypes.data_as(i32p), d.ctypes.data_as(i32p),
                                               int(base_v))

    def validate_parents(self, beg_pos, adj, root, depth, parent):
        b, a = self._csr(beg_pos, adj)
        d = np.ascontiguousarray(depth, dtype=np.int32)
        p = np.ascontiguousarray(parent, dtype=np.int32)
        return self.L.ent_oracle_validate_parents(b.size - 1, b.ctypes.data_as(i64p),
                                                  a.ctypes.data_as(i32p), int(root),
                                                  d.ctypes.data_as(i32p), p.ctypes.data_as(i32p))

    def report(self, beg_pos, depth):
        b = np.ascontiguousarray(beg_pos, dtype=np.int64)
        d = np.ascontiguousarray(depth, dtype=np.int32)
        e, v = C.c_int64(), C.c_int64()
        self.L.ent_oracle_report(b.size - 1, b.ctypes.data_as(i64p), d.ctypes.data_as(i32p),
                                 C.byref(e), C.byref(v))
        return e.value, v.value

    def tuples_parse(self, path, skip_lines=0):
        """(tokens int64[], vmin, vmax) as the reference converter reads a
        text tuple file; None when it cannot (too few lines to skip)."""
        lo, hi = C.c_int64(), C.c_int64()
        cnt = self.L.ent_oracle_tuples_parse(str(path).encode(), int(skip_lines), None, 0, C.byref(lo), C.byref(hi))
        if cnt < 0:
            return None
        t = np.empty(max(cnt, 1), dtype=np.int64)
        self.L.ent_oracle_tuples_parse(str(path).encode(), int(skip_lines), t.ctypes.data_as(i64p), cnt,
                                       C.byref(lo), C.byref(hi))
        return t[:cnt], lo.value, hi.value

    def csr_build(self, n, src, dst, flags, weight=None):
        """(beg_pos int64[n+1], adj int32[m], weight int32[m] or None)."""
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        ne = s.size
        b = np.empty(n + 1, dtype=np.int64)
        a = np.empty(max(2 * ne, 1), dtype=np.int32)
        w = None if weight is None else np.ascontiguousarray(weight, dtype=np.int32)
        wo = None if weight is None else np.empty(max(2 * ne, 1), dtype=np.int32)
        m = self.L.ent_oracle_csr_build(int(n), ne, s.ctypes.data_as(i32p) if ne else None,
                                        d.ctypes.data_as(i32p) if ne else None,
                                        w.ctypes.data_as(i32p) if w is not None and ne else None, int(flags),
                                        b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                        wo.ctypes.data_as(i32p) if wo is not None else None)
        assert m >= 0
        return b, a[:m], (None if wo is None else wo[:m])
