"""Device-resident graph handle and the single-source BFS drop-in.

Mirrors the reference's entry point bfs_gpu_coalescing_mem
(bfs_gpu_opt.cuh:293-447): a CSR is uploaded once (allocator.cuh:33-151) and
each call runs one level-synchronous direction-optimizing BFS
(bfs_tdbu_clfy_sort, bfs_gpu_opt.cuh:40-288) on the GPU through the C ABI of
libent_bfs.so. Depth is int32 with -1 for unreached (the reference's
unsigned char 255)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import BatchStats, BfsConfig, Stats, check, i32p, i64p, lib


def default_config():
    cfg = BfsConfig()
    lib().ent_bfs_default_config(C.byref(cfg))
    return cfg


class DeviceGraph:
    """A CSR graph resident in HBM plus its BFS workspace."""

    def __init__(self, host_graph=None, beg_pos=None, adj=None, device=0, share=None, keep_order=False):
        """keep_order: keep the caller's adjacency order on the device (by
        default rows are reordered by descending neighbour degree class)."""
        if share is not None:
            # Second handle on share's resident CSR, with its own workspace.
            h = C.c_void_p()
            check(lib().ent_graph_share(share.handle, C.byref(h)))
            self._h = h
            self.n, self.m, self.device = share.n, share.m, share.device
            return
        if host_graph is not None:
            beg_pos, adj = host_graph.beg_pos, host_graph.adj
        b = np.ascontiguousarray(beg_pos, dtype=np.int64)
        a = np.ascontiguousarray(adj, dtype=np.int32)
        self.n = int(b.size - 1)
        self.m = int(a.size)
        self.device = device
        h = C.c_void_p()
        check(lib().ent_graph_upload_ex(b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
                                        self.n, self.m, int(device),
                                        L.UPLOAD_KEEP_ORDER if keep_order else 0, C.byref(h)))
        self._h = h

    @classmethod
    def from_edges(cls, n, src, dst, flags=L.CSR_SIMPLE, device=0, keep_order=False):
        """Handle whose CSR is built on the device from tuples (ent_convert.h)."""
        s = np.ascontiguousarray(src, dtype=np.int32)
        d = np.ascontiguousarray(dst, dtype=np.int32)
        if s.shape != d.shape:
            raise ValueError("src and dst differ in length")
        h = C.c_void_p()
        check(lib().ent_graph_upload_edges(s.ctypes.data_as(i32p), d.ctypes.data_as(i32p), int(n), int(s.size),
                                           int(flags), int(device), L.UPLOAD_KEEP_ORDER if keep_order else 0,
                                           C.byref(h)))
        g = cls.__new__(cls)
        g._h = h
        info = g.info()
        g.n, g.m, g.device = info["n"], info["m"], device
        return g

    @property
    def handle(self):
        if not self._h:
            raise RuntimeError("graph already freed")
        return self._h

    def info(self):
        n, m, na, nb = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().ent_graph_info(self.handle, C.byref(n), C.byref(m), C.byref(na), C.byref(nb)))
        return {"n": n.value, "m": m.value, "n_active": na.value, "device_bytes": nb.value}

    @property
    def config(self):
        cfg = BfsConfig()
        check(lib().ent_get_config(self.handle, C.byref(cfg)))
        return cfg

    def configure(self, **kw):
        cfg = self.config
        for k, v in kw.items():
            if not hasattr(cfg, k):
                raise KeyError(k)
            setattr(cfg, k, int(v))
        check(lib().ent_set_config(self.handle, C.byref(cfg)))
        return cfg

    def set_config(self, cfg):
        check(lib().ent_set_config(self.handle, C.byref(cfg)))

    def bfs(self, root, copy=True):
        """One BFS; returns (depth int32[n] or None, stats dict)."""
        st = Stats()
        out = np.empty(self.n, dtype=np.int32) if copy else None
        ptr = out.ctypes.data_as(i32p) if copy else None
        check(lib().ent_bfs(self.handle, int(root), ptr, C.byref(st)))
        return out, st.as_dict()

    def bfs_parents(self, root):
        """One BFS; returns (depth int32[n], parent int32[n], stats)."""
        d = np.empty(self.n, dtype=np.int32)
        p = np.empty(self.n, dtype=np.int32)
        st = Stats()
        check(lib().ent_bfs_parents(self.handle, int(root), d.ctypes.data_as(i32p), p.ctypes.data_as(i32p),
                                    C.byref(st)))
        return d, p, st.as_dict()

    def parents_device(self, root, depth_dev, parent_dev):
        """Parents from a device depth array into a device parent array."""
        check(lib().ent_bfs_parents_device(self.handle, int(root), C.c_void_p(depth_dev),
                                           C.c_void_p(parent_dev)))

    def bfs_device(self, root, depth_dev=None):
        """One BFS writing to device memory (default: the handle's buffer)."""
        st = Stats()
        check(lib().ent_bfs_device(self.handle, int(root),
                                   C.c_void_p(depth_dev) if depth_dev else None, C.byref(st)))
        return st.as_dict()

    def last_device_ms(self):
        """HIP-event device time of the last search (waits for its launches)."""
        ms = C.c_double()
        check(lib().ent_last_device_ms(self.handle, C.byref(ms)))
        return ms.value

    def validate_device(self, root, depth_dev=None):
        """All-vertex BFS check of a device depth array (default: the
        handle's): [root depth != 0, bad values, vertices with a bad edge,
        reached vertices without a parent]; all zero iff it is the BFS from
        root."""
        out = np.zeros(4, dtype=np.int64)
        ptr = depth_dev if depth_dev is not None else self.depth_device_ptr()
        check(lib().ent_bfs_validate_device(self.handle, int(root), C.c_void_p(ptr), out.ctypes.data_as(i64p)))
        return [int(x) for x in out]

    def level_trace(self):
        """Per-launch records of the last search run with count_examined=1:
        list of dicts (launch, level, phase name, frontier vertices and edges)."""
        recs = (L.LevelRecord * 1024)()
        cnt = C.c_int32()
        check(lib().ent_level_trace(self.handle, recs, 1024, C.byref(cnt)))
        return [{"launch": r.launch, "level": r.level, "phase": L.PHASE_NAMES.get(r.phase, str(r.phase)),
                 "frontier_vertices": r.frontier_vertices, "frontier_edges": r.frontier_edges}
                for r in recs[:cnt.value]]

    def download(self):
        """(beg_pos int64[n+1], adj int32[m]) as held on the device."""
        b = np.empty(self.n + 1, dtype=np.int64)
        a = np.empty(max(self.m, 1), dtype=np.int32)
        check(lib().ent_graph_download(self.handle, b.ctypes.data_as(i64p), a.ctypes.data_as(i32p)))
        return b, a[:self.m]

    def depth_device_ptr(self):
        return lib().ent_depth_device_ptr(self.handle)

    def copy_depth(self, src_dev=None, count=None):
        count = self.n if count is None else count
        out = np.empty(count, dtype=np.int32)
        src = src_dev if src_dev is not None else self.depth_device_ptr()
        check(lib().ent_copy_to_host(self.handle, C.c_void_p(src), out.ctypes.data_as(i32p), count))
        return out

    # -- multi-source ------------------------------------------------------
    def batch_reserve(self, k, max_levels=0):
        check(lib().ent_batch_reserve(self.handle, int(k), int(max_levels)))

    def batch_release(self):
        check(lib().ent_batch_release(self.handle))

    def bfs_batch(self, roots):
        """Multi-source BFS; returns (depths int32[k, n], stats)."""
        r = np.ascontiguousarray(roots, dtype=np.int32)
        out = np.empty((r.size, self.n), dtype=np.int32)
        st = BatchStats()
        check(lib().ent_bfs_batch(self.handle, r.ctypes.data_as(i32p), int(r.size),
                                  out.ctypes.data_as(i32p), C.byref(st)))
        return out, st.as_dict()

    def bfs_batch_device(self, roots, depths_dev):
        """Multi-source BFS writing k*n int32 depths to device memory."""
        r = np.ascontiguousarray(roots, dtype=np.int32)
        st = BatchStats()
        check(lib().ent_bfs_batch_device(self.handle, r.ctypes.data_as(i32p), int(r.size),
                                         C.c_void_p(depths_dev), C.byref(st)))
        return st.as_dict()

    def device_alloc(self, nbytes):
        p = C.c_void_p()
        check(lib().ent_device_alloc(self.handle, int(nbytes), C.byref(p)))
        return p.value

    def device_free(self, ptr):
        check(lib().ent_device_free(self.handle, C.c_void_p(ptr)))

    def row_hash(self, dev_ptr, rows, cols):
        """Position-keyed 64-bit checksum of each row of an int32 device array."""
        out = np.empty(rows, dtype=np.uint64)
        check(lib().ent_device_row_hash(self.handle, C.c_void_p(dev_ptr), int(rows), int(cols),
                                        out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def free(self):
        if getattr(self, "_h", None):
            lib().ent_graph_free(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bfs_gpu_coalescing_mem(src_list, beg_pos, csr, gpu_id=0):
    """The reference's entry point: per-root BFS with TEPS printout."""
    s = np.ascontiguousarray(src_list, dtype=np.int32)
    b = np.ascontiguousarray(beg_pos, dtype=np.int64)
    a = np.ascontiguousarray(csr, dtype=np.int32)
    return check(lib().ent_bfs_gpu_coalescing_mem(
        s.ctypes.data_as(i32p), int(s.size), b.ctypes.data_as(i64p), a.ctypes.data_as(i32p),
        int(b.size - 1), int(a.size), int(gpu_id)))


def row_hash_host(a):
    """Host form of the device row checksum (sum over i of mix64((i << 32) ^ value))."""
    a = np.ascontiguousarray(a, dtype=np.int32)
    idx = np.arange(a.shape[-1], dtype=np.uint64)
    z = (idx << np.uint64(32)) ^ a.view(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        return z.sum(axis=-1, dtype=np.uint64)


def selftest_wave(device=0):
    return check(lib().ent_selftest_wave(int(device)))
