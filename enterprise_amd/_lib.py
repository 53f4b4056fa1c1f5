"""ctypes binding of libent_bfs.so (the C ABI declared in include/ent_bfs.h,
include/ent_graph.h, include/ent_batch.h and include/ent_convert.h). No torch types cross the boundary: plain pointers and
sizes only.

The library is loaded from enterprise_amd/lib/ (built in-tree by
__graft_entry__.build()); a missing library raises instead of falling back to
any other implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libent_bfs.so")

ENT_OK = 0
ENT_EINVAL = -1
ENT_ENOMEM = -2
ENT_EDEVICE = -3
ENT_EINTERNAL = -4
ENT_EIO = -5
ENT_EBUSY = -6
ENT_ECANCELED = -7

DIR_AUTO, DIR_REFERENCE, DIR_TOPDOWN, DIR_BOTTOMUP = 0, 1, 2, 3
LAUNCH_PERSISTENT, LAUNCH_LEVELS, LAUNCH_GRAPH, LAUNCH_STREAM = 0, 1, 2, 3
UPLOAD_KEEP_ORDER = 1
CSR_REVERSE, CSR_NO_LOOPS, CSR_SORT, CSR_DEDUP = 1, 2, 4, 8
CSR_SIMPLE = CSR_REVERSE | CSR_NO_LOOPS | CSR_DEDUP

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)


class CSR(C.Structure):
    _fields_ = [("n", C.c_int64), ("m", C.c_int64),
                ("beg_pos", i64p), ("adj", i32p)]


class Tuples(C.Structure):
    _fields_ = [("n", C.c_int64), ("ne", C.c_int64), ("vmin", C.c_int64),
                ("src", i32p), ("dst", i32p)]


class BuildStats(C.Structure):
    _fields_ = [("records", C.c_int64), ("m", C.c_int64), ("key_bits", C.c_int32),
                ("pad", C.c_int32), ("device_ms", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


class BfsConfig(C.Structure):
    _fields_ = [(name, C.c_int32) for name in (
        "direction", "launch", "alpha", "beta", "t_wave", "t_block", "t_grid",
        "piece", "block_size", "blocks_per_cu", "bu_probe", "bu_chunk_words",
        "graph_levels", "queue_small", "batch_alpha", "grid_blocks", "hub_cache", "hub_td", "queue_max_degree", "count_examined",
        "batch_fused")]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("levels", C.c_int32), ("td_levels", C.c_int32),
                ("bu_levels", C.c_int32), ("phases", C.c_int32),
                ("reached_vertices", C.c_int64), ("reached_edges", C.c_int64),
                ("alg_bytes", C.c_int64), ("device_ms", C.c_double),
                ("host_ms", C.c_double), ("bu_level_mask", C.c_uint64),
                ("grid_blocks", C.c_int32), ("block_size", C.c_int32),
                ("examined_edges", C.c_int64), ("examined_bytes", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class LevelRecord(C.Structure):
    _fields_ = [("launch", C.c_int32), ("level", C.c_int32), ("phase", C.c_int32), ("pad", C.c_int32),
                ("frontier_vertices", C.c_int64), ("frontier_edges", C.c_int64)]


PHASE_NAMES = {0: "top-down", 1: "bottom-up", 2: "inspect", 3: "top-down queue", 4: "rebuild"}


class BatchStats(C.Structure):
    _fields_ = [("k", C.c_int32), ("words", C.c_int32), ("levels", C.c_int32),
                ("phases", C.c_int32),
                ("reached_edges", C.c_int64), ("reached_vertices", C.c_int64),
                ("alg_bytes", C.c_int64), ("device_ms", C.c_double),
                ("host_ms", C.c_double), ("bu_level_mask", C.c_uint64),
                ("td_level_mask", C.c_uint64), ("flushes", C.c_int32), ("pad", C.c_int32)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_ if f != "pad"}


# (name, restype, argtypes) for every exported function.
SIGNATURES = [
    ("ent_last_error", C.c_char_p, []),
    ("ent_set_host_threads", None, [C.c_int32]),
    ("ent_gen_rmat", C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double,
                               C.c_double, C.c_uint64, C.c_int32, C.POINTER(CSR)]),
    ("ent_gen_grid", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_uint64,
                               C.c_int32, C.POINTER(CSR)]),
    ("ent_csr_from_edges", C.c_int, [C.c_int64, C.c_int64, i32p, i32p, C.POINTER(CSR)]),
    ("ent_csr_from_arrays", C.c_int, [C.c_int64, C.c_int64, i64p, i32p, C.POINTER(CSR)]),
    ("ent_csr_free", None, [C.POINTER(CSR)]),
    ("ent_csr_check", C.c_int, [C.POINTER(CSR), C.c_int32]),
    ("ent_csr_read_bin", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(CSR)]),
    ("ent_csr_write_bin", C.c_int, [C.POINTER(CSR), C.c_char_p, C.c_char_p]),
    ("ent_pick_roots", C.c_int, [C.POINTER(CSR), C.c_int32, C.c_uint64, i32p]),
    ("ent_pick_roots_glibc", C.c_int, [C.POINTER(CSR), C.c_int32, i32p]),
    ("ent_pick_roots_mixed", C.c_int, [C.POINTER(CSR), C.c_int32, C.c_int32,
                                       C.c_uint64, i32p]),
    ("ent_bfs_default_config", None, [C.POINTER(BfsConfig)]),
    ("ent_graph_upload", C.c_int, [i64p, i32p, C.c_int64, C.c_int64, C.c_int,
                                   C.POINTER(C.c_void_p)]),
    ("ent_graph_upload_ex", C.c_int, [i64p, i32p, C.c_int64, C.c_int64, C.c_int, C.c_int32,
                                      C.POINTER(C.c_void_p)]),
    ("ent_graph_free", C.c_int, [C.c_void_p]),
    ("ent_graph_download", C.c_int, [C.c_void_p, i64p, i32p]),
    ("ent_level_trace", C.c_int, [C.c_void_p, C.POINTER(LevelRecord), C.c_int32, C.POINTER(C.c_int32)]),
    ("ent_bfs_parents_device", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    ("ent_bfs_parents", C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, C.POINTER(Stats)]),
    ("ent_graph_share", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("ent_graph_info", C.c_int, [C.c_void_p, i64p, i64p, i64p, i64p]),
    ("ent_set_config", C.c_int, [C.c_void_p, C.POINTER(BfsConfig)]),
    ("ent_get_config", C.c_int, [C.c_void_p, C.POINTER(BfsConfig)]),
    ("ent_bfs", C.c_int, [C.c_void_p, C.c_int32, i32p, C.POINTER(Stats)]),
    ("ent_bfs_device", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(Stats)]),
    ("ent_depth_device_ptr", C.c_void_p, [C.c_void_p]),
    ("ent_copy_to_host", C.c_int, [C.c_void_p, C.c_void_p, i32p, C.c_int64]),
    ("ent_bfs_gpu_coalescing_mem", C.c_int, [i32p, C.c_int32, i64p, i32p,
                                             C.c_int64, C.c_int64, C.c_int32]),
    ("ent_selftest_wave", C.c_int, [C.c_int]),
    ("ent_calib_copy", C.c_int, [C.c_int, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    ("ent_batch_reserve", C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    ("ent_batch_release", C.c_int, [C.c_void_p]),
    ("ent_batch_workspace_bytes", C.c_int64, [C.c_void_p, C.c_int32, C.c_int32]),
    ("ent_bfs_batch_device", C.c_int, [C.c_void_p, i32p, C.c_int32, C.c_void_p,
                                       C.POINTER(BatchStats)]),
    ("ent_bfs_batch", C.c_int, [C.c_void_p, i32p, C.c_int32, i32p, C.POINTER(BatchStats)]),
    ("ent_device_alloc", C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_void_p)]),
    ("ent_device_free", C.c_int, [C.c_void_p, C.c_void_p]),
    ("ent_device_row_hash", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                      C.POINTER(C.c_uint64)]),
    ("ent_device_row_hash_noalloc", C.c_int, [C.c_int32, C.c_void_p, C.c_int64, C.c_int64,
                                              C.POINTER(C.c_uint64)]),
    ("ent_bfs_validate_device", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, i64p]),
    ("ent_last_device_ms", C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    ("ent_tuples_read", C.c_int, [C.c_char_p, C.c_int64, C.POINTER(Tuples)]),
    ("ent_tuples_free", None, [C.POINTER(Tuples)]),
    ("ent_gen_rmat_tuples", C.c_int, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double,
                                      C.c_uint64, C.c_int32, i32p, i32p]),
    ("ent_csr_build", C.c_int, [C.c_int32, C.c_int64, C.c_int64, i32p, i32p, i32p, C.c_int32,
                                C.POINTER(CSR), C.POINTER(i32p), C.POINTER(BuildStats)]),
    ("ent_csr_build_device", C.c_int, [C.c_int32, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.POINTER(C.c_int64), C.POINTER(BuildStats)]),
    ("ent_graph_upload_edges", C.c_int, [i32p, i32p, C.c_int64, C.c_int64, C.c_int32, C.c_int32,
                                         C.c_int32, C.POINTER(C.c_void_p)]),
    ("ent_csr_write_ref_files", C.c_int, [C.POINTER(CSR), i32p, C.c_char_p]),
]

_lib = None


class EntError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Loads libent_bfs.so once (the in-tree build, or the file named by the
    ENT_BFS_LIB environment variable); raises if it has not been built."""
    global _lib
    if _lib is None:
        path = os.environ.get("ENT_BFS_LIB") or LIB_PATH
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is missing: build it with __graft_entry__.build()")
        handle = C.CDLL(path, mode=C.RTLD_GLOBAL)
        for name, res, args in SIGNATURES:
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc):
    if rc != ENT_OK:
        msg = lib().ent_last_error()
        raise EntError(rc, msg.decode() if msg else "")
    return rc
