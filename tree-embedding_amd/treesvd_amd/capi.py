"""ctypes binding of the C ABI in include/treesvd.h (libtreesvd.so).

The library is the product: HIP kernels behind an extern "C" surface.  This
module only marshals arguments; nothing here computes.  Loading fails loudly
when libtreesvd.so has not been built (python tree-embedding_amd/build.py).
"""
from __future__ import annotations

import ctypes
import os
import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtreesvd.so")

TSVD_DIR_FWD, TSVD_DIR_TRANSPOSE = 0, 1
TSVD_REFRESH_INIT, TSVD_REFRESH_DYNAMIC = 0, 1
TSVD_FP32, TSVD_FP64 = 0, 1
TSVD_TRANSFORM_NC, TSVD_TRANSFORM_LP = 0, 1

# Every symbol include/treesvd.h declares.
EXPORTS = [
    "tsvd_last_error", "tsvd_device_info", "tsvd_device_alloc", "tsvd_device_free", "tsvd_memcpy",
    "tsvd_profile_enable", "tsvd_profile_reset", "tsvd_profile_read", "tsvd_gemm_tuning", "tsvd_gemm_tuning_read",
    "tsvd_graph_create", "tsvd_graph_destroy", "tsvd_graph_append", "tsvd_graph_num_edges",
    "tsvd_graph_degrees", "tsvd_graph_export_csr", "tsvd_graph_set_staging_limit", "tsvd_read_edge_file", "tsvd_read_subset_file",
    "tsvd_read_lines", "tsvd_write_embedding_csv",
    "tsvd_ppr_create", "tsvd_ppr_destroy", "tsvd_ppr_refresh", "tsvd_ppr_push", "tsvd_ppr_read",
    "tsvd_ppr_last_push_stats", "tsvd_ppr_set_two_sided", "tsvd_ppr_refresh_edges",
    "tsvd_ppr_get_push_setting", "tsvd_ppr_set_push_setting", "tsvd_ppr_set_graphs", "tsvd_ppr_graph_stats",
    "frPCA", "tsvd_frpca_csr", "tsvd_frpca_csr_batch", "tsvd_frpca_csr_noalloc",
    "tsvd_pipeline_create", "tsvd_pipeline_destroy", "tsvd_pipeline_snapshot",
    "tsvd_pipeline_embedding", "tsvd_pipeline_root_sigma", "tsvd_pipeline_graph", "tsvd_pipeline_ppr",
    "tsvd_pipeline_block_matrix", "tsvd_pipeline_block_norms", "tsvd_pipeline_block_transform", "tsvd_pipeline_node", "tsvd_pipeline_block_columns",
    "tsvd_pipeline_dense_gram", "tsvd_pipeline_gram_apply", "tsvd_pipeline_right_embedding",
    "tsvd_pipeline_lp_scores", "tsvd_pipeline_source_state", "tsvd_pipeline_snapshot_tuned",
    "tsvd_server_create", "tsvd_server_destroy", "tsvd_server_submit", "tsvd_server_wait",
    "tsvd_server_pipeline", "tsvd_server_edge_budget", "tsvd_probe_atomic_rate", "tsvd_probe_atomic_lanes", "tsvd_gen_rmat",
    "tsvd_gen_chung_lu", "tsvd_eig_batched",
]


class TreeSVDError(RuntimeError):
    pass


class Mat(ctypes.Structure):
    _fields_ = [("nrows", ctypes.c_int), ("ncols", ctypes.c_int), ("d", ctypes.POINTER(ctypes.c_double))]


class MatCSR(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int), ("nrows", ctypes.c_int), ("ncols", ctypes.c_int),
                ("values", ctypes.POINTER(ctypes.c_double)), ("cols", ctypes.POINTER(ctypes.c_int)),
                ("pointerB", ctypes.POINTER(ctypes.c_int)), ("pointerE", ctypes.POINTER(ctypes.c_int))]


class Config(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("directed", ctypes.c_int32), ("nsrc", ctypes.c_int32),
                ("sources", ctypes.POINTER(ctypes.c_int32)), ("alpha", ctypes.c_double),
                ("rmax", ctypes.c_double), ("precision", ctypes.c_int32), ("nparts", ctypes.c_int32),
                ("fanout", ctypes.c_int32), ("rank", ctypes.c_int32), ("q", ctypes.c_int32),
                ("transform", ctypes.c_int32), ("delta", ctypes.c_double), ("dyn_start", ctypes.c_int32),
                ("lazy_start", ctypes.c_int32), ("slack", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("two_sided", ctypes.c_int32), ("right_embedding", ctypes.c_int32),
                ("sparse_state", ctypes.c_int32), ("sparse_groups", ctypes.c_int32),
                ("merge_bf16x3", ctypes.c_int32)]


class SnapshotStats(ctypes.Structure):
    _fields_ = [("snapshot", ctypes.c_int32), ("ndirty", ctypes.c_int32), ("nmerged", ctypes.c_int32),
                ("reused", ctypes.c_int32), ("push_rounds", ctypes.c_int64), ("push_pops", ctypes.c_int64),
                ("push_relax", ctypes.c_int64), ("push_items", ctypes.c_int64),
                ("push_wave_relax", ctypes.c_int64), ("ms_total", ctypes.c_double), ("ms_ingest", ctypes.c_double),
                ("ms_refresh", ctypes.c_double), ("ms_push", ctypes.c_double), ("ms_blocks", ctypes.c_double),
                ("ms_frpca", ctypes.c_double), ("ms_merge", ctypes.c_double),
                ("state_entries", ctypes.c_int64), ("state_batches", ctypes.c_int32),
                ("push_line_relax", ctypes.c_int64), ("norm_items", ctypes.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PushSetting(ctypes.Structure):
    _fields_ = [("hub_chunk", ctypes.c_int32), ("ride_frac", ctypes.c_double), ("items_per_wave", ctypes.c_int32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("max_step_edges", ctypes.c_int64), ("max_step_requests", ctypes.c_int32),
                ("no_overlap", ctypes.c_int32)]


class RequestResult(ctypes.Structure):
    _fields_ = [("ticket", ctypes.c_int64), ("snapshot", ctypes.c_int32), ("step_requests", ctypes.c_int32),
                ("step_edges", ctypes.c_int64), ("last_ticket", ctypes.c_int64), ("submit_ms", ctypes.c_double),
                ("start_ms", ctypes.c_double), ("done_ms", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("step", SnapshotStats)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "step"}
        d["step"] = self.step.as_dict()
        return d


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise TreeSVDError(f"{LIB_PATH} is not built; run python tree-embedding_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, i64, f64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
        L.tsvd_last_error.restype = ctypes.c_char_p
        L.tsvd_device_info.argtypes = [ctypes.c_char_p, ctypes.c_int, vp, vp]
        L.tsvd_device_alloc.argtypes = [i64, vp]
        L.tsvd_device_free.argtypes = [vp]
        L.tsvd_memcpy.argtypes = [vp, vp, i64, ctypes.c_int]
        L.tsvd_profile_enable.argtypes = [ctypes.c_int]
        L.tsvd_profile_read.argtypes = [ctypes.c_char_p, ctypes.c_int]
        if hasattr(L, "tsvd_gemm_tuning"):
            L.tsvd_gemm_tuning.argtypes = [ctypes.c_int]
            L.tsvd_gemm_tuning_read.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.tsvd_graph_create.argtypes = [i32, ctypes.c_int, f64, vp]
        L.tsvd_graph_destroy.argtypes = [vp]
        L.tsvd_graph_append.argtypes = [vp, vp, vp, i64, ctypes.c_int]
        L.tsvd_graph_num_edges.restype = i64
        L.tsvd_graph_set_staging_limit.argtypes = [vp, i64]
        L.tsvd_graph_num_edges.argtypes = [vp]
        L.tsvd_graph_degrees.argtypes = [vp, ctypes.c_int, vp, vp]
        L.tsvd_graph_export_csr.argtypes = [vp, ctypes.c_int, vp, vp]
        L.tsvd_read_edge_file.argtypes = [ctypes.c_char_p, ctypes.c_int, vp, vp, i64, vp]
        L.tsvd_read_subset_file.argtypes = [ctypes.c_char_p, vp, i64, vp]
        L.tsvd_read_lines.argtypes = [ctypes.c_char_p, vp, i64, vp, vp]
        L.tsvd_write_embedding_csv.argtypes = [ctypes.c_char_p, vp, i64, i32]
        L.tsvd_ppr_create.argtypes = [vp, vp, i32, f64, f64, ctypes.c_int, vp]
        L.tsvd_ppr_destroy.argtypes = [vp]
        L.tsvd_ppr_refresh.argtypes = [vp, ctypes.c_int, ctypes.c_int, i32, i32]
        L.tsvd_ppr_push.argtypes = [vp, ctypes.c_int, i32, i32]
        L.tsvd_ppr_read.argtypes = [vp, ctypes.c_int, ctypes.c_int, vp]
        L.tsvd_ppr_last_push_stats.argtypes = [vp, ctypes.c_int, vp, vp, vp, vp]
        L.tsvd_ppr_set_two_sided.argtypes = [vp, ctypes.c_int]
        L.tsvd_ppr_refresh_edges.argtypes = [vp, ctypes.c_int, vp, vp, ctypes.c_int64, ctypes.c_int,
                                             ctypes.c_int32, ctypes.c_int32]
        L.tsvd_ppr_get_push_setting.argtypes = [vp, vp]
        L.tsvd_ppr_set_push_setting.argtypes = [vp, vp]
        L.tsvd_ppr_set_graphs.argtypes = [vp, ctypes.c_int]
        L.tsvd_ppr_graph_stats.argtypes = [vp, vp, vp, vp]
        if hasattr(L, "frPCA"):
            L.frPCA.restype = None
            L.frPCA.argtypes = [ctypes.POINTER(MatCSR), vp, vp, vp, ctypes.c_int, ctypes.c_int]
        if hasattr(L, "tsvd_eig_batched"):
            L.tsvd_eig_batched.argtypes = [vp, i32, i32, i32, vp, vp, vp]
        if hasattr(L, "tsvd_frpca_csr_noalloc"):
            L.tsvd_frpca_csr_noalloc.argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32, i32, ctypes.c_uint64]
        if hasattr(L, "tsvd_frpca_csr"):
            L.tsvd_frpca_csr.argtypes = [ctypes.POINTER(MatCSR), ctypes.POINTER(Mat), ctypes.POINTER(Mat),
                                         ctypes.POINTER(Mat), ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
            L.tsvd_frpca_csr_batch.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                               ctypes.c_uint64]
        if hasattr(L, "tsvd_pipeline_create"):
            L.tsvd_pipeline_create.argtypes = [ctypes.POINTER(Config), vp]
            L.tsvd_pipeline_destroy.argtypes = [vp]
            L.tsvd_pipeline_snapshot.argtypes = [vp, vp, vp, i64, ctypes.c_int, ctypes.POINTER(SnapshotStats)]
            L.tsvd_pipeline_embedding.argtypes = [vp, vp]
            L.tsvd_pipeline_root_sigma.argtypes = [vp, vp]
            L.tsvd_pipeline_graph.restype = vp
            L.tsvd_pipeline_graph.argtypes = [vp]
            L.tsvd_pipeline_ppr.restype = vp
            L.tsvd_pipeline_ppr.argtypes = [vp]
            L.tsvd_pipeline_block_matrix.argtypes = [vp, i32, vp]
            L.tsvd_pipeline_block_norms.argtypes = [vp, vp, vp, vp, vp]
            L.tsvd_pipeline_block_transform.argtypes = [vp, i32, i32, i32, ctypes.c_double, i32, vp, vp, vp]
            L.tsvd_pipeline_block_columns.argtypes = [vp, i32, vp, vp]
            L.tsvd_pipeline_node.argtypes = [vp, i32, i32, vp]
            L.tsvd_pipeline_dense_gram.argtypes = [vp, vp]
            L.tsvd_pipeline_gram_apply.argtypes = [vp, vp, i32, vp]
            L.tsvd_pipeline_right_embedding.argtypes = [vp, i32, i32, vp]
            L.tsvd_pipeline_lp_scores.argtypes = [vp, vp, vp, i64, vp]
            L.tsvd_pipeline_source_state.argtypes = [vp, i32, i32, vp, vp]
            L.tsvd_pipeline_snapshot_tuned.argtypes = [vp, vp, vp, i64, ctypes.c_int, vp, i32, vp, vp, vp]
            L.tsvd_server_create.argtypes = [vp, vp, vp]
            L.tsvd_server_destroy.argtypes = [vp]
            L.tsvd_server_submit.argtypes = [vp, vp, vp, i64, vp]
            L.tsvd_server_wait.argtypes = [vp, i64, vp, vp]
            L.tsvd_server_pipeline.restype = vp
            L.tsvd_server_pipeline.argtypes = [vp]
            L.tsvd_server_edge_budget.restype = i64
            L.tsvd_server_edge_budget.argtypes = [vp]
            L.tsvd_probe_atomic_rate.argtypes = [i64, i32, vp]
            L.tsvd_probe_atomic_lanes.argtypes = [i64, i32, i32, i32, vp]
            L.tsvd_gen_rmat.argtypes = [i32, i32, i64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_uint64, vp, vp]
            L.tsvd_gen_chung_lu.argtypes = [i32, i64, ctypes.c_double, ctypes.c_double, ctypes.c_uint64, vp, vp]
        _lib = L
    return _lib


def check(rc: int):
    if rc != 0:
        msg = lib().tsvd_last_error()
        raise TreeSVDError(f"treesvd error {rc}: {msg.decode() if msg else ''}")


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class DeviceArray:
    """A device buffer owned through the C ABI (HBM-resident inputs)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.dtype, self.shape, self.nbytes = host.dtype, host.shape, max(host.nbytes, 1)
        h = ctypes.c_void_p()
        check(lib().tsvd_device_alloc(self.nbytes, ctypes.byref(h)))
        self.ptr = h
        if host.nbytes:
            check(lib().tsvd_memcpy(self.ptr, ctypes.c_void_p(host.ctypes.data), host.nbytes, 0))

    @classmethod
    def adopt(cls, ptr, dtype, shape):
        """Take ownership of a device allocation made by the library."""
        self = cls.__new__(cls)
        self.dtype, self.shape = np.dtype(dtype), tuple(shape)
        self.nbytes = max(int(np.prod(shape)) * self.dtype.itemsize, 1)
        self.ptr = ptr
        return self

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if out.nbytes:
            check(lib().tsvd_memcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, out.nbytes, 1))
        return out

    def free(self):
        if self.ptr:
            lib().tsvd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceSlice:
    """A view [offset, offset+count) of a device array of int32 (no ownership)."""

    def __init__(self, base, offset, count):
        self.base, self.offset, self.count = base, int(offset), int(count)
        self.ptr = ctypes.c_void_p(base.ptr.value + 4 * self.offset)
        self.shape, self.size, self.dtype = (self.count,), self.count, np.dtype(np.int32)

    def numpy(self):
        out = np.empty(self.count, np.int32)
        if self.count:
            check(lib().tsvd_memcpy(ctypes.c_void_p(out.ctypes.data), self.ptr, 4 * self.count, 1))
        return out

    def free(self):
        pass


def gen_rmat(scale, n, m, a=0.57, b=0.19, c=0.19, seed=0):
    """Seeded directed R-MAT edges drawn in HBM (tsvd_gen_rmat): m distinct
    (src, dst) pairs, ids mod n, no self-loops, in a seeded random order.
    Returns two DeviceArray objects of m int32 each."""
    ps, pd = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().tsvd_gen_rmat(int(scale), int(n), int(m), float(a), float(b), float(c), ctypes.c_uint64(seed),
                              ctypes.byref(ps), ctypes.byref(pd)))
    return DeviceArray.adopt(ps, np.int32, (int(m),)), DeviceArray.adopt(pd, np.int32, (int(m),))


def gen_chung_lu(n, m, out_exp, in_exp, seed=0):
    """Seeded directed Chung-Lu edges drawn in HBM (tsvd_gen_chung_lu): m
    distinct (src, dst) pairs, out-weight (u+1)^(-1/(out_exp-1)), in-weight
    (pi(v)+1)^(-1/(in_exp-1)) for a seeded permutation pi, no self-loops, in
    a seeded random order.  Returns two DeviceArray objects of m int32 each."""
    ps, pd = ctypes.c_void_p(), ctypes.c_void_p()
    check(lib().tsvd_gen_chung_lu(int(n), int(m), float(out_exp), float(in_exp), ctypes.c_uint64(seed),
                                  ctypes.byref(ps), ctypes.byref(pd)))
    return DeviceArray.adopt(ps, np.int32, (int(m),)), DeviceArray.adopt(pd, np.int32, (int(m),))


def _dptr(t):
    """(pointer, on_device) for a numpy array or a DeviceArray."""
    if isinstance(t, np.ndarray):
        return ctypes.c_void_p(t.ctypes.data), 0
    return t.ptr, 1


def device_info():
    buf = ctypes.create_string_buffer(256)
    tot, fr = ctypes.c_int64(), ctypes.c_int64()
    check(lib().tsvd_device_info(buf, 256, ctypes.byref(tot), ctypes.byref(fr)))
    return buf.value.decode(), tot.value, fr.value


class Graph:
    def __init__(self, n: int, directed: bool, slack: float = 0.1, handle=None):
        self.n, self.directed = n, directed
        self.owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().tsvd_graph_create(n, int(directed), slack, ctypes.byref(h)))
            handle = h
        self.h = handle

    def close(self):
        if self.owned and self.h:
            lib().tsvd_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def append(self, src, dst):
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, dtype=np.int32)
            dst = np.ascontiguousarray(dst, dtype=np.int32)
        ps, dev = _dptr(src)
        pd, _ = _dptr(dst)
        check(lib().tsvd_graph_append(self.h, ps, pd, int(src.shape[0]), dev))

    def num_edges(self):
        return lib().tsvd_graph_num_edges(self.h)

    def set_staging_limit(self, edges):
        """Host batches above `edges` edges free their device staging after use."""
        check(lib().tsvd_graph_set_staging_limit(self.h, int(edges)))

    def degrees(self, d=0):
        deg = np.zeros(self.n, np.int32); former = np.zeros(self.n, np.int32)
        check(lib().tsvd_graph_degrees(self.h, d, _p(deg), _p(former)))
        return deg, former

    def csr(self, d=0):
        rp = np.zeros(self.n + 1, np.int64)
        check(lib().tsvd_graph_export_csr(self.h, d, _p(rp), None))
        col = np.zeros(max(int(rp[-1]), 1), np.int32)
        check(lib().tsvd_graph_export_csr(self.h, d, _p(rp), _p(col)))
        return rp, col[:int(rp[-1])]


def read_edge_file(path: str, dup_last: bool = True):
    cnt = ctypes.c_int64()
    check(lib().tsvd_read_edge_file(path.encode(), int(dup_last), None, None, 0, ctypes.byref(cnt)))
    s = np.zeros(max(cnt.value, 1), np.int32); d = np.zeros(max(cnt.value, 1), np.int32)
    check(lib().tsvd_read_edge_file(path.encode(), int(dup_last), _p(s), _p(d), cnt.value, ctypes.byref(cnt)))
    return s[:cnt.value], d[:cnt.value]


def read_subset_file(path: str):
    """Subset vertex ids (Tree_u_NC.cpp:363-371)."""
    cnt = ctypes.c_int64()
    check(lib().tsvd_read_subset_file(path.encode(), None, 0, ctypes.byref(cnt)))
    ids = np.zeros(max(cnt.value, 1), np.int32)
    check(lib().tsvd_read_subset_file(path.encode(), _p(ids), cnt.value, ctypes.byref(cnt)))
    return ids[:cnt.value]


def read_lines(path: str):
    """config.txt lines (Tree_u_NC.cpp:338-351)."""
    cnt, need = ctypes.c_int32(), ctypes.c_int64()
    check(lib().tsvd_read_lines(path.encode(), None, 0, ctypes.byref(cnt), ctypes.byref(need)))
    buf = ctypes.create_string_buffer(max(need.value, 1))
    check(lib().tsvd_read_lines(path.encode(), buf, need.value, ctypes.byref(cnt), ctypes.byref(need)))
    raw = buf.raw[:need.value]
    return [x.decode() for x in raw.split(b"\0")[:cnt.value]]


def write_embedding_csv(path: str, U):
    """The drivers' embedding file (Tree_u_NC.cpp:63, 898-909)."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    if U.ndim != 2:
        raise ValueError("U must be a matrix")
    check(lib().tsvd_write_embedding_csv(path.encode(), _p(U), U.shape[0], U.shape[1]))


class PPR:
    def __init__(self, graph: Graph, sources, alpha, rmax, precision=TSVD_FP32, handle=None):
        self.g = graph
        self.sources = np.ascontiguousarray(sources, dtype=np.int32)
        self.nsrc = self.sources.size
        self.owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().tsvd_ppr_create(graph.h, _p(self.sources), self.nsrc, alpha, rmax, precision,
                                        ctypes.byref(h)))
            handle = h
        self.h = handle

    def close(self):
        if self.owned and self.h:
            lib().tsvd_ppr_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def push_setting(self):
        """(hub_chunk, ride_frac, items_per_wave) of the push
        (tsvd_ppr_get_push_setting)."""
        x = PushSetting()
        check(lib().tsvd_ppr_get_push_setting(self.h, ctypes.byref(x)))
        return x.hub_chunk, x.ride_frac, x.items_per_wave

    @push_setting.setter
    def push_setting(self, hr):
        """(hub_chunk, ride_frac[, items_per_wave]); items_per_wave kept when omitted."""
        ipw = int(hr[2]) if len(hr) > 2 else self.push_setting[2]
        x = PushSetting(int(hr[0]), float(hr[1]), ipw)
        check(lib().tsvd_ppr_set_push_setting(self.h, ctypes.byref(x)))

    def refresh_edges(self, src, dst, d=0, s0=0, s1=None):
        """The n >= 1e7-node path's insertion update in batch order
        (tsvd_ppr_refresh_edges), after the same batch was appended."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        s1 = self.nsrc if s1 is None else s1
        check(lib().tsvd_ppr_refresh_edges(self.h, int(d), _p(src), _p(dst), int(src.size), 0, int(s0), int(s1)))

    def set_two_sided(self, on=True):
        check(lib().tsvd_ppr_set_two_sided(self.h, int(on)))

    def set_graphs(self, on=True):
        """Push rounds replayed from captured hipGraphs, or launched one by
        one (tsvd_ppr_set_graphs)."""
        check(lib().tsvd_ppr_set_graphs(self.h, int(on)))

    def graph_stats(self):
        """(replays, captures, failed captures) of this handle's push graphs."""
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        check(lib().tsvd_ppr_graph_stats(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def refresh(self, mode, d=0, s0=0, s1=None):
        check(lib().tsvd_ppr_refresh(self.h, mode, d, s0, self.nsrc if s1 is None else s1))

    def push(self, d=0, s0=0, s1=None):
        check(lib().tsvd_ppr_push(self.h, d, s0, self.nsrc if s1 is None else s1))

    def read(self, d=0, what=0):
        out = np.zeros((self.nsrc, self.g.n), np.float64)
        check(lib().tsvd_ppr_read(self.h, d, what, _p(out)))
        return out

    def p(self, d=0):
        return self.read(d, 0)

    def r(self, d=0):
        return self.read(d, 1)

    def last_push_stats(self, d=0):
        a, b, c, t = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        check(lib().tsvd_ppr_last_push_stats(self.h, d, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c),
                                             ctypes.byref(t)))
        return {"rounds": a.value, "pops": b.value, "relax": c.value, "ms": t.value}


def _mat_csr(A_dense_or_csr):
    """The reference's 1-based mat_csr (pointerB/pointerE) of a dense or scipy
    matrix; returns the struct and the arrays it points into."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A_dense_or_csr)
    A.sort_indices()
    m, n = A.shape
    vals = np.ascontiguousarray(A.data, dtype=np.float64)
    cols = np.ascontiguousarray(A.indices + 1, dtype=np.int32)
    pb = np.ascontiguousarray(A.indptr[:-1] + 1, dtype=np.int32)
    pe = np.ascontiguousarray(A.indptr[1:] + 1, dtype=np.int32)
    if vals.size == 0:
        vals = np.zeros(1); cols = np.ones(1, np.int32)
    csr = MatCSR(int(A.nnz), m, n, vals.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                 cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), pb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                 pe.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return csr, (vals, cols, pb, pe), (m, n)


def _outputs(m, n, k):
    U = np.zeros((k, m)); S = np.zeros(k); V = np.zeros((k, n))     # column-major buffers
    mats = (Mat(m, k, U.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            Mat(k, 1, S.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
            Mat(n, k, V.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return (U, S, V), mats


def frpca_csr(A_dense_or_csr, k: int, q: int, seed: int = 1):
    """frPCA through the drop-in C ABI (tsvd_frpca_csr).  Accepts a dense
    numpy matrix or a scipy CSR matrix; builds the reference's 1-based
    mat_csr (pointerB/pointerE) and returns U (m x k), S (k, ascending), V."""
    csr, keep, (m, n) = _mat_csr(A_dense_or_csr)
    (U, S, V), (mu, ms, mv) = _outputs(m, n, k)
    check(lib().tsvd_frpca_csr(ctypes.byref(csr), ctypes.byref(mu), ctypes.byref(ms), ctypes.byref(mv), k, q, seed))
    return U.T.copy(), S, V.T.copy()


def eig_batched(mats, top: int = 0):
    """tsvd_eig_batched on a (count, n, n) stack of symmetric FP64 matrices.
    Returns (lam (count, n) ascending, V (count, n, n) with V[b][:, i] the
    eigenvector of lam[b, i], info).  With 0 < top < n only the last `top`
    columns of each V are defined."""
    A = np.ascontiguousarray(mats, dtype=np.float64)
    if A.ndim == 2:
        A = A[None]
    count, n, n2 = A.shape
    if n != n2:
        raise ValueError("matrices must be square")
    Vt = np.zeros_like(A)
    lam = np.zeros((count, n))
    info = ctypes.c_int32(0)
    check(lib().tsvd_eig_batched(_p(A), n, count, int(top), _p(Vt), _p(lam), ctypes.byref(info)))
    return lam, np.swapaxes(Vt, 1, 2).copy(), int(info.value)


def frpca_csr_noalloc(A_dense_or_csr, k: int, q: int, seed: int = 1):
    """tsvd_frpca_csr_noalloc: the matrix is staged in device arrays (1-based
    CSR as mat_csr), factored with no allocation inside the call, and U, S, V
    read back.  Returns U (m x k), S (k, ascending), V (n x k)."""
    import scipy.sparse as sp
    A = sp.csr_matrix(A_dense_or_csr, dtype=np.float64)
    A.sort_indices()
    m, n = A.shape
    vals = DeviceArray(A.data.astype(np.float64) if A.nnz else np.zeros(1))
    cols = DeviceArray((A.indices + 1).astype(np.int32) if A.nnz else np.zeros(1, np.int32))
    pb = DeviceArray((A.indptr[:-1] + 1).astype(np.int32))
    pe = DeviceArray((A.indptr[1:] + 1).astype(np.int32))
    U = DeviceArray(np.zeros(m * k))
    S = DeviceArray(np.zeros(k))
    V = DeviceArray(np.zeros(n * k))
    try:
        check(lib().tsvd_frpca_csr_noalloc(m, n, int(A.nnz), vals.ptr, cols.ptr, pb.ptr, pe.ptr, U.ptr, S.ptr, V.ptr,
                                           k, q, ctypes.c_uint64(seed)))
        return U.numpy().reshape(k, m).T.copy(), S.numpy().copy(), V.numpy().reshape(k, n).T.copy()
    finally:
        for x in (vals, cols, pb, pe, U, S, V):
            x.free()


def frpca_csr_batch(mats, k: int, q: int, seed: int = 1):
    """tsvd_frpca_csr_batch over a list of matrices; item i uses seed + i.
    Returns [(U, S, V), ...]."""
    csrs, keep, outs, om = [], [], [], []
    for A in mats:
        csr, arrs, (m, n) = _mat_csr(A)
        csrs.append(csr); keep.append(arrs)
        bufs, ms_ = _outputs(m, n, k)
        outs.append(bufs); om.append(ms_)
    cnt = len(csrs)
    if cnt == 0:
        check(lib().tsvd_frpca_csr_batch(None, None, None, None, 0, k, q, seed))
    PA = (ctypes.POINTER(MatCSR) * cnt)(*[ctypes.pointer(c) for c in csrs])
    PU = (ctypes.POINTER(Mat) * cnt)(*[ctypes.pointer(o[0]) for o in om])
    PS = (ctypes.POINTER(Mat) * cnt)(*[ctypes.pointer(o[1]) for o in om])
    PV = (ctypes.POINTER(Mat) * cnt)(*[ctypes.pointer(o[2]) for o in om])
    check(lib().tsvd_frpca_csr_batch(PA, PU, PS, PV, cnt, k, q, seed))
    return [(U.T.copy(), S, V.T.copy()) for U, S, V in outs]


class Pipeline:
    """The plain per-snapshot method (tsvd_pipeline_*), one handle."""

    def __init__(self, n, directed, sources, alpha, rmax, nparts, fanout, rank, q=6, precision=TSVD_FP32,
                 transform=TSVD_TRANSFORM_NC, delta=0.65, dyn_start=2, lazy_start=2, slack=0.1, seed=0,
                 handle=None, two_sided=False, right_embedding=False, sparse_state=False, sparse_groups=0,
                 merge_bf16x3=False):
        self.sources = np.ascontiguousarray(sources, dtype=np.int32)
        self.nsrc, self.rank, self.n, self.nparts = self.sources.size, rank, n, nparts
        self.cfg = Config(n, int(directed), self.nsrc, self.sources.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          alpha, rmax, precision, nparts, fanout, rank, q, transform, delta, dyn_start, lazy_start,
                          slack, seed, int(two_sided), int(right_embedding), int(sparse_state), int(sparse_groups),
                          int(merge_bf16x3))
        self.owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().tsvd_pipeline_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        else:
            h = handle
        self.h = h
        self.graph = Graph(n, directed, handle=ctypes.c_void_p(lib().tsvd_pipeline_graph(h)))
        self.ppr = PPR(self.graph, self.sources, alpha, rmax, precision,
                       handle=ctypes.c_void_p(lib().tsvd_pipeline_ppr(h)))

    def close(self):
        if self.h and self.owned:
            lib().tsvd_pipeline_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def snapshot(self, src, dst):
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, dtype=np.int32)
            dst = np.ascontiguousarray(dst, dtype=np.int32)
        ps, dev = _dptr(src)
        pd, _ = _dptr(dst)
        st = SnapshotStats()
        check(lib().tsvd_pipeline_snapshot(self.h, ps, pd, int(src.shape[0]), dev, ctypes.byref(st)))
        return st.as_dict()

    def snapshot_tuned(self, src, dst, candidates):
        """One snapshot whose push is first timed under each (hub_chunk,
        ride_frac[, items_per_wave]) candidate (tsvd_pipeline_snapshot_tuned;
        items_per_wave 8 when omitted).  Returns (stats, chosen setting as
        (hub_chunk, ride_frac, items_per_wave), per-candidate push ms; -1
        where not timed)."""
        if isinstance(src, np.ndarray):
            src = np.ascontiguousarray(src, dtype=np.int32)
            dst = np.ascontiguousarray(dst, dtype=np.int32)
        ps, dev = _dptr(src)
        pd, _ = _dptr(dst)
        cand = (PushSetting * max(len(candidates), 1))(*[PushSetting(int(c[0]), float(c[1]), int(c[2]) if len(c) > 2 else 8)
                                                          for c in candidates])
        ms = np.zeros(max(len(candidates), 1))
        chosen = PushSetting()
        st = SnapshotStats()
        check(lib().tsvd_pipeline_snapshot_tuned(self.h, ps, pd, int(src.shape[0]), dev, cand, len(candidates),
                                                 ctypes.byref(chosen), _p(ms), ctypes.byref(st)))
        return st.as_dict(), (chosen.hub_chunk, chosen.ride_frac, chosen.items_per_wave), ms[:len(candidates)]

    def embedding(self):
        E = np.zeros((self.nsrc, self.rank))
        check(lib().tsvd_pipeline_embedding(self.h, _p(E)))
        return E

    def root_sigma(self):
        s = np.zeros(self.rank)
        check(lib().tsvd_pipeline_root_sigma(self.h, _p(s)))
        return s

    def block_matrix(self, j, width):
        A = np.zeros((width, self.nsrc))         # column-major nsrc x width
        check(lib().tsvd_pipeline_block_matrix(self.h, j, _p(A)))
        return A.T.copy()

    def block_transform(self, j0, j1, iter_, delta, test=True):
        """tsvd_pipeline_block_transform over blocks [j0, j1): (a_norm,
        e_norm, record) with record -1 (clean) or iter_ (dirty, matrix
        replaced)."""
        cnt = max(j1 - j0, 1)
        a, e, rec = np.zeros(cnt), np.zeros(cnt), np.zeros(cnt, np.int32)
        check(lib().tsvd_pipeline_block_transform(self.h, int(j0), int(j1), int(iter_), float(delta), int(test),
                                                  _p(a), _p(e), _p(rec)))
        return a[:j1 - j0], e[:j1 - j0], rec[:j1 - j0]

    def block_columns(self, j):
        """Vertex ids of block j's stored columns, in sketch-row order."""
        cnt = ctypes.c_int32()
        check(lib().tsvd_pipeline_block_columns(self.h, int(j), None, ctypes.byref(cnt)))
        cols = np.zeros(max(cnt.value, 1), np.int32)
        check(lib().tsvd_pipeline_block_columns(self.h, int(j), _p(cols), ctypes.byref(cnt)))
        return cols[:cnt.value]

    def block_norms(self):
        a, e, r = np.zeros(self.nparts), np.zeros(self.nparts), np.zeros(self.nparts)
        d = np.zeros(self.nparts, np.int32)
        check(lib().tsvd_pipeline_block_norms(self.h, _p(a), _p(e), _p(r), _p(d)))
        return a, e, r, d.astype(bool)

    def node(self, level, index):
        out = np.zeros((self.rank, self.nsrc))
        check(lib().tsvd_pipeline_node(self.h, level, index, _p(out)))
        return out.T.copy()

    def dense_gram(self):
        """FP64 A A^T of the current transformed matrix (nsrc x nsrc)."""
        G = np.zeros((self.nsrc, self.nsrc))
        check(lib().tsvd_pipeline_dense_gram(self.h, _p(G)))
        return G

    def right_embedding(self, v0=0, v1=None):
        """V rows [v0, v1) of the LP right embedding (n x rank)."""
        v1 = self.n if v1 is None else v1
        V = np.zeros((v1 - v0, self.rank))
        check(lib().tsvd_pipeline_right_embedding(self.h, int(v0), int(v1), _p(V)))
        return V

    def lp_scores(self, src, dst):
        """Link scores U_j[row(src)] . V[dst] (Tree_u_LP.cpp:1047-1090)."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        out = np.zeros(src.size)
        check(lib().tsvd_pipeline_lp_scores(self.h, _p(src), _p(dst), int(src.size), _p(out)))
        return out

    def source_state(self, i, d=0):
        """(reserves, residues) of subset source i in push direction d, n
        values each (with sparse_state: the stored positive pair lists)."""
        p, r = np.zeros(self.n), np.zeros(self.n)
        check(lib().tsvd_pipeline_source_state(self.h, int(d), int(i), _p(p), _p(r)))
        return p, r

    def gram_apply(self, Y):
        """Z = A A^T Y in FP64 for an nsrc x l block Y, without forming A A^T."""
        Y = np.ascontiguousarray(Y, dtype=np.float64)
        if Y.ndim != 2 or Y.shape[0] != self.nsrc:
            raise ValueError(f"Y must be {self.nsrc} x l")
        Z = np.zeros_like(Y)
        check(lib().tsvd_pipeline_gram_apply(self.h, _p(Y), Y.shape[1], _p(Z)))
        return Z


def profile_enable(on: bool = True):
    check(lib().tsvd_profile_enable(int(on)))


def profile_reset():
    check(lib().tsvd_profile_reset())


def gemm_tuning(on: bool = True):
    """Tile-width autotuning of the MFMA products (tsvd_gemm_tuning)."""
    check(lib().tsvd_gemm_tuning(int(on)))


def gemm_tuning_read():
    """{kind: {"tile_n", "ms_tile80", "ms_tile144", "calls"}} chosen so far."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().tsvd_gemm_tuning_read(buf, 1 << 16))
    return json.loads(buf.value.decode())


def probe_atomic_rate(rows=1 << 22, iters=64):
    """Wave-level (64-lane, 256-byte) returning FP32 atomics per second on
    this device (tsvd_probe_atomic_rate)."""
    r = ctypes.c_double()
    check(lib().tsvd_probe_atomic_rate(int(rows), int(iters), ctypes.byref(r)))
    return r.value


def probe_atomic_lanes(returning=True, lanes=64, rows=1 << 22, iters=64):
    """tsvd_probe_atomic_lanes: wave-level FP32 atomics per second with the
    returning or non-returning form and only `lanes` active lanes."""
    r = ctypes.c_double()
    check(lib().tsvd_probe_atomic_lanes(int(rows), int(iters), int(returning), int(lanes), ctypes.byref(r)))
    return r.value


def profile_read():
    """{kernel: (launches, total_ms, algorithmic_bytes, flops)} since the last reset."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    check(lib().tsvd_profile_read(buf, 1 << 16))
    return {k: (int(v[0]), float(v[1]), float(v[2]), float(v[3])) for k, v in json.loads(buf.value.decode()).items()}


class Server:
    """The served form (tsvd_server_*): continuous batching of insertion
    batches over one resident pipeline.  max_step_requests=1 is the plain
    per-snapshot method under the same arrival stream."""

    def __init__(self, n, directed, sources, alpha, rmax, nparts, fanout, rank, q=6, precision=TSVD_FP32,
                 transform=TSVD_TRANSFORM_NC, delta=0.65, dyn_start=2, lazy_start=2, slack=0.1, seed=0,
                 max_step_edges=0, max_step_requests=0, two_sided=False, right_embedding=False,
                 sparse_state=False, sparse_groups=0, overlap=True, merge_bf16x3=False):
        self.sources = np.ascontiguousarray(sources, dtype=np.int32)
        self.nsrc, self.rank = self.sources.size, rank
        self.cfg = Config(n, int(directed), self.nsrc, self.sources.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                          alpha, rmax, precision, nparts, fanout, rank, q, transform, delta, dyn_start, lazy_start,
                          slack, seed, int(two_sided), int(right_embedding), int(sparse_state), int(sparse_groups),
                          int(merge_bf16x3))
        self.scfg = ServerConfig(int(max_step_edges), int(max_step_requests), int(not overlap))
        h = ctypes.c_void_p()
        check(lib().tsvd_server_create(ctypes.byref(self.cfg), ctypes.byref(self.scfg), ctypes.byref(h)))
        self.h = h
        self.pipeline = Pipeline(n, directed, self.sources, alpha, rmax, nparts, fanout, rank, q, precision,
                                 transform, delta, dyn_start, lazy_start, slack, seed,
                                 handle=ctypes.c_void_p(lib().tsvd_server_pipeline(h)), two_sided=two_sided,
                                 right_embedding=right_embedding, sparse_state=sparse_state,
                                 sparse_groups=sparse_groups, merge_bf16x3=merge_bf16x3)

    @property
    def edge_budget(self):
        return int(lib().tsvd_server_edge_budget(self.h))

    def submit(self, src, dst):
        src = np.ascontiguousarray(src, dtype=np.int32)
        dst = np.ascontiguousarray(dst, dtype=np.int32)
        if src.shape != dst.shape:
            raise ValueError("src and dst differ in length")
        t = ctypes.c_int64()
        check(lib().tsvd_server_submit(self.h, _p(src), _p(dst), int(src.size), ctypes.byref(t)))
        return t.value

    def wait(self, ticket, embedding=True):
        E = np.zeros((self.nsrc, self.rank)) if embedding else None
        r = RequestResult()
        check(lib().tsvd_server_wait(self.h, int(ticket), _p(E), ctypes.byref(r)))
        return E, r.as_dict()

    def close(self):
        if self.h:
            lib().tsvd_server_destroy(self.h)
        self.h = None
        if getattr(self, "pipeline", None) is not None:
            self.pipeline.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
