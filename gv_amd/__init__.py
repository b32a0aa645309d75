"""MI355X-native GraphVine: dynamic graph store + analytics on one GPU.

Python mirror of the reference's host interface over the C ABI in
include/gv.h (libgv.so, hand-written HIP kernels for gfx950).  Names follow
the reference:

    GraphVine(...)            include/dynamicStructure/graphvine.cuh:29-105
      .bulkBuild / .batchInsert / .batchDelete / .compaction / .getVertexSize
    static_pagerank(...)      include/algorithm/pagerank/pagerank.cuh:11
    static_tc(...)            include/algorithm/tc/tc.cuh:10
    bc(...)                   include/algorithm/bc/bc.cuh:8

Differences from the reference that are part of the contract: vertex IDs are
0-based; update batches are COO arrays (host numpy or device tensors) rather
than a CSR object; errors raise GraphVineError instead of exiting.  There is
no CPU fallback: if libgv.so is missing or the GPU is absent, calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libgv.so")

GV_OK = 0
GV_E_INVALID = -1
GV_E_RANGE = -2
GV_E_NOMEM = -3
GV_E_POOL = -4
GV_E_HIP = -5
GV_E_UNSUPPORTED = -6
GV_MEM_HOST = 0
GV_MEM_DEVICE = 1
GV_OP_INSERT = 0
GV_OP_DELETE = 1

EXPORTED = [
    "gv_config_default", "gv_create", "gv_destroy", "gv_bulk_build", "gv_batch_insert",
    "gv_batch_delete", "gv_batch_apply", "gv_batch_apply_device", "gv_compact", "gv_num_vertices", "gv_num_edges", "gv_degrees",
    "gv_pool_stats", "gv_export_csr", "gv_neighbors", "gv_neighbors_noalloc", "gv_pagerank",
    "gv_pagerank_ex",
    "gv_pagerank_opts_default", "gv_pagerank_run", "gv_pagerank_personalized",
    "gv_triangle_count",
    "gv_bc", "gv_gen_rmat", "gv_gen_uniform", "gv_gen_sample", "gv_gen_chung_lu",
    "gv_get_config", "gv_set_tuning", "gv_reserve", "gv_pagerank_reserve", "gv_pagerank_last",
    "gv_capture_begin", "gv_capture_end",
    "gv_capture_launch", "gv_capture_destroy", "gv_read_matrix_market", "gv_sync",
    "gv_get_stream", "gv_test_set_sort_spin_limit",
    "gv_last_error",
]


class GvConfig(ctypes.Structure):
    _fields_ = [
        ("block_slots", ctypes.c_uint32),
        ("pool_blocks", ctypes.c_uint64),
        ("pool_fraction", ctypes.c_double),
        ("compact_threshold", ctypes.c_float),
        ("pr_thread_max", ctypes.c_uint32),
        ("pr_group_max", ctypes.c_uint32),
        ("pr_wave_max", ctypes.c_uint32),
        ("maintain_in_edges", ctypes.c_int),
        ("device", ctypes.c_int),
        ("pr_layout", ctypes.c_int),
        ("bc_bu_ratio", ctypes.c_float),
    ]


class GvBatchStats(ctypes.Structure):
    _fields_ = [
        ("requested", ctypes.c_uint64),
        ("directed_entries", ctypes.c_uint64),
        ("applied", ctypes.c_uint64),
        ("touched", ctypes.c_uint64),
        ("blocks_allocated", ctypes.c_uint64),
        ("live_edges", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


class GvPagerankStats(ctypes.Structure):
    _fields_ = [
        ("iters", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("kernel_ms", ctypes.c_double),
        ("edges", ctypes.c_uint64),
    ]


class GvPagerankOpts(ctypes.Structure):
    _fields_ = [
        ("alpha", ctypes.c_float),
        ("eps", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("init", ctypes.c_void_p),
        ("init_mem", ctypes.c_int),
        ("compensated", ctypes.c_int),
    ]


class GraphVineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"gv status {status}: {msg}")
        self.status = status


_lib = None


def lib():
    """Load libgv.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run `python -m gv_amd.build`")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, u64, i32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
    sig = {
        "gv_config_default": [ctypes.POINTER(GvConfig)],
        "gv_create": [ctypes.POINTER(GvConfig), u32, i32, ctypes.POINTER(P)],
        "gv_destroy": [P],
        "gv_bulk_build": [P, P, P, u64, i32, ctypes.POINTER(GvBatchStats)],
        "gv_batch_insert": [P, P, P, u64, i32, ctypes.POINTER(GvBatchStats)],
        "gv_batch_delete": [P, P, P, u64, i32, ctypes.POINTER(GvBatchStats)],
        "gv_compact": [P, ctypes.c_float, ctypes.POINTER(u64)],
        "gv_batch_apply": [P, P, P, P, P, u32, i32, ctypes.c_float,
                           ctypes.POINTER(GvBatchStats)],
        "gv_batch_apply_device": [P, P, P, P, P, u32, u64, ctypes.c_float],
        "gv_num_vertices": [P, ctypes.POINTER(u32)],
        "gv_num_edges": [P, ctypes.POINTER(u64)],
        "gv_degrees": [P, P, i32],
        "gv_pool_stats": [P, ctypes.POINTER(u64), ctypes.POINTER(u64)],
        "gv_export_csr": [P, P, P, i32],
        "gv_neighbors": [P, P, u32, P, P, i32],
        "gv_neighbors_noalloc": [P, P, u32, P, P],
        "gv_pagerank": [P, ctypes.c_float, ctypes.c_double, i32, P, i32, ctypes.POINTER(i32),
                        ctypes.POINTER(ctypes.c_double)],
        "gv_pagerank_ex": [P, ctypes.c_float, ctypes.c_double, i32, P, i32,
                           ctypes.POINTER(GvPagerankStats)],
        "gv_pagerank_opts_default": [ctypes.POINTER(GvPagerankOpts)],
        "gv_pagerank_run": [P, ctypes.POINTER(GvPagerankOpts), P, i32,
                            ctypes.POINTER(GvPagerankStats)],
        "gv_pagerank_personalized": [P, ctypes.c_float, ctypes.c_double, i32, P, u32, P, i32,
                                     ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double)],
        "gv_triangle_count": [P, P, i32, ctypes.POINTER(u64)],
        "gv_triangle_count_update": [P, i32, P, P, u64, i32, P, i32, ctypes.POINTER(u64),
                                     ctypes.POINTER(GvBatchStats)],
        "gv_bc": [P, P, u32, P, i32],
        "gv_gen_rmat": [P, u64, i32, u64, u32, u32, u32, P, P],
        "gv_gen_uniform": [P, u64, u32, u64, P, P],
        "gv_gen_sample": [P, u64, u32, P, P, u64, P, P],
        "gv_gen_chung_lu": [P, u64, u32, ctypes.c_double, ctypes.c_double, ctypes.c_double, u64,
                            P, P],
        "gv_get_config": [P, ctypes.POINTER(GvConfig)],
        "gv_set_tuning": [P, ctypes.POINTER(GvConfig)],
        "gv_reserve": [P, u64, u32],
        "gv_pagerank_reserve": [P, u64],
        "gv_pagerank_last": [P, ctypes.POINTER(GvPagerankStats)],
        "gv_capture_begin": [P],
        "gv_capture_end": [P, ctypes.POINTER(P)],
        "gv_capture_launch": [P, ctypes.POINTER(GvBatchStats)],
        "gv_capture_destroy": [P],
        "gv_read_matrix_market": [ctypes.c_char_p, ctypes.POINTER(u32), ctypes.POINTER(u64), P,
                                  P, ctypes.POINTER(i32)],
        "gv_sync": [P],
        "gv_get_stream": [P, ctypes.POINTER(P)],
        "gv_test_set_sort_spin_limit": [u32],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.gv_last_error.argtypes = []
    L.gv_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != GV_OK:
        raise GraphVineError(rc, lib().gv_last_error().decode(errors="replace"))


def default_config(**overrides):
    cfg = GvConfig()
    _check(lib().gv_config_default(ctypes.byref(cfg)))
    for k, v in overrides.items():
        if not hasattr(cfg, k):
            raise KeyError(k)
        setattr(cfg, k, v)
    return cfg


def _is_device(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


def _ptr_u32(x):
    """(pointer, keepalive, mem) for a uint32 host array, a gv_amd.hip
    DeviceArray or a device tensor."""
    if hasattr(x, "ptr") and hasattr(x, "nbytes") and not isinstance(x, np.ndarray):
        if np.dtype(x.dtype) != np.uint32:
            raise TypeError("device arrays must be uint32")
        return x.ptr, x, GV_MEM_DEVICE
    if _is_device(x):
        import torch
        if x.dtype not in (torch.int32, torch.uint32):
            raise TypeError("device arrays must be int32/uint32")
        if not x.is_contiguous():
            raise ValueError("device arrays must be contiguous")
        return ctypes.c_void_p(x.data_ptr()), x, GV_MEM_DEVICE
    a = np.ascontiguousarray(x, dtype=np.uint32)
    return a.ctypes.data_as(ctypes.c_void_p), a, GV_MEM_HOST


def _dev_len(x):
    return int(x.n) if hasattr(x, "ptr") and hasattr(x, "nbytes") else int(x.numel())


def _dev_ptr(x, dtypes, name):
    """(pointer, length) of a DeviceArray or contiguous device tensor whose
    element type is one of dtypes (numpy types)."""
    if hasattr(x, "ptr") and hasattr(x, "nbytes") and not isinstance(x, np.ndarray):
        if np.dtype(x.dtype) not in [np.dtype(t) for t in dtypes]:
            raise TypeError(f"{name}: device array of {dtypes} expected")
        return x.ptr, int(x.n)
    if _is_device(x):
        import torch
        tmap = {np.dtype(np.uint64): torch.uint64, np.dtype(np.int64): torch.int64,
                np.dtype(np.int32): torch.int32, np.dtype(np.uint32): torch.uint32}
        if x.dtype not in [tmap[np.dtype(t)] for t in dtypes] or not x.is_contiguous():
            raise TypeError(f"{name}: contiguous device tensor of {dtypes} expected")
        return ctypes.c_void_p(x.data_ptr()), int(x.numel())
    raise TypeError(f"{name}: device array expected")


class GraphVine:
    """Device-resident dynamic graph (reference: class GraphVine)."""

    def __init__(self, num_vertices, directed=False, config=None, **overrides):
        cfg = config if config is not None else default_config(**overrides)
        h = ctypes.c_void_p()
        _check(lib().gv_create(ctypes.byref(cfg), int(num_vertices), 1 if directed else 0,
                               ctypes.byref(h)))
        self._h = h
        self.num_vertices = int(num_vertices)
        self.directed = bool(directed)
        self.config = cfg

    @property
    def handle(self):
        if self._h is None:
            raise GraphVineError(GV_E_INVALID, "graph is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().gv_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _update(self, fn, src, dst, with_stats):
        ps, ks, ms = _ptr_u32(src)
        pd, kd, md = _ptr_u32(dst)
        if ms != md:
            raise ValueError("src and dst must both be host or both be device arrays")
        n = int(len(src))
        if int(len(dst)) != n:
            raise ValueError("src and dst lengths differ")
        st = GvBatchStats()
        _check(fn(self.handle, ps, pd, n, ms, ctypes.byref(st) if with_stats else None))
        return st.as_dict() if with_stats else None

    # reference: GraphVine::bulkBuild (graphvine.cu:181-220)
    def bulkBuild(self, src, dst, stats=True):
        return self._update(lib().gv_bulk_build, src, dst, stats)

    # reference: GraphVine::batchInsert (graphvine.cu:222-371)
    def batchInsert(self, src, dst, stats=True):
        return self._update(lib().gv_batch_insert, src, dst, stats)

    # reference: GraphVine::batchDelete (graphvine.cu:373-451)
    def batchDelete(self, src, dst, stats=True):
        return self._update(lib().gv_batch_delete, src, dst, stats)

    def applyBatches(self, src, dst, offsets, kinds, compact_fraction=-1.0, stats=True):
        """Batched form: batches [offsets[b], offsets[b+1]) of src/dst with
        kinds[b] in {GV_OP_INSERT, GV_OP_DELETE}, applied as one pipeline with
        the same final graph as applying them in order."""
        ps, ks, ms = _ptr_u32(src)
        pd, kd, md = _ptr_u32(dst)
        if ms != md:
            raise ValueError("src and dst must both be host or both be device arrays")
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        kin = np.ascontiguousarray(kinds, dtype=np.int32)
        nb = len(kin)
        if len(off) != nb + 1 or int(off[-1]) > min(len(src), len(dst)):
            raise ValueError("offsets must have n_batches+1 entries within the arrays")
        st = GvBatchStats()
        _check(lib().gv_batch_apply(self.handle, ps, pd, off.ctypes.data_as(ctypes.c_void_p),
                                    kin.ctypes.data_as(ctypes.c_void_p), nb, ms,
                                    float(compact_fraction), ctypes.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def applyBatchesDevice(self, src, dst, offsets, kinds, compact_fraction=-1.0):
        """Batched form over a device-resident queue (no host reads, capturable):
        src/dst are uint32 device arrays of the queue's capacity (max_pairs),
        offsets a uint64 device array of max_batches+1 entries and kinds an
        int32 device array of max_batches entries (gv_amd.hip.DeviceArray or
        contiguous device tensors), all read when the call runs on the
        handle's stream.  Inside a capture it needs reserve(max_pairs,
        max_batches) first; errors found on the device surface at the next
        call that reads the status (sync())."""
        ps, ks, ms = _ptr_u32(src)
        pd, kd, md = _ptr_u32(dst)
        if ms != GV_MEM_DEVICE or md != GV_MEM_DEVICE:
            raise ValueError("src and dst must be device arrays")
        po, no = _dev_ptr(offsets, (np.uint64, np.int64), "offsets")
        pk, nk = _dev_ptr(kinds, (np.int32,), "kinds")
        if no != nk + 1:
            raise ValueError("offsets must have max_batches+1 entries")
        n = _dev_len(src)
        if _dev_len(dst) != n:
            raise ValueError("src and dst lengths differ")
        _check(lib().gv_batch_apply_device(self.handle, ps, pd, po, pk, nk, n,
                                           float(compact_fraction)))

    # reference: GraphVine::compaction (graphvine.cu:453-464)
    def compaction(self, tombstone_fraction=0.0):
        n = ctypes.c_uint64()
        _check(lib().gv_compact(self.handle, float(tombstone_fraction), ctypes.byref(n)))
        return int(n.value)

    # reference: GraphVine::getVertexSize (graphvine.cuh:56)
    def getVertexSize(self):
        return self.num_vertices

    def sync(self):
        _check(lib().gv_sync(self.handle))

    def get_config(self):
        cfg = GvConfig()
        _check(lib().gv_get_config(self.handle, ctypes.byref(cfg)))
        return cfg

    def set_tuning(self, **fields):
        """Change compact_threshold / pr_thread_max / pr_group_max /
        pr_wave_max / pr_layout / bc_bu_ratio of the live handle."""
        cfg = self.get_config()
        for k, v in fields.items():
            if k not in ("compact_threshold", "pr_thread_max", "pr_group_max", "pr_wave_max",
                         "pr_layout", "bc_bu_ratio"):
                raise ValueError(f"{k} is not a live tuning field")
            setattr(cfg, k, v)
        _check(lib().gv_set_tuning(self.handle, ctypes.byref(cfg)))
        self.config = self.get_config()

    def reserve(self, max_pairs, max_batches=1):
        """Size staging buffers so captured updates never allocate."""
        _check(lib().gv_reserve(self.handle, int(max_pairs), int(max_batches)))

    def reserve_pagerank(self, max_edges):
        """Size the default PageRank layout for up to max_edges live edges, so
        PageRank can be captured (and never allocates up to that size)."""
        _check(lib().gv_pagerank_reserve(self.handle, int(max_edges)))

    def pagerank_last(self):
        """Stats dict (iters, residual, edges) of the last default-layout
        PageRank call, captured replays included."""
        st = GvPagerankStats()
        _check(lib().gv_pagerank_last(self.handle, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in st._fields_}

    def capture(self):
        """Context manager recording the update/compaction/PageRank calls
        made inside it; on exit .result holds a Capture that replays them."""
        return _CaptureScope(self)

    def stream(self):
        p = ctypes.c_void_p()
        _check(lib().gv_get_stream(self.handle, ctypes.byref(p)))
        return p.value

    def num_edges(self):
        n = ctypes.c_uint64()
        _check(lib().gv_num_edges(self.handle, ctypes.byref(n)))
        return int(n.value)

    def degrees(self):
        out = np.empty(self.num_vertices, dtype=np.uint32)
        _check(lib().gv_degrees(self.handle, out.ctypes.data_as(ctypes.c_void_p), GV_MEM_HOST))
        return out

    def pool_stats(self):
        f, t = ctypes.c_uint64(), ctypes.c_uint64()
        _check(lib().gv_pool_stats(self.handle, ctypes.byref(f), ctypes.byref(t)))
        return int(f.value), int(t.value)

    def export_csr(self):
        """Sorted CSR of the live graph: (offsets[V+1] uint64, cols uint32)."""
        off = np.empty(self.num_vertices + 1, dtype=np.uint64)
        _check(lib().gv_export_csr(self.handle, off.ctypes.data_as(ctypes.c_void_p), None,
                                   GV_MEM_HOST))
        cols = np.empty(int(off[-1]), dtype=np.uint32)
        _check(lib().gv_export_csr(self.handle, off.ctypes.data_as(ctypes.c_void_p),
                                   cols.ctypes.data_as(ctypes.c_void_p), GV_MEM_HOST))
        return off, cols

    def export_csr_device(self, offsets, cols):
        """Sorted CSR into device tensors (uint64/int64 offsets[V+1], int32 cols)."""
        _check(lib().gv_export_csr(self.handle, ctypes.c_void_p(offsets.data_ptr()),
                                   ctypes.c_void_p(cols.data_ptr()) if cols is not None else None,
                                   GV_MEM_DEVICE))

    def neighbors(self, vs):
        """1-hop lookups: (offs[n+1], sorted neighbours concatenated)."""
        v = np.ascontiguousarray(vs, dtype=np.uint32)
        n = len(v)
        offs = np.empty(n + 1, dtype=np.uint64)
        _check(lib().gv_neighbors(self.handle, v.ctypes.data_as(ctypes.c_void_p), n,
                                  offs.ctypes.data_as(ctypes.c_void_p), None, GV_MEM_HOST))
        out = np.empty(int(offs[-1]), dtype=np.uint32)
        _check(lib().gv_neighbors(self.handle, v.ctypes.data_as(ctypes.c_void_p), n,
                                  offs.ctypes.data_as(ctypes.c_void_p),
                                  out.ctypes.data_as(ctypes.c_void_p), GV_MEM_HOST))
        return offs, out


# reference: static_pagerank (pagerank.cu:3-68)
def static_pagerank(graph, alpha=0.85, epsilon=1e-6, max_iter=100):
    """Returns (x float32[V], iterations, final L1 residual)."""
    x = np.empty(graph.num_vertices, dtype=np.float32)
    it = ctypes.c_int()
    res = ctypes.c_double()
    _check(lib().gv_pagerank(graph.handle, float(alpha), float(epsilon), int(max_iter),
                             x.ctypes.data_as(ctypes.c_void_p), GV_MEM_HOST, ctypes.byref(it),
                             ctypes.byref(res)))
    return x, int(it.value), float(res.value)


def pagerank_device(graph, out_dev, alpha=0.85, epsilon=1e-6, max_iter=100):
    """PageRank into a device float32[V] buffer; returns the stats dict
    (iters, residual, kernel_ms = summed iteration-kernel device time, edges)."""
    st = GvPagerankStats()
    _check(lib().gv_pagerank_ex(graph.handle, float(alpha), float(epsilon), int(max_iter),
                                out_dev.ptr, GV_MEM_DEVICE, ctypes.byref(st)))
    return {f: getattr(st, f) for f, _ in st._fields_}


def pagerank_device_run(graph, out_dev, init_dev=None, alpha=0.85, epsilon=1e-6, max_iter=100,
                        compensated=True, stats=True):
    """PageRank into a device float32[V] buffer, warm-started from a device
    float32[V] vector (init_dev; None = 1/V); returns the stats dict, or
    None with stats=False (the form a capture records)."""
    o = GvPagerankOpts()
    _check(lib().gv_pagerank_opts_default(ctypes.byref(o)))
    o.alpha, o.eps, o.max_iter = float(alpha), float(epsilon), int(max_iter)
    o.compensated = 1 if compensated else 0
    if init_dev is not None:
        o.init = init_dev.ptr
        o.init_mem = GV_MEM_DEVICE
    st = GvPagerankStats()
    _check(lib().gv_pagerank_run(graph.handle, ctypes.byref(o), out_dev.ptr, GV_MEM_DEVICE,
                                 ctypes.byref(st) if stats else None))
    return {f: getattr(st, f) for f, _ in st._fields_} if stats else None


def pagerank(graph, alpha=0.85, epsilon=1e-6, max_iter=100, init=None, compensated=True,
             stats=False):
    """PageRank with a warm start (init: float32[V] host array) and the choice
    of compensated or plain FP32 sums.  Returns (x, iters, residual[, stats])."""
    o = GvPagerankOpts()
    _check(lib().gv_pagerank_opts_default(ctypes.byref(o)))
    o.alpha, o.eps, o.max_iter = float(alpha), float(epsilon), int(max_iter)
    o.compensated = 1 if compensated else 0
    keep = None
    if init is not None:
        keep = np.ascontiguousarray(init, dtype=np.float32)
        if len(keep) != graph.num_vertices:
            raise ValueError("init must have num_vertices entries")
        o.init = keep.ctypes.data_as(ctypes.c_void_p)
        o.init_mem = GV_MEM_HOST
    x = np.empty(graph.num_vertices, dtype=np.float32)
    st = GvPagerankStats()
    _check(lib().gv_pagerank_run(graph.handle, ctypes.byref(o), x.ctypes.data_as(ctypes.c_void_p),
                                 GV_MEM_HOST, ctypes.byref(st)))
    d = {f: getattr(st, f) for f, _ in st._fields_}
    return (x, d["iters"], d["residual"], d) if stats else (x, d["iters"], d["residual"])


def pagerank_personalized(graph, seeds, alpha=0.85, epsilon=1e-6, max_iter=100):
    """Up to 64 personalized PageRank vectors at once: returns (X[V, n], iters,
    largest per-column residual)."""
    sd = np.ascontiguousarray(seeds, dtype=np.uint32)
    X = np.empty((graph.num_vertices, len(sd)), dtype=np.float32)
    it = ctypes.c_int()
    r = ctypes.c_double()
    _check(lib().gv_pagerank_personalized(graph.handle, float(alpha), float(epsilon),
                                          int(max_iter), sd.ctypes.data_as(ctypes.c_void_p),
                                          len(sd), X.ctypes.data_as(ctypes.c_void_p),
                                          GV_MEM_HOST, ctypes.byref(it), ctypes.byref(r)))
    return X, int(it.value), float(r.value)


# reference: static_tc (tc.cu:3-27)
def static_tc(graph, per_vertex=True):
    """Returns (per-vertex uint64 counts or None, total)."""
    pv = np.empty(graph.num_vertices, dtype=np.uint64) if per_vertex else None
    tot = ctypes.c_uint64()
    _check(lib().gv_triangle_count(graph.handle,
                                   pv.ctypes.data_as(ctypes.c_void_p) if per_vertex else None,
                                   GV_MEM_HOST, ctypes.byref(tot)))
    return pv, int(tot.value)


# reference: dynamic_tc (tc.cu:29-64); the batch is explicit (the reference
# reads the affected set its last batchInsert left behind)
def dynamic_tc(graph, src, dst, per_vertex, total, delete=False, stats=False):
    """Apply one insert (or, delete=True, delete) batch to an undirected graph
    and update counts taken before it: per_vertex (uint64 host array of V,
    updated in place; may be None) and total.  Returns (total, stats)."""
    ps, ks, ms = _ptr_u32(src)
    pd, kd, md = _ptr_u32(dst)
    if ms != md:
        raise ValueError("src and dst must both be host or both be device arrays")
    n = int(len(src))
    if int(len(dst)) != n:
        raise ValueError("src and dst lengths differ")
    ppv = None
    if per_vertex is not None:
        if (not isinstance(per_vertex, np.ndarray) or per_vertex.dtype != np.uint64
                or per_vertex.shape != (graph.num_vertices,)
                or not per_vertex.flags.c_contiguous):
            raise TypeError("per_vertex must be a contiguous uint64 array of V entries")
        ppv = per_vertex.ctypes.data_as(ctypes.c_void_p)
    tot = ctypes.c_uint64(int(total))
    st = GvBatchStats()
    _check(lib().gv_triangle_count_update(graph.handle, GV_OP_DELETE if delete else GV_OP_INSERT,
                                          ps, pd, n, ms, ppv, GV_MEM_HOST, ctypes.byref(tot),
                                          ctypes.byref(st) if stats else None))
    return int(tot.value), (st.as_dict() if stats else None)


# reference: bc (bc.cu:3-88); sources are explicit (the reference ignores its
# argument and runs every vertex, bc.cu:42)
def bc(graph, sources):
    s = np.ascontiguousarray(sources, dtype=np.uint32)
    out = np.empty(graph.num_vertices, dtype=np.float32)
    _check(lib().gv_bc(graph.handle, s.ctypes.data_as(ctypes.c_void_p), len(s),
                       out.ctypes.data_as(ctypes.c_void_p), GV_MEM_HOST))
    return out


def stream_key(seed, stream):
    """Key of generator stream `stream` under `seed`: mix(mix(seed) ^ stream*C)
    with the SplitMix64 finaliser, so every (seed, stream) pair names an
    independent counter-based stream."""
    M = (1 << 64) - 1

    def fmix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    return fmix(fmix(int(seed)) ^ ((int(stream) * 0xD6E8FEB86659FD93) & M))


# Device generators (counter-based SplitMix64 streams; identical streams to
# the host restatement, so batches can be produced on the GPU at scale).
def gen_rmat(graph, key, scale, n, a=0.57, b=0.19, c=0.19):
    from .hip import DeviceArray
    s4 = 4294967296.0
    src, dst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    _check(lib().gv_gen_rmat(graph.handle, key, scale, n, int(a * s4), int((a + b) * s4),
                             int((a + b + c) * s4), src.ptr, dst.ptr))
    return src, dst


def gen_uniform(graph, key, V, n):
    from .hip import DeviceArray
    src, dst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    _check(lib().gv_gen_uniform(graph.handle, key, V, n, src.ptr, dst.ptr))
    return src, dst


def gen_sample(graph, key, V, offsets_dev, cols_dev, n):
    """Sample n edges of a sorted device CSR (offsets uint64[V+1], cols uint32)."""
    from .hip import DeviceArray
    src, dst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    _check(lib().gv_gen_sample(graph.handle, key, V, offsets_dev.ptr, cols_dev.ptr, n, src.ptr,
                               dst.ptr))
    return src, dst


def read_matrix_market(path):
    """(num_vertices, src, dst, symmetric) of a MatrixMarket coordinate file,
    0-based uint32 host arrays (reference: readFile, marketReader.cu:3-121)."""
    V, n, sym = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_int()
    p = os.fsencode(path)
    _check(lib().gv_read_matrix_market(p, ctypes.byref(V), ctypes.byref(n), None, None,
                                       ctypes.byref(sym)))
    src = np.empty(max(n.value, 1), np.uint32)
    dst = np.empty(max(n.value, 1), np.uint32)
    _check(lib().gv_read_matrix_market(p, ctypes.byref(V), ctypes.byref(n),
                                       src.ctypes.data_as(ctypes.c_void_p),
                                       dst.ctypes.data_as(ctypes.c_void_p), ctypes.byref(sym)))
    return int(V.value), src[: n.value], dst[: n.value], bool(sym.value)


def gen_chung_lu(graph, key, V, n, gamma=2.1, avg_deg=24.0, max_deg=200000.0):
    """n Chung-Lu edges (endpoints drawn in proportion to power-law weights)."""
    from .hip import DeviceArray
    src, dst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    _check(lib().gv_gen_chung_lu(graph.handle, key, V, float(gamma), float(avg_deg),
                                 float(max_deg), n, src.ptr, dst.ptr))
    return src, dst


class Capture:
    """A recorded update/compaction pipeline (gv_capture)."""

    def __init__(self, graph, handle):
        self.graph = graph
        self._c = handle

    def launch(self, stats=False):
        st = GvBatchStats()
        _check(lib().gv_capture_launch(self._c, ctypes.byref(st) if stats else None))
        return st.as_dict() if stats else None

    def close(self):
        if getattr(self, "_c", None) is not None:
            lib().gv_capture_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _CaptureScope:
    def __init__(self, graph):
        self.graph = graph
        self.result = None

    def __enter__(self):
        _check(lib().gv_capture_begin(self.graph.handle))
        return self

    def __exit__(self, et, ev, tb):
        c = ctypes.c_void_p()
        rc = lib().gv_capture_end(self.graph.handle, ctypes.byref(c))
        if et is None:
            _check(rc)
            self.result = Capture(self.graph, c)
        elif rc == 0:
            lib().gv_capture_destroy(c)
        return False


def export_csr_device(graph):
    """Sorted CSR as device arrays (offsets uint64[V+1], cols uint32)."""
    from .hip import DeviceArray
    off = DeviceArray(graph.num_vertices + 1, np.uint64)
    _check(lib().gv_export_csr(graph.handle, off.ptr, None, GV_MEM_DEVICE))
    E = int(off.numpy()[-1])
    cols = DeviceArray(max(E, 1), np.uint32)
    _check(lib().gv_export_csr(graph.handle, off.ptr, cols.ptr, GV_MEM_DEVICE))
    return off, cols, E
