"""ctypes wrapper of the CPU oracle (gv_oracle.c) and the OpenMP CSR baseline.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Never part of the product path.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

# streams of the counter-based generators
STREAM_RMAT = 1
STREAM_INSERT = 2
STREAM_DELETE = 3
STREAM_CHUNG_LU = 4
STREAM_SOURCES = 5


def build():
    r = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    u32, u64, i32, f64 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
    sig = {
        "orc_fmix64": ([u64], u64),
        "orc_stream_key": ([u64, u64], u64),
        "orc_rand64": ([u64, u64], u64),
        "orc_gen_rmat": ([u64, i32, u64, u32, u32, u32, P, P], None),
        "orc_gen_uniform": ([u64, u32, u64, P, P], None),
        "orc_gen_sample": ([u64, u32, P, P, u64, P, P], None),
        "orc_gen_chung_lu": ([u64, u32, f64, f64, f64, u64, P, P], None),
        "orc_create": ([u32, i32], P),
        "orc_destroy": ([P], None),
        "orc_bulk_build": ([P, P, P, u64, ctypes.POINTER(u64)], i32),
        "orc_batch_insert": ([P, P, P, u64, ctypes.POINTER(u64)], i32),
        "orc_batch_delete": ([P, P, P, u64, ctypes.POINTER(u64)], i32),
        "orc_num_edges": ([P], u64),
        "orc_num_vertices": ([P], u32),
        "orc_degrees": ([P, P], None),
        "orc_export_csr": ([P, P, P], None),
        "orc_pagerank": ([P, f64, f64, i32, i32, P, P, ctypes.POINTER(f64)], i32),
        "orc_pagerank_personalized": ([P, f64, f64, i32, i32, P, u32, P, ctypes.POINTER(f64)], i32),
        "orc_triangle_count": ([P, P], u64),
        "orc_bc": ([P, P, u32, P], None),
        "omp_apply_batch": ([u32, i32, P, P, P, P, u64, i32, P, P], u64),
        "omp_pagerank_sym": ([u32, P, P, f64, f64, i32, i32, P, P, ctypes.POINTER(f64)], i32),
        "omp_triangle_count": ([u32, P, P, P], u64),
        "omp_threads": ([], i32),
        "omp_transpose": ([u32, P, P, P, P], None),
        "omp_pagerank_dir": ([u32, P, P, P, f64, f64, i32, i32, P, ctypes.POINTER(f64)], i32),
        "omp_bc_dir": ([u32, P, P, P, u32, P], None),
        "omp_pagerank_personalized": ([u32, P, P, f64, f64, i32, i32, P, u32, P,
                                       ctypes.POINTER(f64)], i32),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def stream_key(seed, stream):
    return int(lib().orc_stream_key(seed, stream))


def thresholds(a=0.57, b=0.19, c=0.19):
    """R-MAT cumulative quadrant thresholds scaled by 2^32."""
    s = 4294967296.0
    return int(a * s), int((a + b) * s), int((a + b + c) * s)


def gen_rmat(key, scale, n, a=0.57, b=0.19, c=0.19):
    src = np.empty(n, np.uint32)
    dst = np.empty(n, np.uint32)
    ta, tab, tabc = thresholds(a, b, c)
    lib().orc_gen_rmat(key, scale, n, ta, tab, tabc, _p(src), _p(dst))
    return src, dst


def gen_uniform(key, V, n):
    src = np.empty(n, np.uint32)
    dst = np.empty(n, np.uint32)
    lib().orc_gen_uniform(key, V, n, _p(src), _p(dst))
    return src, dst


def gen_sample(key, V, offsets, cols, n):
    src = np.empty(n, np.uint32)
    dst = np.empty(n, np.uint32)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    c = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    lib().orc_gen_sample(key, V, _p(off), _p(c), n, _p(src), _p(dst))
    return src, dst


def gen_chung_lu(key, V, gamma, avg_deg, max_deg, n):
    src = np.empty(n, np.uint32)
    dst = np.empty(n, np.uint32)
    lib().orc_gen_chung_lu(key, V, gamma, avg_deg, max_deg, n, _p(src), _p(dst))
    return src, dst


class OracleGraph:
    """Host-sequential set-based dynamic graph (the contract)."""

    def __init__(self, V, directed=False):
        self.V = int(V)
        self.directed = bool(directed)
        self._g = lib().orc_create(self.V, 1 if directed else 0)

    def close(self):
        if self._g:
            lib().orc_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _apply(self, fn, src, dst):
        s, d = u32(src), u32(dst)
        c = ctypes.c_uint64()
        rc = fn(self._g, _p(s), _p(d), len(s), ctypes.byref(c))
        if rc != 0:
            raise ValueError("vertex ID out of range")
        return int(c.value)

    def bulk_build(self, src, dst):
        return self._apply(lib().orc_bulk_build, src, dst)

    def insert(self, src, dst):
        return self._apply(lib().orc_batch_insert, src, dst)

    def delete(self, src, dst):
        return self._apply(lib().orc_batch_delete, src, dst)

    def num_edges(self):
        return int(lib().orc_num_edges(self._g))

    def degrees(self):
        d = np.empty(self.V, np.uint32)
        lib().orc_degrees(self._g, _p(d))
        return d

    def export_csr(self):
        off = np.empty(self.V + 1, np.uint64)
        lib().orc_export_csr(self._g, _p(off), None)
        cols = np.empty(max(int(off[-1]), 1), np.uint32)
        lib().orc_export_csr(self._g, _p(off), _p(cols))
        return off, cols[: int(off[-1])]

    def pagerank(self, alpha=0.85, eps=1e-6, max_iter=100, fixed_iters=0, x0=None):
        """x0: start vector (default 1/V), e.g. the float32 result of an
        earlier run for a warm-started refresh."""
        x = np.empty(self.V, np.float64)
        r = ctypes.c_double()
        start = None
        if x0 is not None:
            start = np.ascontiguousarray(x0, dtype=np.float64)
            if len(start) != self.V:
                raise ValueError("x0 must have V entries")
        it = lib().orc_pagerank(self._g, alpha, eps, max_iter, fixed_iters,
                                _p(start) if start is not None else None, _p(x), ctypes.byref(r))
        return x, int(it), float(r.value)

    def pagerank_personalized(self, seeds, alpha=0.85, eps=1e-6, max_iter=100, fixed_iters=0):
        sd = u32(seeds)
        x = np.empty((self.V, len(sd)), np.float64)
        r = ctypes.c_double()
        it = lib().orc_pagerank_personalized(self._g, alpha, eps, max_iter, fixed_iters, _p(sd),
                                             len(sd), _p(x), ctypes.byref(r))
        return x, int(it), float(r.value)

    def triangles(self):
        pv = np.empty(self.V, np.uint64)
        tot = lib().orc_triangle_count(self._g, _p(pv))
        return pv, int(tot)

    def bc(self, sources):
        s = u32(sources)
        out = np.empty(self.V, np.float64)
        lib().orc_bc(self._g, _p(s) if len(s) else None, len(s), _p(out))
        return out


def omp_apply_batch(V, directed, off, cols, src, dst, delete):
    """OpenMP CSR rebuild: returns the new (offsets, cols)."""
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    s, d = u32(src), u32(dst)
    cap = int(off[-1]) + (0 if delete else (len(s) if directed else 2 * len(s)))
    off_out = np.empty(V + 1, np.uint64)
    cols_out = np.empty(max(cap, 1), np.uint32)
    n = lib().omp_apply_batch(V, 1 if directed else 0, _p(off), _p(cols), _p(s), _p(d), len(s),
                              1 if delete else 0, _p(off_out), _p(cols_out))
    return off_out, cols_out[:n]


def omp_pagerank_sym(V, off, cols, alpha=0.85, eps=1e-6, max_iter=100, fixed_iters=0, x0=None,
                     residual=False):
    """FP64 pull PageRank of a symmetric CSR (the oracle's rule); x0: start
    vector (default 1/V); fixed_iters > 0 runs exactly that many updates.
    Returns (x, iters) or, with residual=True, (x, iters, last residual)."""
    x = np.empty(V, np.float64)
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    start = None if x0 is None else np.ascontiguousarray(x0, np.float64)
    if start is not None and len(start) != V:
        raise ValueError("x0 must have V entries")
    r = ctypes.c_double()
    it = lib().omp_pagerank_sym(V, _p(off), _p(cols), alpha, eps, max_iter, fixed_iters,
                                _p(start) if start is not None else None, _p(x), ctypes.byref(r))
    return (x, int(it), float(r.value)) if residual else (x, int(it))


def omp_pagerank_personalized(V, off, cols, seeds, alpha=0.85, eps=1e-6, max_iter=100,
                              fixed_iters=0):
    """The oracle's personalized PageRank (OracleGraph.pagerank_personalized)
    on the out-adjacency CSR, columns spread over OpenMP threads: same
    operations in the same order per column, so the same bits.  Returns
    (X[V, ncols] FP64, iters, largest per-column residual)."""
    sd = u32(seeds)
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    X = np.empty((V, len(sd)), np.float64)
    r = ctypes.c_double()
    it = lib().omp_pagerank_personalized(V, _p(off), _p(cols), alpha, eps, max_iter, fixed_iters,
                                         _p(sd), len(sd), _p(X), ctypes.byref(r))
    return X, int(it), float(r.value)


def omp_triangle_count(V, off, cols, per_vertex=True):
    """Exact triangle count of a sorted symmetric CSR (OpenMP, degree-oriented
    lists, one marked neighbour set per thread); returns (per-vertex min-ID
    counts or None, total)."""
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    pv = np.empty(V, np.uint64) if per_vertex else None
    tot = lib().omp_triangle_count(V, _p(off), _p(cols), _p(pv) if pv is not None else None)
    return pv, int(tot)


def omp_transpose(V, off, cols):
    """In-lists (offsets, cols) of a CSR of out-lists."""
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    ioff = np.empty(V + 1, np.uint64)
    icols = np.empty(max(int(off[-1]), 1), np.uint32)
    lib().omp_transpose(V, _p(off), _p(cols), _p(ioff), _p(icols))
    return ioff, icols[: int(off[-1])]


def omp_pagerank_dir(V, off, cols, alpha=0.85, eps=1e-6, max_iter=100, fixed_iters=0):
    """FP64 pull PageRank of the directed graph given by out-list CSR
    (off, cols); returns (x, iters, residual)."""
    off = np.ascontiguousarray(off, np.uint64)
    ioff, icols = omp_transpose(V, off, cols)
    icols = icols if len(icols) else np.zeros(1, np.uint32)
    od = np.diff(off).astype(np.uint32)
    x = np.empty(V, np.float64)
    r = ctypes.c_double()
    it = lib().omp_pagerank_dir(V, _p(ioff), _p(icols), _p(od), alpha, eps, max_iter, fixed_iters,
                                _p(x), ctypes.byref(r))
    return x, int(it), float(r.value)


def omp_bc_dir(V, off, cols, sources):
    off = np.ascontiguousarray(off, np.uint64)
    cols = u32(cols) if len(cols) else np.zeros(1, np.uint32)
    s = u32(sources)
    out = np.empty(V, np.float64)
    lib().omp_bc_dir(V, _p(off), _p(cols), _p(s) if len(s) else None, len(s), _p(out))
    return out


def omp_threads():
    return int(lib().omp_threads())
