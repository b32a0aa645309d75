"""Python binding of the request loop (include/gv_serve.h).

    srv = Server(graph)
    reqs = [insert(src, dst), neighbors(vs, out_cap=...), pagerank(V), bc(V, 7)]
    resps = srv.replay(reqs, arrival_us)      # open-loop arrivals
    reqs[1].neighbours()                      # results live on the request

Request arrays must stay alive until the request completes; the Request
objects hold them.
"""
import ctypes

import numpy as np

from . import GV_MEM_DEVICE, GV_MEM_HOST, _check, _ptr_u32, lib

GV_REQ_INSERT = 0
GV_REQ_DELETE = 1
GV_REQ_NEIGHBORS = 2
GV_REQ_PAGERANK = 3
GV_REQ_BC = 4

P = ctypes.c_void_p


class GvRequest(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int),
        ("mem", ctypes.c_int),
        ("src", P),
        ("dst", P),
        ("n", ctypes.c_uint64),
        ("vs", P),
        ("offs", P),
        ("out", P),
        ("out_cap", ctypes.c_uint64),
        ("source", ctypes.c_uint32),
        ("x", P),
    ]


class GvResponse(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("kind", ctypes.c_int),
        ("arrival_us", ctypes.c_double),
        ("start_us", ctypes.c_double),
        ("done_us", ctypes.c_double),
        ("step", ctypes.c_uint64),
        ("step_size", ctypes.c_uint32),
        ("iters", ctypes.c_int),
        ("residual", ctypes.c_double),
        ("out_len", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class GvServerOpts(ctypes.Structure):
    _fields_ = [
        ("max_step_pairs", ctypes.c_uint64),
        ("max_step_requests", ctypes.c_uint32),
        ("headroom_blocks", ctypes.c_uint64),
        ("compact_fraction", ctypes.c_float),
        ("pr_alpha", ctypes.c_float),
        ("pr_eps", ctypes.c_double),
        ("pr_max_iter", ctypes.c_int),
        ("analytics_cu_share", ctypes.c_float),
    ]


class GvServerStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in ("requests_done", "steps", "update_pairs_applied",
                                               "refused", "free_blocks", "total_blocks")]


_bound = False


def _lib():
    global _bound
    L = lib()
    if not _bound:
        u64 = ctypes.c_uint64
        sig = {
            "gv_server_opts_default": [ctypes.POINTER(GvServerOpts)],
            "gv_server_create": [P, ctypes.POINTER(GvServerOpts), ctypes.POINTER(P)],
            "gv_server_submit": [P, ctypes.POINTER(GvRequest), ctypes.POINTER(u64)],
            "gv_server_wait": [P, u64, ctypes.POINTER(GvResponse)],
            "gv_server_replay": [P, ctypes.POINTER(GvRequest), ctypes.POINTER(ctypes.c_double),
                                 u64, ctypes.POINTER(GvResponse)],
            "gv_server_drain": [P],
            "gv_server_get_stats": [P, ctypes.POINTER(GvServerStats)],
            "gv_server_destroy": [P],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        _bound = True
    return L


class Request:
    """One request and the buffers it reads and writes."""

    def __init__(self, kind):
        self.c = GvRequest()
        self.c.kind = kind
        self.keep = []
        self.response = None

    # results (valid once the request is done)
    def neighbours(self):
        """(offsets[n+1], concatenated sorted neighbour lists)."""
        r = self.response
        return self._offs, self._out[: int(r["out_len"])] if r else None

    def vector(self):
        return self._x


def _update(kind, src, dst):
    r = Request(kind)
    ps, ks, ms = _ptr_u32(src)
    pd, kd, md = _ptr_u32(dst)
    if ms != md:
        raise ValueError("src and dst must both be host or both be device arrays")
    if len(src) != len(dst):
        raise ValueError("src and dst lengths differ")
    r.c.mem = ms
    r.c.src, r.c.dst, r.c.n = ps, pd, len(src)
    r.keep += [ks, kd]
    return r


def insert(src, dst):
    return _update(GV_REQ_INSERT, src, dst)


def delete(src, dst):
    return _update(GV_REQ_DELETE, src, dst)


def neighbors(vs, out_cap=None):
    """1-hop lookup of vs; out_cap=None keeps only the offsets (the lists
    are produced but not copied out)."""
    r = Request(GV_REQ_NEIGHBORS)
    v = np.ascontiguousarray(vs, dtype=np.uint32)
    r._offs = np.zeros(len(v) + 1, np.uint64)
    r._out = np.empty(max(int(out_cap or 0), 1), np.uint32)
    r.c.vs = v.ctypes.data_as(P)
    r.c.n = len(v)
    r.c.offs = r._offs.ctypes.data_as(P)
    if out_cap is not None:
        r.c.out = r._out.ctypes.data_as(P)
        r.c.out_cap = int(out_cap)
    r.keep += [v]
    return r


def pagerank(V=None):
    """PageRank refresh; V given -> the vector is copied out."""
    r = Request(GV_REQ_PAGERANK)
    r._x = None
    if V is not None:
        r._x = np.empty(int(V), np.float32)
        r.c.x = r._x.ctypes.data_as(P)
    return r


def bc(source, V=None):
    r = Request(GV_REQ_BC)
    r.c.source = int(source)
    r._x = None
    if V is not None:
        r._x = np.empty(int(V), np.float32)
        r.c.x = r._x.ctypes.data_as(P)
    return r


class Server:
    def __init__(self, graph, **opts):
        o = GvServerOpts()
        _check(_lib().gv_server_opts_default(ctypes.byref(o)))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise KeyError(k)
            setattr(o, k, v)
        h = P()
        _check(_lib().gv_server_create(graph.handle, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self.graph = graph
        self.opts = o
        self._live = {}

    def submit(self, req):
        t = ctypes.c_uint64()
        _check(_lib().gv_server_submit(self._h, ctypes.byref(req.c), ctypes.byref(t)))
        self._live[int(t.value)] = req
        return int(t.value)

    def wait(self, ticket):
        r = GvResponse()
        _check(_lib().gv_server_wait(self._h, ticket, ctypes.byref(r)))
        req = self._live.pop(ticket, None)
        d = r.as_dict()
        if req is not None:
            req.response = d
        return d

    def replay(self, reqs, arrival_us):
        n = len(reqs)
        arr = (GvRequest * n)(*[r.c for r in reqs])
        t = np.ascontiguousarray(arrival_us, dtype=np.float64)
        if len(t) != n:
            raise ValueError("one arrival time per request")
        out = (GvResponse * n)()
        _check(_lib().gv_server_replay(self._h, arr, t.ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), n, out))
        res = [out[i].as_dict() for i in range(n)]
        for r, d in zip(reqs, res):
            r.response = d
        return res

    def drain(self):
        _check(_lib().gv_server_drain(self._h))

    def stats(self):
        s = GvServerStats()
        _check(_lib().gv_server_get_stats(self._h, ctypes.byref(s)))
        return {f: getattr(s, f) for f, _ in s._fields_}

    def close(self):
        if getattr(self, "_h", None) is not None:
            _lib().gv_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
