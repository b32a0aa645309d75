"""GPU test of gv_neighbors_noalloc, the lookup entry point that allocates
no device memory: called through the C ABI with device buffers, its answers
must equal gv_neighbors' (bit-exact offsets and sorted neighbour lists) on
two calls in a row, the second with a smaller input, and sizes beyond its
stated maximum must be refused."""
import ctypes

import numpy as np
import pytest

import gv_amd as G
from gv_amd.hip import DeviceArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def noalloc(g, vs):
    """Two-call protocol of gv_neighbors_noalloc on device buffers."""
    L = G.lib()
    n = len(vs)
    dvs = DeviceArray.from_numpy(np.ascontiguousarray(vs, dtype=np.uint32))
    doffs = DeviceArray(n + 1, np.uint64)
    rc = L.gv_neighbors_noalloc(g.handle, dvs.ptr, n, doffs.ptr, None)
    if rc:
        return rc, None, None
    offs = doffs.numpy()
    dout = DeviceArray(max(int(offs[-1]), 1), np.uint32)
    rc = L.gv_neighbors_noalloc(g.handle, dvs.ptr, n, doffs.ptr, dout.ptr)
    if rc:
        return rc, None, None
    return rc, doffs.numpy(), dout.numpy(int(offs[-1]))


def test_noalloc_equals_neighbors_on_two_calls():
    V = 1 << 14
    cfg = G.default_config()
    cfg.pool_blocks = 8 * V
    g = G.GraphVine(V, directed=False, config=cfg)
    s, d = O.gen_rmat(O.stream_key(31, O.STREAM_RMAT), 14, 16 * V)
    g.bulkBuild(s, d)
    rng = np.random.default_rng(31)
    # first call: 5000 queries including the hubs (low IDs), repeats and
    # isolated vertices; second call: a smaller, different set
    q1 = np.concatenate([np.arange(64, dtype=np.uint32),
                         rng.integers(0, V, 5000 - 64).astype(np.uint32)])
    q2 = rng.integers(0, V, 300).astype(np.uint32)
    for q in (q1, q2):
        rc, offs, out = noalloc(g, q)
        assert rc == 0, G.lib().gv_last_error()
        eo, ev = g.neighbors(q)
        assert np.array_equal(offs, eo)
        assert np.array_equal(out, ev)
    g.close()


def test_noalloc_refuses_sizes_beyond_its_maximum():
    V = 1 << 16
    cfg = G.default_config()
    cfg.pool_blocks = 8 * V
    g = G.GraphVine(V, directed=False, config=cfg)
    hub = np.zeros(V - 1, dtype=np.uint32)
    g.bulkBuild(hub, np.arange(1, V, dtype=np.uint32))  # vertex 0 has V-1 neighbours
    L = G.lib()
    # 33 queries of the hub: 33 * 65535 > 2^21 neighbours
    rc, _, _ = noalloc(g, np.zeros(33, dtype=np.uint32))
    assert rc == G.GV_E_INVALID
    assert b"maximum" in L.gv_last_error()
    # more than 65536 queries
    rc, _, _ = noalloc(g, np.ones(65537, dtype=np.uint32))
    assert rc == G.GV_E_INVALID
    # a call within the limits still works after the refusals
    q = np.array([0, 5, 7], dtype=np.uint32)
    rc, offs, out = noalloc(g, q)
    assert rc == 0
    eo, ev = g.neighbors(q)
    assert np.array_equal(offs, eo) and np.array_equal(out, ev)
    g.close()
