"""Captured update -> compact -> PageRank pipelines (gv_pagerank_reserve,
gv_pagerank_run inside gv_capture_*, gv_pagerank_last).  Every replay is
checked against the oracle (graph: sorted sets; PageRank: L1 <= 1e-6 at the
replay's iteration count) and against an eager PageRank call on the same
graph (bitwise), with eager work and synchronisations between replays."""
import numpy as np
import pytest

import gv_amd as G
from gv_amd.hip import DeviceArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu

SIZES = [3000, 1200, 400, 2500]
KINDS = [G.GV_OP_INSERT, G.GV_OP_DELETE, G.GV_OP_INSERT, G.GV_OP_DELETE]


def make(V, directed, **kw):
    cfg = G.default_config(maintain_in_edges=1 if directed else 0, **kw)
    cfg.pool_blocks = 8 * V + 4096
    return G.GraphVine(V, directed=directed, config=cfg)


def assert_same(g, o, where=""):
    off, cols = g.export_csr()
    eo, ec = o.export_csr()
    assert np.array_equal(off, eo), f"offsets differ {where}"
    assert np.array_equal(cols, ec), f"neighbour sets differ {where}"


def fill_group(o, V, seed):
    srcs, dsts = [], []
    for b, (n, k) in enumerate(zip(SIZES, KINDS)):
        if k == G.GV_OP_INSERT:
            s, d = O.gen_uniform(O.stream_key(seed, 10_000 + b), V, n)
            o.insert(s, d)
        else:
            off, cols = o.export_csr()
            s, d = O.gen_sample(O.stream_key(seed, 20_000 + b), V, off, cols, n)
            o.delete(s, d)
        srcs.append(s)
        dsts.append(d)
    return np.concatenate(srcs), np.concatenate(dsts)


@pytest.mark.parametrize("directed,scale", [(False, 16), (True, 15), (False, 13)])
def test_captured_update_compact_pagerank_replays(directed, scale):
    # scale 16 / 15: vertices past the hot labels, so cold source blocks and
    # bins exist; scale 13: every source label is hot
    V = 1 << scale
    s, d = O.gen_rmat(O.stream_key(41, O.STREAM_RMAT), scale, 8 * V)
    g = make(V, directed)
    o = O.OracleGraph(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    off = np.concatenate([[0], np.cumsum(SIZES)]).astype(np.uint64)
    n = int(off[-1])
    dsrc, ddst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    out = DeviceArray(V, np.float32)
    g.reserve(n, len(SIZES))
    per = 1 if directed else 2
    g.reserve_pagerank(o.num_edges() + 4 * per * n)
    with g.capture() as cap:
        g.applyBatches(dsrc, ddst, off, KINDS, compact_fraction=0.25, stats=False)
        G.pagerank_device_run(g, out, alpha=0.85, epsilon=1e-6, max_iter=100, stats=False)
    c = cap.result
    for rep in range(3):
        hs, hd = fill_group(o, V, 900 + rep)
        g.sync()
        dsrc.upload(hs)
        ddst.upload(hd)
        c.launch()
        st = g.pagerank_last()
        assert 0 < st["iters"] <= 100
        x = out.numpy().astype(np.float64)
        assert_same(g, o, f"replay {rep}")
        xr, _, _ = o.pagerank(0.85, 1e-6, 100, fixed_iters=st["iters"])
        l1 = float(np.abs(x - xr).sum())
        assert l1 <= 1e-6, (rep, l1)
        # an eager call on the same graph: same iterations, same bits
        xe, ite, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
        assert ite == st["iters"]
        assert np.array_equal(xe, out.numpy()), f"replay {rep} differs from the eager call"
        # eager updates between replays
        es, ed = O.gen_uniform(O.stream_key(950 + rep, 1), V, 500)
        g.batchInsert(es, ed)
        o.insert(es, ed)
    c.close()
    g.close()


def test_captured_pagerank_beyond_reservation_reports_and_recovers():
    V = 1 << 15
    s, d = O.gen_rmat(O.stream_key(42, O.STREAM_RMAT), 15, 8 * V)
    g = make(V, False)
    o = O.OracleGraph(V, False)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    n = 1000
    dsrc, ddst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    out = DeviceArray(V, np.float32)
    g.reserve(n, 1)
    g.reserve_pagerank(o.num_edges() + 2 * n)
    with g.capture() as cap:
        g.batchInsert(dsrc, ddst, stats=False)
        G.pagerank_device_run(g, out, stats=False)
    c = cap.result
    hs, hd = O.gen_uniform(O.stream_key(43, 1), V, n)
    o.insert(hs, hd)
    dsrc.upload(hs)
    ddst.upload(hd)
    c.launch()
    st = g.pagerank_last()
    xr, _, _ = o.pagerank(0.85, 1e-6, 100, fixed_iters=st["iters"])
    assert float(np.abs(out.numpy().astype(np.float64) - xr).sum()) <= 1e-6
    # the graph grows past the reservation by eager inserts: the replay
    # applies its batch, its PageRank computes nothing and says so
    ws, wd = O.gen_uniform(O.stream_key(44, 1), V, 200_000)
    g.batchInsert(ws, wd)
    o.insert(ws, wd)
    hs, hd = O.gen_uniform(O.stream_key(45, 1), V, n)
    o.insert(hs, hd)
    g.sync()
    dsrc.upload(hs)
    ddst.upload(hd)
    c.launch()
    with pytest.raises(G.GraphVineError):
        g.pagerank_last()
    assert_same(g, o, "after the refused PageRank")
    # an eager call still works (its buffers grow), which retires the capture
    x, it, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    xr, _, _ = o.pagerank(0.85, 1e-6, 100, fixed_iters=it)
    assert float(np.abs(x.astype(np.float64) - xr).sum()) <= 1e-6
    with pytest.raises(G.GraphVineError):
        c.launch()
    c.close()
    g.close()


def test_pagerank_capture_rules():
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(46, O.STREAM_RMAT), 12, 8 * V)
    g = make(V, False)
    g.bulkBuild(s, d)
    out = DeviceArray(V, np.float32)
    # no reservation: the captured call is refused and the capture fails
    with pytest.raises(G.GraphVineError):
        with g.capture():
            G.pagerank_device_run(g, out, stats=False)
    g.reserve_pagerank(g.num_edges() + 1000)
    # statistics cannot be returned inside a capture
    with pytest.raises(G.GraphVineError):
        with g.capture():
            G.pagerank_device_run(g, out, stats=True)
    # host vectors cannot be recorded
    with pytest.raises(G.GraphVineError):
        with g.capture():
            G.static_pagerank(g)
    # the handle still works eagerly
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    x, it, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    xr, _, _ = o.pagerank(0.85, 1e-6, 100, fixed_iters=it)
    assert float(np.abs(x.astype(np.float64) - xr).sum()) <= 1e-6
    g.close()
