"""GPU parity of PageRank, triangle counting and betweenness centrality.

Tolerances (FP32 device arithmetic vs the FP64 oracle):
  PageRank  L1 <= 1e-6 at the same iteration count (and the device's own
            stopping iteration within 1 of the oracle's);
  TC        exact (uint64 per vertex and total);
  BC        relative 1e-5 per vertex (|g - o| <= 1e-5 * max(|o|, 1)).
"""
import itertools
import math

import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

PR_L1 = 1e-6
BC_REL = 1e-5


def make(V, directed, pool=None, **kw):
    cfg = G.default_config(**kw)
    cfg.pool_blocks = pool if pool is not None else 4 * V + 4096
    return G.GraphVine(V, directed=directed, config=cfg)


def build_pair(V, directed, s, d, **kw):
    g = make(V, directed, **kw)
    o = O.OracleGraph(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    return g, o


def check_pagerank(g, o, eps=1e-6, max_iter=100, same_stop=True):
    x, it, r = G.static_pagerank(g, 0.85, eps, max_iter)
    xo, ito, _ = o.pagerank(0.85, eps, max_iter)
    if same_stop:
        assert abs(it - ito) <= 1
    else:
        # eps below the FP32 iterate's rounding floor: the device runs to
        # max_iter or stops on an exact FP32 fixed point
        assert it >= ito - 1
    xr, _, _ = o.pagerank(0.85, eps, max_iter, fixed_iters=it)
    l1 = float(np.abs(x.astype(np.float64) - xr).sum())
    assert l1 <= PR_L1, l1
    return l1, it


def check_bc(g, o, sources):
    b = G.bc(g, sources).astype(np.float64)
    bo = o.bc(sources)
    err = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
    assert err.max() <= BC_REL, err.max()


def test_pagerank_closed_forms():
    V = 9
    g, o = build_pair(V, True, np.arange(V), (np.arange(V) + 1) % V)
    x, it, r = G.static_pagerank(g, 0.85, 1e-6, 50)
    assert it == 1 and np.allclose(x, 1.0 / V, rtol=1e-6)
    g2, _ = build_pair(2, True, [0], [1])
    x, it, _ = G.static_pagerank(g2, 0.85, 1e-30, 3)
    a = 0.85
    xa, xb = 0.5, 0.5
    for _ in range(3):
        xa, xb = (1 - a) / 2, (1 - a) / 2 + a * xa
    assert it == 3 and abs(x[0] - xa) < 1e-7 and abs(x[1] - xb) < 1e-7


@pytest.mark.parametrize("layout", [1, 4, 5])
@pytest.mark.parametrize("directed", [True, False])
def test_pagerank_config1_after_updates(directed, layout):
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 16 * V)
    g, o = build_pair(V, directed, s, d, pool=6 * V, pr_layout=layout)
    check_pagerank(g, o)
    for b in range(3):
        si, di = O.gen_uniform(O.stream_key(1, 1000 + b), V, 4096)
        g.batchInsert(si, di)
        o.insert(si, di)
        off, cols = o.export_csr()
        sd, dd = O.gen_sample(O.stream_key(1, 2000 + b), V, off, cols, 4096)
        g.batchDelete(sd, dd)  # tombstones left in place on purpose
        o.delete(sd, dd)
        check_pagerank(g, o)
    g.compaction(0.0)
    check_pagerank(g, o)


def test_pagerank_iteration_limits_and_repeated_calls():
    """The flat form issues iterations in groups of 8 and reads the done flag
    one group behind: the iteration count must still stop exactly at max_iter
    on either side of a group boundary, and the launches queued after a stop
    must leave the next call on the same handle unaffected."""
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(5, O.STREAM_RMAT), 12, 16 * V)
    g, o = build_pair(V, False, s, d)
    for max_iter in (1, 7, 8, 9, 15, 16, 17, 25):
        # eps 0 never converges: every call runs to its limit
        x, it, _ = G.static_pagerank(g, 0.85, 0.0, max_iter)
        assert it == max_iter
        xr, _, _ = o.pagerank(0.85, 0.0, max_iter, fixed_iters=max_iter)
        assert float(np.abs(x.astype(np.float64) - xr).sum()) <= PR_L1
    # converging calls back to back: same answer every time
    first = None
    for _ in range(3):
        l1, it = check_pagerank(g, o)
        x, it2, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
        assert it2 == it
        if first is None:
            first = x
        assert np.array_equal(x, first)


@pytest.mark.parametrize("layout", [1, 4, 5])
def test_pagerank_all_degree_bins_and_chunked_hubs(layout):
    # small thresholds push vertices through every bin incl. chunked hubs
    V = 3000
    s, d = O.gen_rmat(O.stream_key(3, O.STREAM_RMAT), 12, 60000)
    s, d = s % V, d % V
    star = np.zeros(V - 1, np.uint32), np.arange(1, V, dtype=np.uint32)
    s = np.concatenate([s, star[0]])
    d = np.concatenate([d, star[1]])
    g, o = build_pair(V, False, s, d, pool=8 * V, pr_thread_max=2, pr_group_max=16,
                      pr_wave_max=128, pr_layout=layout)
    check_pagerank(g, o)
    check_pagerank(g, o, eps=1e-9, max_iter=200, same_stop=False)
    # tombstones on the hub (degree > one snapshot chunk) and elsewhere
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(3, 2000), V, off, cols, 2500)
    sd = np.concatenate([sd, np.zeros(700, np.uint32)])
    dd = np.concatenate([dd, np.arange(1, 1400, 2, dtype=np.uint32)])
    g.batchDelete(sd, dd)
    o.delete(sd, dd)
    check_pagerank(g, o)


@pytest.mark.parametrize("layout", [1, 4, 5])
@pytest.mark.parametrize("V", [40000, 1 << 17])
@pytest.mark.parametrize("directed", [False, True])
def test_pagerank_staged_hot_and_cold_sources(V, directed, layout):
    # more vertices than the 32768 contributions staged in LDS, so gathers
    # take both the LDS and the global path; the star hub's row spans many
    # stream tiles (layout 1) or chunks (layout 4, with small bin limits that
    # put vertices in every bin); directed graphs have empty in-rows
    s, d = O.gen_rmat(O.stream_key(5, O.STREAM_RMAT), 17, 12 * V)
    s, d = s % V, d % V
    s = np.concatenate([s, np.full(V - 1, 7, np.uint32)])
    d = np.concatenate([d, np.arange(V, dtype=np.uint32)[np.arange(V) != 7]])
    g, o = build_pair(V, directed, s, d, pool=24 * V, pr_thread_max=4, pr_group_max=48,
                      pr_wave_max=1024, pr_layout=layout)
    check_pagerank(g, o)
    x1, it1, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    x2, it2, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    assert it1 == it2 and np.array_equal(x1, x2)  # deterministic
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(5, 2000), V, off, cols, V // 4)
    g.batchDelete(sd, dd)  # tombstones: the snapshot keeps live entries only
    o.delete(sd, dd)
    check_pagerank(g, o)
    g.close()


@pytest.mark.parametrize("directed", [False, True])
def test_pagerank_propagation_blocked_many_blocks_and_bins(directed):
    """Layout 5 at a size with many cold source blocks (16384 labels each,
    past the 32767 hot labels) and many 31-tile bins: rows empty in one
    stream or both (padding entries), a hub row spanning many tiles of the
    cold stream and hub sources in many hot rows, a warm start, plain sums,
    and the result after deletes and inserts, all at L1 <= 1e-6 against the
    oracle at equal iterations; reruns are bit-identical."""
    V = 300_000
    s, d = O.gen_rmat(O.stream_key(52, O.STREAM_RMAT), 19, 10 * V)
    s, d = s % V, d % V
    hub = 12345
    s = np.concatenate([s, np.full(60000, hub, np.uint32)])
    d = np.concatenate([d, np.arange(100000, 160000, dtype=np.uint32)])
    g, o = build_pair(V, directed, s, d, pool=24 * V, pr_layout=5)
    check_pagerank(g, o)
    x1, it1, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    x2, it2, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
    assert it1 == it2 and np.array_equal(x1, x2)
    xw, itw, _ = G.pagerank(g, init=x1 * 1.5)
    xo, _, _ = o.pagerank(fixed_iters=itw, x0=(x1 * 1.5).astype(np.float64))
    assert np.abs(xw.astype(np.float64) - xo).sum() <= PR_L1
    xp, itp, _ = G.pagerank(g, compensated=False)
    xo, _, _ = o.pagerank(fixed_iters=itp)
    assert np.abs(xp.astype(np.float64) - xo).sum() <= PR_L1
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(52, 2000), V, off, cols, V)
    g.batchDelete(sd, dd)
    o.delete(sd, dd)
    si, di = O.gen_uniform(O.stream_key(52, 3000), V, V)
    g.batchInsert(si, di)
    o.insert(si, di)
    check_pagerank(g, o)
    g.close()


@pytest.mark.parametrize("layout", [1, 4, 5])
@pytest.mark.parametrize("V", [5, 9, 20001, 36863, 70000])
def test_pagerank_odd_sizes_and_large_hub(V, layout):
    # hub with > 16384 in-edges, vertex counts that are not powers of two,
    # tiny graphs; for layout 5, 20001 leaves a partial last group of 4 in
    # the LDS-staged hot labels and 36863 a partial last group in the last
    # source block's expansion window (3 labels)
    rng = np.random.default_rng(V)
    s = rng.integers(0, V, 6 * V).astype(np.uint32)
    d = rng.integers(0, V, 6 * V).astype(np.uint32)
    s = np.concatenate([s, np.zeros(V - 1, np.uint32)])
    d = np.concatenate([d, np.arange(1, V, dtype=np.uint32)])
    for directed in (False, True):
        g, o = build_pair(V, directed, s, d, pool=16 * V + 4096, pr_layout=layout)
        check_pagerank(g, o)
        x1, it1, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
        x2, it2, _ = G.static_pagerank(g, 0.85, 1e-6, 100)
        assert it1 == it2 and np.array_equal(x1, x2)  # deterministic
        g.close()


def test_pagerank_cycle_of_2_pow_24_vertices():
    # the relabelling sort runs on exactly 2^24 keys (the largest single-sweep
    # sort) with a workspace sized for V + 1 keys, and the build stages 2^24
    # keys; a directed cycle has x = 1/V at every iteration
    V = 1 << 24
    s = np.arange(V, dtype=np.uint32)
    d = (s + 1) % V
    g = make(V, True, pool=V + 4096, block_slots=32)
    g.bulkBuild(s, d, stats=False)
    x, it, r = G.static_pagerank(g, 0.85, 1e-6, 20)
    assert it == 1 and r < 1e-6
    assert np.allclose(x, 1.0 / V, rtol=1e-6)
    g.close()


@pytest.mark.parametrize("n", [3, 7, 40])
def test_tc_complete_graph(n):
    e = list(itertools.combinations(range(n), 2))
    g, o = build_pair(n, False, [a for a, _ in e], [b for _, b in e])
    pv, tot = G.static_tc(g)
    assert tot == math.comb(n, 3)
    assert pv.tolist() == [math.comb(n - 1 - i, 2) for i in range(n)]


def test_tc_config1_symmetrized_exact():
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 16 * V)
    g, o = build_pair(V, False, s, d, pool=6 * V)
    pv, tot = G.static_tc(g)
    pvo, toto = o.triangles()
    assert tot == toto and np.array_equal(pv, pvo)
    _, tot2 = G.static_tc(g, per_vertex=False)
    assert tot2 == toto
    # after updates with tombstones
    si, di = O.gen_uniform(O.stream_key(1, 1000), V, 4096)
    g.batchInsert(si, di)
    o.insert(si, di)
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(1, 2000), V, off, cols, 4096)
    g.batchDelete(sd, dd)
    o.delete(sd, dd)
    pv, tot = G.static_tc(g)
    pvo, toto = o.triangles()
    assert tot == toto and np.array_equal(pv, pvo)


def test_tc_rejects_directed():
    g, _ = build_pair(10, True, [0, 1], [1, 2])
    with pytest.raises(G.GraphVineError) as e:
        G.static_tc(g)
    assert e.value.status == G.GV_E_UNSUPPORTED


def test_bc_closed_forms():
    n = 12
    g, _ = build_pair(n, False, np.arange(n - 1), np.arange(1, n))
    b = G.bc(g, np.arange(n))
    assert np.allclose(b, [2 * i * (n - 1 - i) for i in range(n)], rtol=1e-6)
    g2, _ = build_pair(n, False, np.zeros(n - 1), np.arange(1, n))
    b2 = G.bc(g2, np.arange(n))
    assert abs(b2[0] - (n - 1) * (n - 2)) < 1e-3 and np.allclose(b2[1:], 0)


def test_bc_config1_16_sources():
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 16 * V)
    g, o = build_pair(V, True, s, d, pool=6 * V)
    rng = np.random.default_rng(1)
    src = rng.choice(V, 16, replace=False).astype(np.uint32)
    check_bc(g, o, src)
    # undirected view, with tombstones present
    gu, ou = build_pair(V, False, s, d, pool=6 * V)
    off, cols = ou.export_csr()
    sd, dd = O.gen_sample(O.stream_key(1, 2000), V, off, cols, 4096)
    gu.batchDelete(sd, dd)
    ou.delete(sd, dd)
    check_bc(gu, ou, src)


@pytest.mark.parametrize("directed", [True, False])
def test_bc_single_source_calls(directed):
    # one source per call: the level-synchronous kernels (not the 32-source
    # batches), each level's frontier issued and extended in one piece
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 16 * V)
    g, o = build_pair(V, directed, s, d, pool=6 * V)
    rng = np.random.default_rng(7)
    for src in rng.choice(V, 4, replace=False).astype(np.uint32):
        check_bc(g, o, np.array([src], np.uint32))


@pytest.mark.parametrize("directed", [False, True])
def test_bc_single_source_hub_rows(directed):
    """Frontier rows longer than 8192 slots are spread over the whole grid in
    both sweeps (partials joined in block order backward): two hubs of 9000
    and 30000 neighbours over an R-MAT graph, sources on and off the hubs."""
    V = 40000
    rng = np.random.default_rng(8)
    s, d = O.gen_rmat(O.stream_key(23, O.STREAM_RMAT), 15, 120000)
    s, d = s % V, d % V
    n0 = rng.choice(np.arange(2, V), 9000, replace=False).astype(np.uint32)
    n1 = rng.choice(np.arange(2, V), 30000, replace=False).astype(np.uint32)
    s = np.concatenate([s, np.zeros(9000, np.uint32), np.ones(30000, np.uint32), n1[:500]])
    d = np.concatenate([d, n0, n1, np.zeros(500, np.uint32)]).astype(np.uint32)
    g, o = build_pair(V, directed, s.astype(np.uint32), d, pool=16 * V)
    for src in ([0], [1], [int(n1[7])], [5, 1]):
        check_bc(g, o, np.array(src, np.uint32))
    g.close()


def test_bc_rejects_bad_source():
    g, _ = build_pair(10, False, [0, 1], [1, 2])
    with pytest.raises(G.GraphVineError) as e:
        G.bc(g, [3, 10])
    assert e.value.status == G.GV_E_RANGE
