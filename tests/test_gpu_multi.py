"""GPU parity of the many-vector analytics and PageRank options: the
64-column personalized PageRank block, multi-source BC (32 sources per
sweep), warm-started and plain-sum PageRank."""
import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def build_pair(V, directed, s, d, **kw):
    cfg = G.default_config(**kw)
    cfg.pool_blocks = 8 * V + 4096
    g = G.GraphVine(V, directed=directed, config=cfg)
    o = O.OracleGraph(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    return g, o


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("ncols", [1, 7, 64])
def test_personalized_pagerank_block(directed, ncols):
    V = 3000
    s, d = O.gen_rmat(O.stream_key(21, O.STREAM_RMAT), 12, 40000)
    g, o = build_pair(V, directed, s % V, d % V)
    seeds = np.random.default_rng(ncols).choice(V, ncols, replace=False).astype(np.uint32)
    X, it, r = G.pagerank_personalized(g, seeds, 0.85, 1e-6, 100)
    Xo, ito, _ = o.pagerank_personalized(seeds, 0.85, 1e-6, 100)
    assert abs(it - ito) <= 1
    Xr, _, _ = o.pagerank_personalized(seeds, 0.85, 1e-6, 100, fixed_iters=it)
    l1 = np.abs(X.astype(np.float64) - Xr).sum(axis=0)
    assert l1.max() <= 1e-6, l1.max()
    g.close()


@pytest.mark.parametrize("directed", [False, True])
def test_personalized_pagerank_chunked_hub_rows(directed):
    """Rows longer than 4096 slots are summed in chunks by separate
    wavefronts and merged in chunk order: two hubs of 5000 and 13000
    in-neighbours (2 and 4 chunks), one of them a seed, over an R-MAT graph."""
    V = 20000
    rng = np.random.default_rng(3)
    s, d = O.gen_rmat(O.stream_key(22, O.STREAM_RMAT), 14, 60000)
    s, d = s % V, d % V
    hub_src = [rng.choice(np.arange(2, V), 5000, replace=False).astype(np.uint32),
               rng.choice(np.arange(2, V), 13000, replace=False).astype(np.uint32)]
    s = np.concatenate([s, hub_src[0], hub_src[1]]).astype(np.uint32)
    d = np.concatenate([d, np.zeros(5000, np.uint32), np.ones(13000, np.uint32)]).astype(np.uint32)
    g, o = build_pair(V, directed, s, d)
    seeds = np.concatenate([[1], rng.choice(np.arange(2, V), 63, replace=False)]).astype(np.uint32)
    X, it, r = G.pagerank_personalized(g, seeds, 0.85, 1e-6, 100)
    Xo, ito, _ = o.pagerank_personalized(seeds, 0.85, 1e-6, 100)
    assert abs(it - ito) <= 1
    Xr, _, _ = o.pagerank_personalized(seeds, 0.85, 1e-6, 100, fixed_iters=it)
    l1 = np.abs(X.astype(np.float64) - Xr).sum(axis=0)
    assert l1.max() <= 1e-6, l1.max()
    g.close()


def test_personalized_matches_global_when_averaged():
    V = 64
    s, d = O.gen_rmat(O.stream_key(22, O.STREAM_RMAT), 6, 500)
    g, o = build_pair(V, False, s % V, d % V)
    X, it, _ = G.pagerank_personalized(g, np.arange(V), 0.85, 1e-30, 12)
    x, itg, _ = G.static_pagerank(g, 0.85, 1e-30, 12)
    assert it == itg == 12
    assert np.abs(X.mean(axis=1) - x).max() < 1e-6
    g.close()


# bc_bu_ratio: 0 = every level top-down, 1e9 = bottom-up wherever the
# in-adjacency allows (every level but the sources'), 1 = the default rule
@pytest.mark.parametrize("ratio", [0.0, 1.0, 1e9])
@pytest.mark.parametrize("nsrc", [2, 16, 33, 70])
def test_multi_source_bc(nsrc, ratio):
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(23, O.STREAM_RMAT), 12, 16 * V)
    for directed in (True, False):
        g, o = build_pair(V, directed, s, d, bc_bu_ratio=ratio)
        src = np.random.default_rng(nsrc).choice(V, nsrc, replace=False).astype(np.uint32)
        b = G.bc(g, src).astype(np.float64)
        bo = o.bc(src)
        err = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
        assert err.max() <= 1e-5, err.max()
        # the result does not depend on the direction chosen per level
        b2 = G.bc(g, src).astype(np.float64)
        assert np.array_equal(b, b2)
        g.close()


@pytest.mark.parametrize("ratio", [0.0, 1e9])
@pytest.mark.parametrize("directed", [True, False])
def test_multi_source_bc_hub_items_both_directions(directed, ratio):
    """Vertices with many work items in every role: a hub H reached at level
    2 through 700 level-1 vertices (a bottom-up candidate with 700 in-slots,
    three 256-slot items that add atomically), whose 1500 out-neighbours form
    level 3 (a top-down vertex with six items and a backward vertex whose
    partials are merged); sources at the hub's in-neighbours, the hub and a
    far vertex; a chain hangs off level 3 so the sweep runs past 8 levels."""
    n1, n3, chain = 700, 1500, 12
    H = 1 + n1
    V = H + 1 + n3 + chain + 5
    s = [0] * n1 + list(range(1, n1 + 1))
    d = list(range(1, n1 + 1)) + [H] * n1
    lvl3 = list(range(H + 1, H + 1 + n3))
    s += [H] * n3
    d += lvl3
    c0 = H + 1 + n3
    s += [lvl3[0]] + list(range(c0, c0 + chain - 1))
    d += list(range(c0, c0 + chain))
    g, o = build_pair(V, directed, np.array(s, np.uint32), np.array(d, np.uint32),
                      bc_bu_ratio=ratio)
    for src in ([0], [0, 5], [0, 5, H, c0 + 3, V - 1]):
        src = np.array(src + [0], np.uint32)  # at least two sources (the multi-source path)
        b = G.bc(g, src).astype(np.float64)
        bo = o.bc(src)
        err = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
        assert err.max() <= 1e-5, (src, err.max())
    g.close()


def test_multi_source_bc_repeated_source_and_path():
    n = 40
    g, o = build_pair(n, False, np.arange(n - 1), np.arange(1, n))
    src = np.array([0, 5, 5, 39, 20], np.uint32)
    b = G.bc(g, src).astype(np.float64)
    assert np.allclose(b, o.bc(src), rtol=1e-6, atol=1e-4)
    g.close()


def test_pagerank_warm_start_and_plain_sums():
    V = 1 << 13
    s, d = O.gen_rmat(O.stream_key(24, O.STREAM_RMAT), 13, 16 * V)
    g, o = build_pair(V, False, s, d)
    x0, it0, _ = G.pagerank(g, epsilon=1e-7)
    # restarting from the converged vector stops after the first update, whose
    # residual is already below eps
    again, ita, ra = G.pagerank(g, epsilon=1e-7, init=x0)
    assert ita <= 2 and ra < 1e-7
    assert np.abs(again - x0).sum() < 1e-6
    # after a perturbation a warm start need not take fewer iterations (the
    # uniform edges join isolated vertices into small bipartite pieces whose
    # error decays at rate alpha from any start), but it reaches the same point
    si, di = O.gen_uniform(O.stream_key(24, 1000), V, 2000)
    g.batchInsert(si, di)
    o.insert(si, di)
    cold, itc, _ = G.pagerank(g)
    warm, itw, _, st = G.pagerank(g, init=x0, stats=True)
    assert st["iters"] == itw
    # each against the oracle from the same start at the same iteration count
    xc, _, _ = o.pagerank(fixed_iters=itc)
    xw, _, _ = o.pagerank(fixed_iters=itw, x0=x0.astype(np.float64))
    assert np.abs(cold - xc).sum() <= 1e-6
    assert np.abs(warm - xw).sum() <= 1e-6
    # and a warm start stopped by the rule agrees with the oracle's own stop
    _, itwo, _ = o.pagerank(x0=x0.astype(np.float64))
    assert abs(itwo - itw) <= 1
    # plain FP32 sums (the uncompensated comparison path of the precision
    # workload) at equal iterations
    plain, itp, _ = G.pagerank(g, compensated=False)
    xr, _, _ = o.pagerank(fixed_iters=itp)
    assert np.abs(plain - xr).sum() <= 1e-6
    g.close()
