"""Pins the oracle's personalized PageRank (orc_pagerank_personalized):
closed form on a directed cycle, and linearity against the already pinned
global PageRank (the mean of the personalized iterates over all V seeds is the
global iterate at every update, since their start vectors and teleport terms
average to 1/V)."""
import numpy as np

from oracle import oracle as O


def test_ppr_directed_cycle_closed_form():
    n, a = 7, 0.85
    g = O.OracleGraph(n, True)
    g.bulk_build(np.arange(n), (np.arange(n) + 1) % n)
    x, it, r = g.pagerank_personalized([2], alpha=a, eps=1e-15, max_iter=5000)
    # fixed point: x[2+k] = (1-a) a^k / (1 - a^n)
    ref = np.array([(1 - a) * a ** ((v - 2) % n) / (1 - a ** n) for v in range(n)])
    assert np.abs(x[:, 0] - ref).max() < 1e-12


def test_ppr_mean_over_all_seeds_equals_global_pagerank():
    V = 40
    s, d = O.gen_rmat(O.stream_key(3, O.STREAM_RMAT), 6, 300)
    g = O.OracleGraph(V, True)
    g.bulk_build(s % V, d % V)
    for k in (1, 3, 10):
        X, it, _ = g.pagerank_personalized(np.arange(V), fixed_iters=k)
        xg, itg, _ = g.pagerank(fixed_iters=k)
        assert it == itg == k
        assert np.abs(X.mean(axis=1) - xg).max() < 1e-15


def test_ppr_stopping_rule():
    V = 30
    s, d = O.gen_rmat(O.stream_key(5, O.STREAM_RMAT), 5, 200)
    g = O.OracleGraph(V, False)
    g.bulk_build(s % V, d % V)
    X, it, r = g.pagerank_personalized([0, 5, 7], eps=1e-8, max_iter=1000)
    assert r < 1e-8
    _, _, r_prev = g.pagerank_personalized([0, 5, 7], fixed_iters=it - 1)
    assert r_prev >= 1e-8
