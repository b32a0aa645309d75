"""GPU parity of the per-batch triangle-count update (gv_triangle_count_update,
the exact replacement of dynamic_tc, tc.cu:29-64).

Oracle: after every batch the updated counts must equal the oracle's full
recount of the updated graph (per vertex and total, exact uint64), and the
device's own full recount (gv_triangle_count).  The oracle's count is pinned
by closed forms and brute force in the oracle tests.
"""
import itertools
import math

import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make(V, pool=None, **kw):
    cfg = G.default_config(**kw)
    cfg.pool_blocks = pool if pool is not None else 4 * V + 4096
    return G.GraphVine(V, directed=False, config=cfg)


def pair(V, s, d, pool=None, **kw):
    g = make(V, pool, **kw)
    o = O.OracleGraph(V, False)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    return g, o


def step(g, o, pv, tot, s, d, delete):
    tot, _ = G.dynamic_tc(g, s, d, pv, tot, delete=delete)
    (o.delete if delete else o.insert)(s, d)
    pvo, toto = o.triangles()
    assert tot == toto
    assert np.array_equal(pv, pvo)
    return tot


def check_device(g, pv, tot):
    pvd, totd = G.static_tc(g)
    assert totd == tot and np.array_equal(pvd, pv)


def test_one_batch_with_all_three_edges():
    # every edge of the triangle changes in one batch: counted once
    g, o = pair(4, [0], [3])
    pv, tot = G.static_tc(g)
    assert tot == 0
    tot = step(g, o, pv, tot, [0, 1, 2, 1], [1, 2, 0, 0], False)  # (1,0) repeats (0,1)
    assert tot == 1 and pv[0] == 1
    tot = step(g, o, pv, tot, [2, 0, 1], [0, 1, 2], True)
    assert tot == 0 and not pv.any()


@pytest.mark.parametrize("n", [6, 40])
def test_complete_graph_rebuilt_and_torn_down(n):
    e = list(itertools.combinations(range(n), 2))
    rng = np.random.default_rng(n)
    rng.shuffle(e)
    half = len(e) // 2
    a = np.array(e, np.uint32)
    g, o = pair(n, a[:half, 0], a[:half, 1])
    pv, tot = G.static_tc(g)
    # the other half, plus repeats, reversed copies, self-loops and present edges
    ins_s = np.concatenate([a[half:, 1], a[half:, 0], a[:5, 0], np.arange(3)])
    ins_d = np.concatenate([a[half:, 0], a[half:, 1], a[:5, 1], np.arange(3)])
    tot = step(g, o, pv, tot, ins_s, ins_d, False)
    assert tot == math.comb(n, 3)
    assert pv.tolist() == [math.comb(n - 1 - i, 2) for i in range(n)]
    check_device(g, pv, tot)
    # deletes of absent edges change nothing
    tot = step(g, o, pv, tot, [0, 1], [0, 1], True)
    assert tot == math.comb(n, 3)
    tot = step(g, o, pv, tot, a[: 2 * half // 3, 0], a[: 2 * half // 3, 1], True)
    check_device(g, pv, tot)
    tot = step(g, o, pv, tot, a[:, 1], a[:, 0], True)
    assert tot == 0 and not pv.any()


def test_config1_stream_of_batches():
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 16 * V)
    g, o = pair(V, s, d, pool=8 * V)
    pv, tot = G.static_tc(g)
    for b in range(3):
        si, di = O.gen_uniform(O.stream_key(1, 1000 + b), V, 4096)
        # re-insert some present edges too
        off, cols = o.export_csr()
        sp, dp = O.gen_sample(O.stream_key(1, 3000 + b), V, off, cols, 512)
        tot = step(g, o, pv, tot, np.concatenate([si, sp]), np.concatenate([di, dp]), False)
        off, cols = o.export_csr()
        sd, dd = O.gen_sample(O.stream_key(1, 2000 + b), V, off, cols, 4096)
        # and delete some absent edges
        sa, da = O.gen_uniform(O.stream_key(1, 4000 + b), V, 256)
        tot = step(g, o, pv, tot, np.concatenate([sd, sa]), np.concatenate([dd, da]), True)
    check_device(g, pv, tot)
    g.compaction(0.0)
    check_device(g, pv, tot)


def test_hubs_take_the_block_path():
    # hub 0 joined to almost every vertex (more than one 16384-slot part),
    # hub 1 of a few thousand slots (one part), and batches with hundreds of
    # keys on each hub (many key chunks) and hub-hub and hub-leaf edges
    V = 40000
    rng = np.random.default_rng(11)
    leaves0 = np.arange(2, V, dtype=np.uint32)
    keep0 = leaves0[rng.random(len(leaves0)) < 0.9]
    leaves1 = rng.choice(np.arange(2, V, dtype=np.uint32), 3000, replace=False)
    rs = rng.integers(2, V, 60000).astype(np.uint32)
    rd = rng.integers(2, V, 60000).astype(np.uint32)
    s = np.concatenate([np.zeros(len(keep0), np.uint32), np.ones(len(leaves1), np.uint32), rs])
    d = np.concatenate([keep0, leaves1, rd])
    g, o = pair(V, s, d, pool=8 * V)
    pv, tot = G.static_tc(g)
    miss0 = np.setdiff1d(leaves0, keep0)[:1500]
    ins_s = np.concatenate([np.zeros(len(miss0), np.uint32), [0, 1],
                            rng.integers(2, V, 20000).astype(np.uint32)])
    ins_d = np.concatenate([miss0, [1, 0], rng.integers(2, V, 20000).astype(np.uint32)])
    tot = step(g, o, pv, tot, ins_s, ins_d, False)
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(11, 2000), V, off, cols, 20000)
    del_s = np.concatenate([sd, np.zeros(800, np.uint32), np.ones(400, np.uint32), [0]])
    del_d = np.concatenate([dd, keep0[:800], leaves1[:400], [1]])
    tot = step(g, o, pv, tot, del_s, del_d, True)
    check_device(g, pv, tot)


def test_total_only_and_errors():
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(2, O.STREAM_RMAT), 12, 16 * V)
    g, o = pair(V, s, d)
    _, tot = G.static_tc(g, per_vertex=False)
    si, di = O.gen_uniform(O.stream_key(2, 1000), V, 2048)
    tot, st = G.dynamic_tc(g, si, di, None, tot, stats=True)
    o.insert(si, di)
    assert tot == o.triangles()[1]
    assert st["requested"] == 2048 and st["applied"] > 0
    # a bad ID: nothing applied, counts unchanged
    with pytest.raises(G.GraphVineError) as e:
        G.dynamic_tc(g, [1, V], [2, 3], None, tot)
    assert e.value.status == G.GV_E_RANGE
    pvd, totd = G.static_tc(g)
    assert totd == tot and np.array_equal(pvd, o.triangles()[0])
    gd = G.GraphVine(10, directed=True)
    with pytest.raises(G.GraphVineError) as e:
        G.dynamic_tc(gd, [0], [1], None, 0)
    assert e.value.status == G.GV_E_UNSUPPORTED
