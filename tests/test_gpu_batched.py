"""Batched form (gv_batch_apply): a group of uneven insert/delete batches in
one pipeline must leave exactly the graph that applying the batches one after
another leaves (oracle), bit-exact sorted neighbour sets."""
import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make(V, directed, pool=None, **kw):
    cfg = G.default_config(**kw)
    cfg.pool_blocks = pool if pool is not None else 8 * V + 4096
    return G.GraphVine(V, directed=directed, config=cfg)


def assert_same(g, o, where=""):
    off, cols = g.export_csr()
    eo, ec = o.export_csr()
    assert np.array_equal(off, eo), f"offsets differ {where}"
    assert np.array_equal(cols, ec), f"neighbour sets differ {where}"
    assert g.num_edges() == o.num_edges()


def uneven_group(o, V, seed, nb, lo=1, hi=3000):
    """Random group: log-uniform sizes, 70% inserts / 30% deletes; deletes
    sampled from the oracle graph as it evolves; some inserts re-add edges
    deleted earlier in the group and some deletes hit edges inserted earlier."""
    rng = np.random.default_rng(seed)
    srcs, dsts, offs, kinds = [], [], [0], []
    recent = None
    for b in range(nb):
        n = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if rng.random() < 0.7:
            s, d = O.gen_uniform(O.stream_key(seed, 10_000 + b), V, n)
            if recent is not None and len(recent[0]):
                k = min(len(recent[0]), n // 3)
                s[:k], d[:k] = recent[0][:k], recent[1][:k]
            kinds.append(G.GV_OP_INSERT)
            o.insert(s, d)
        else:
            off, cols = o.export_csr()
            s, d = O.gen_sample(O.stream_key(seed, 20_000 + b), V, off, cols, n)
            kinds.append(G.GV_OP_DELETE)
            o.delete(s, d)
        recent = (s.copy(), d.copy())
        srcs.append(s)
        dsts.append(d)
        offs.append(offs[-1] + len(s))
    return np.concatenate(srcs), np.concatenate(dsts), np.array(offs, np.uint64), kinds


@pytest.mark.parametrize("directed", [False, True])
def test_group_equals_sequential(directed):
    V = 1 << 13
    s, d = O.gen_rmat(O.stream_key(4, O.STREAM_RMAT), 13, 16 * V)
    o = O.OracleGraph(V, directed)
    g = make(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    for grp in range(3):
        gs, gd, off, kinds = uneven_group(o, V, 100 + grp, nb=40)
        g.applyBatches(gs, gd, off, kinds, compact_fraction=0.25 if grp == 1 else -1.0)
        assert_same(g, o, f"group {grp}")
    g.compaction(0.0)
    assert_same(g, o, "after compaction")
    g.close()


def test_group_insert_delete_same_edge_orders():
    V = 50
    o = O.OracleGraph(V, True)
    g = make(V, True)
    base_s = np.array([1, 2, 3], np.uint32)
    base_d = np.array([2, 3, 4], np.uint32)
    g.bulkBuild(base_s, base_d)
    o.bulk_build(base_s, base_d)
    # (1,2): delete then insert -> present; (5,6): insert then delete -> absent;
    # (2,3): delete, insert, delete -> absent; (7,8): insert twice -> present
    batches = [([1, 2], [2, 3], G.GV_OP_DELETE), ([1, 5, 2, 7], [2, 6, 3, 8], G.GV_OP_INSERT),
               ([5, 2], [6, 3], G.GV_OP_DELETE), ([7], [8], G.GV_OP_INSERT), ([], [], G.GV_OP_DELETE)]
    s = np.concatenate([np.array(b[0], np.uint32) for b in batches])
    d = np.concatenate([np.array(b[1], np.uint32) for b in batches])
    off = np.cumsum([0] + [len(b[0]) for b in batches]).astype(np.uint64)
    for bs, bd, k in batches:
        (o.insert if k == G.GV_OP_INSERT else o.delete)(np.array(bs, np.uint32), np.array(bd, np.uint32))
    g.applyBatches(s, d, off, [b[2] for b in batches])
    assert_same(g, o)
    offs, out = g.neighbors([1, 2, 5, 7])
    assert out.tolist() == [2, 8]
    g.close()


def test_group_rejects_bad_ids_atomically():
    V = 100
    g = make(V, False)
    o = O.OracleGraph(V, False)
    s, d = O.gen_uniform(O.stream_key(7, 1), V, 300)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    with pytest.raises(G.GraphVineError) as e:
        g.applyBatches(np.array([1, 2, 3], np.uint32), np.array([4, 5, 100], np.uint32),
                       np.array([0, 2, 3], np.uint64), [G.GV_OP_INSERT, G.GV_OP_DELETE])
    assert e.value.status == G.GV_E_RANGE
    assert_same(g, o)
    g.close()


def test_group_above_single_sweep_limit_equals_sequential():
    """A group of more than 2^24 directed keys takes the classic radix path
    and sorts only the (source, destination) bits: equal edges must still
    come out in batch order, so the last operation on an edge wins.  Inserts
    re-add edges deleted by the batch before, deletes hit edges inserted by
    the batch before."""
    V = 1 << 16
    s, d = O.gen_rmat(O.stream_key(9, O.STREAM_RMAT), 16, 16 * V)
    o = O.OracleGraph(V, False)
    g = make(V, False, pool=64 * V)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    rng = np.random.default_rng(9)
    srcs, dsts, offs, kinds = [], [], [0], []
    prev = None
    n = 300_000
    for b in range(40):
        if b % 2 == 0:
            bs, bd = O.gen_uniform(O.stream_key(9, 30_000 + b), V, n)
            if prev is not None:
                k = len(prev[0]) // 4
                bs[:k], bd[:k] = prev[0][:k], prev[1][:k]
            kinds.append(G.GV_OP_INSERT)
            o.insert(bs, bd)
        else:
            eo, ec = o.export_csr()
            bs, bd = O.gen_sample(O.stream_key(9, 40_000 + b), V, eo, ec, n // 2)
            k = min(len(bs), n // 8)
            pick = rng.choice(len(prev[0]), k, replace=False)
            bs = np.concatenate([bs, prev[0][pick]])
            bd = np.concatenate([bd, prev[1][pick]])
            kinds.append(G.GV_OP_DELETE)
            o.delete(bs, bd)
        prev = (bs.copy(), bd.copy())
        srcs.append(bs)
        dsts.append(bd)
        offs.append(offs[-1] + len(bs))
    assert 2 * offs[-1] > (1 << 24)
    g.applyBatches(np.concatenate(srcs), np.concatenate(dsts), np.array(offs, np.uint64), kinds,
                   compact_fraction=0.25)
    assert_same(g, o, "group above the single-sweep limit")
    g.close()
