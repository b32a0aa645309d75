"""Sorted-prefix probes of the store against the CPU oracle.

A bulk-built adjacency (the first insert run into an empty vertex) is a
sorted prefix, cut into chunks of 16 slots that each keep the value of their
first slot; the update probe reads only the chunks a batch key can lie in.
These tests drive hubs through every way that prefix changes -- deletes of
chunk-first values, absent keys inside chunk ranges, duplicate keys, inserts
above and inside the prefix's range (a tail), compaction, merged groups, keys
dense enough that fence items are scanned instead -- and compare
sorted neighbour sets, degrees and applied counts with the oracle after every
step, at every block size.  Hub sizes cover one fence item (256..1023 prefix
slots), many fence items, and segments of at most 8, at most 64 and more than
64 keys."""
import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu

V = 1 << 17
HUBS = {0: 300, 1: 1000, 2: 5000, 3: 70000, 4: 2048}


def make(directed, block_slots):
    cfg = G.default_config(block_slots=block_slots)
    cfg.pool_blocks = 3 * V + 64 * 100000 // block_slots + 8192
    return G.GraphVine(V, directed=directed, config=cfg)


def assert_same(g, o, where):
    off, cols = g.export_csr()
    eo, ec = o.export_csr()
    assert np.array_equal(off, eo), f"offsets differ {where}"
    assert np.array_equal(cols, ec), f"neighbour sets differ {where}"
    assert np.array_equal(g.degrees(), o.degrees()), f"degrees differ {where}"


def hub_graph(rng):
    src, dst = [], []
    for h, d in HUBS.items():
        # hubs 0 and 4 stay below V/2, so runs above their largest neighbour fit
        hi = V // 2 if h in (0, 4) else V
        nb = rng.choice(np.arange(16, hi), d, replace=False)
        src.append(np.full(d, h, np.uint32))
        dst.append(nb.astype(np.uint32))
    # background: a sparse random graph over the other vertices
    n = 4 * V
    src.append(rng.integers(16, V, n, dtype=np.uint32))
    dst.append(rng.integers(16, V, n, dtype=np.uint32))
    return np.concatenate(src), np.concatenate(dst)


def neighbours(o, u):
    off, cols = o.export_csr()
    return cols[off[u]:off[u + 1]].astype(np.int64)


def delete_keys(o, rng, many):
    """Per hub: chunk-first values (every 16th sorted neighbour), the largest,
    random ones, absent values inside the hub's range and duplicates; `many`
    selects how many keys the big hubs get (<= 8, <= 64 or well above)."""
    src, dst = [], []
    for h in HUBS:
        nb = np.sort(neighbours(o, h))
        if len(nb) == 0:
            continue
        k = {0: 6, 1: 40, 2: 600}[many] if h in (2, 3) else 5
        pick = list(nb[::16][:k // 2]) + [nb[-1]]
        pick += list(rng.choice(nb, min(k, len(nb)), replace=False))
        present = set(nb.tolist())
        absent = [x for x in rng.integers(nb[0], nb[-1] + 1, 3 * k) if x not in present][:3]
        pick += absent + pick[:2]  # duplicates
        src.append(np.full(len(pick), h, np.uint32))
        dst.append(np.array(pick, np.uint32))
    s, d = np.concatenate(src), np.concatenate(dst)
    p = rng.permutation(len(s))
    return s[p], d[p]


def insert_keys(o, rng):
    """Hub 0 and 4: a run above their largest neighbour; hubs 1-3: values
    spread over their range; both go to a tail after the prefix."""
    src, dst = [], []
    for h in HUBS:
        nb = neighbours(o, h)
        top = int(nb.max()) if len(nb) else 16
        if h in (0, 4) and top + 100 < V:
            new = np.arange(top + 1, top + 1 + 70, dtype=np.int64)
        else:
            new = rng.integers(16, V, 200)
        src.append(np.full(len(new), h, np.uint32))
        dst.append(new.astype(np.uint32))
    s, d = np.concatenate(src), np.concatenate(dst)
    return s, d


@pytest.mark.parametrize("directed", [True, False])
@pytest.mark.parametrize("block_slots", [64, 32, 128])
def test_sorted_prefix_probe_sequence(block_slots, directed):
    rng = np.random.default_rng(11 + block_slots + (1 if directed else 0))
    g = make(directed, block_slots)
    o = O.OracleGraph(V, directed)
    s, d = hub_graph(rng)
    st = g.bulkBuild(s, d)
    assert st["applied"] == o.bulk_build(s, d)
    assert_same(g, o, "after bulk build")
    deleted = []
    for step, many in enumerate((0, 1, 2, 2, 1)):
        ds, dd = delete_keys(o, rng, many)
        st = g.batchDelete(ds, dd)
        assert st["applied"] == o.delete(ds, dd), f"delete applied, step {step}"
        assert_same(g, o, f"after delete step {step}")
        deleted.append((ds, dd))
        is_, id_ = insert_keys(o, rng)
        # some deleted values come back (into a tail)
        rs, rd = deleted[-1]
        is_ = np.concatenate([is_, rs[:20]])
        id_ = np.concatenate([id_, rd[:20]])
        st = g.batchInsert(is_, id_)
        assert st["applied"] == o.insert(is_, id_), f"insert applied, step {step}"
        assert_same(g, o, f"after insert step {step}")
        if step in (1, 3):
            g.compaction(0.0)
            assert_same(g, o, f"after compaction step {step}")
    # a merged group over the same hubs: deletes and inserts in one pipeline
    srcs, dsts, offs, kinds = [], [], [0], []
    shadow = O.OracleGraph(V, directed)
    off, cols = o.export_csr()
    ss = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    shadow.bulk_build(ss, cols)
    for b in range(6):
        if b % 2 == 0:
            bs, bd = delete_keys(shadow, rng, b % 3)
            kinds.append(G.GV_OP_DELETE)
            shadow.delete(bs, bd)
        else:
            bs, bd = insert_keys(shadow, rng)
            kinds.append(G.GV_OP_INSERT)
            shadow.insert(bs, bd)
        srcs.append(bs)
        dsts.append(bd)
        offs.append(offs[-1] + len(bs))
    g.applyBatches(np.concatenate(srcs), np.concatenate(dsts), np.array(offs, np.uint64), kinds,
                   compact_fraction=0.25, stats=False)
    assert_same(g, shadow, "after merged group")
    # and plain deletes again on the result
    ds, dd = delete_keys(shadow, rng, 2)
    st = g.batchDelete(ds, dd)
    assert st["applied"] == shadow.delete(ds, dd)
    assert_same(g, shadow, "after final delete")
    g.close()


def test_prefix_of_insert_runs_into_empty_vertices():
    """Vertices that start empty get their first insert run as a sorted
    prefix; later runs (above it or inside its range) form a tail."""
    rng = np.random.default_rng(5)
    g = make(True, 64)
    o = O.OracleGraph(V, True)
    for step in range(6):
        s, d = [], []
        for h in range(4):
            lo = 16 + step * 3000 if h % 2 == 0 else int(rng.integers(16, V - 4000))
            vals = np.unique(rng.integers(lo, lo + 3000, 600 + 300 * h))
            s.append(np.full(len(vals), h, np.uint32))
            d.append(vals.astype(np.uint32))
        s, d = np.concatenate(s), np.concatenate(d)
        st = g.batchInsert(s, d)
        assert st["applied"] == o.insert(s, d)
        assert_same(g, o, f"after insert run {step}")
        ds, dd = delete_keys(o, rng, step % 3)
        st = g.batchDelete(ds, dd)
        assert st["applied"] == o.delete(ds, dd)
        assert_same(g, o, f"after delete {step}")
    g.close()
