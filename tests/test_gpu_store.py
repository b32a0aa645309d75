"""GPU parity of the dynamic graph store against the CPU oracle.

Every comparison is bit-exact: sorted neighbour sets (via gv_export_csr),
degrees and per-batch applied counts.  Workloads are seeded and identical on
both sides (host-generated batches fed to both)."""
import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make(V, directed, pool=None, **kw):
    cfg = G.default_config(**kw)
    cfg.pool_blocks = pool if pool is not None else 4 * V + 4096
    return G.GraphVine(V, directed=directed, config=cfg)


def assert_same(g, o, where=""):
    off, cols = g.export_csr()
    eo, ec = o.export_csr()
    assert np.array_equal(off, eo), f"offsets differ {where}"
    assert np.array_equal(cols, ec), f"neighbour sets differ {where}"
    assert np.array_equal(g.degrees(), o.degrees()), f"degrees differ {where}"
    assert g.num_edges() == o.num_edges()


def config1_stream(seed=1):
    """Plain small config: directed R-MAT scale 14, edge factor 16, then 10
    insert batches of 4096 uniform edges and 10 delete batches of 4096 sampled
    edges (each followed by compaction)."""
    V = 1 << 14
    s, d = O.gen_rmat(O.stream_key(seed, O.STREAM_RMAT), 14, 16 * V)
    return V, s, d


@pytest.mark.parametrize("block_slots", [64, 32, 128])
def test_config1_update_sequence_bit_exact(block_slots):
    V, s, d = config1_stream()
    o = O.OracleGraph(V, True)
    g = make(V, True, pool=3 * V + 16 * V * 2 // block_slots + 4096, block_slots=block_slots)
    st = g.bulkBuild(s, d)
    assert st["applied"] == o.bulk_build(s, d)
    assert_same(g, o, "after bulk build")
    for b in range(10):
        si, di = O.gen_uniform(O.stream_key(1, 1000 + b), V, 4096)
        st = g.batchInsert(si, di)
        assert st["applied"] == o.insert(si, di)
        assert_same(g, o, f"after insert {b}")
    for b in range(10):
        off, cols = o.export_csr()
        sd, dd = O.gen_sample(O.stream_key(1, 2000 + b), V, off, cols, 4096)
        st = g.batchDelete(sd, dd)
        assert st["applied"] == o.delete(sd, dd)
        g.compaction(0.0)
        assert_same(g, o, f"after delete {b}")
    free, total = g.pool_stats()
    assert free <= total
    g.close()


def test_undirected_updates_bit_exact():
    V = 5000
    s, d = O.gen_rmat(O.stream_key(2, O.STREAM_RMAT), 13, 40000)
    s, d = s % V, d % V
    o = O.OracleGraph(V, False)
    g = make(V, False, pool=8 * V)
    assert g.bulkBuild(s, d)["applied"] == o.bulk_build(s, d)
    assert_same(g, o)
    for b in range(6):
        si, di = O.gen_uniform(O.stream_key(2, 3000 + b), V, 3000)
        assert g.batchInsert(si, di)["applied"] == o.insert(si, di)
        off, cols = o.export_csr()
        sd, dd = O.gen_sample(O.stream_key(2, 4000 + b), V, off, cols, 3000)
        assert g.batchDelete(sd, dd)["applied"] == o.delete(sd, dd)
        if b % 2:
            g.compaction(0.25)
        assert_same(g, o, f"batch {b}")
    g.close()


def test_duplicates_selfloops_and_noops():
    g = make(10, True)
    o = O.OracleGraph(10, True)
    s = np.array([1, 1, 1, 2, 3, 3], np.uint32)
    d = np.array([2, 2, 1, 3, 3, 4], np.uint32)
    assert g.bulkBuild(s, d)["applied"] == o.bulk_build(s, d) == 3
    assert g.batchInsert([1], [2])["applied"] == 0  # existing edge
    assert g.batchDelete([5], [6])["applied"] == 0  # missing edge
    assert g.batchInsert([], [])["applied"] == 0  # empty batch
    assert g.batchInsert([4, 4], [4, 4])["applied"] == 0  # self-loops only
    assert_same(g, o)
    # delete then re-insert the same edge while its tombstone is still there
    assert g.batchDelete([1], [2])["applied"] == 1
    assert g.batchInsert([1], [2])["applied"] == 1
    o.delete([1], [2])
    o.insert([1], [2])
    assert_same(g, o)
    g.close()


def test_out_of_range_rejects_whole_batch():
    g = make(100, False)
    o = O.OracleGraph(100, False)
    s, d = O.gen_uniform(O.stream_key(3, 1), 100, 500)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    with pytest.raises(G.GraphVineError) as e:
        g.batchInsert([1, 2, 3], [4, 100, 5])
    assert e.value.status == G.GV_E_RANGE
    with pytest.raises(G.GraphVineError) as e:
        g.batchDelete([int(s[0]), 7], [int(d[0]), 1000])
    assert e.value.status == G.GV_E_RANGE
    assert_same(g, o, "after rejected batches")
    assert g.batchInsert([1], [4])["applied"] == o.insert([1], [4])
    assert_same(g, o)
    g.close()


def test_pool_exhaustion_rejects_and_recovers():
    V = 1000
    g = make(V, True, pool=V + 10, block_slots=32)  # V+10 blocks per store
    o = O.OracleGraph(V, True)
    s = np.arange(V, dtype=np.uint32)
    d = (s + 1) % V
    assert g.bulkBuild(s, d)["applied"] == o.bulk_build(s, d)  # V blocks per store
    free, total = g.pool_stats()
    assert total == 2 * (V + 10) and free == 20
    # vertex 0 grows to 999 slots = 32 blocks: 31 more > 10 free -> rejected
    big_d = np.arange(2, V, dtype=np.uint32)
    with pytest.raises(G.GraphVineError) as e:
        g.batchInsert(np.zeros(V - 2, np.uint32), big_d)
    assert e.value.status == G.GV_E_POOL
    assert g.pool_stats()[0] == free
    assert_same(g, o, "after rejected batch")
    assert g.batchInsert([5], [9])["applied"] == o.insert([5], [9])
    assert_same(g, o)
    g.close()


@pytest.mark.parametrize("directed", [True, False])
def test_failed_sort_is_reported_and_nothing_applied(directed):
    """The staging sort's bounded look-back wait, forced to give up at once
    from the host (every tile after the first), must fail the call with
    GV_E_HIP and leave the graph as it was -- for inserts, deletes and
    groups, and in both stores of a directed graph.  Every kernel exits
    normally; the wait bound is restored afterwards."""
    V = 1 << 12
    g = make(V, directed)
    o = O.OracleGraph(V, directed)
    s, d = O.gen_rmat(O.stream_key(21, O.STREAM_RMAT), 12, 8 * V)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    si, di = O.gen_uniform(O.stream_key(21, 1), V, 12000)  # > 2 sort tiles of keys
    off, cols = o.export_csr()
    sd, dd = O.gen_sample(O.stream_key(21, 2), V, off, cols, 12000)
    L = G.lib()
    try:
        assert L.gv_test_set_sort_spin_limit(0) == 0
        for fn in (g.batchInsert, g.batchDelete):
            with pytest.raises(G.GraphVineError) as e:
                fn(si, di) if fn == g.batchInsert else fn(sd, dd)
            assert e.value.status == G.GV_E_HIP
            assert "sort" in str(e.value)
        with pytest.raises(G.GraphVineError) as e:
            g.applyBatches(np.concatenate([sd, si]), np.concatenate([dd, di]),
                           [0, len(sd), len(sd) + len(si)], [G.GV_OP_DELETE, G.GV_OP_INSERT],
                           compact_fraction=0.25)
        assert e.value.status == G.GV_E_HIP
    finally:
        L.gv_test_set_sort_spin_limit(1 << 26)
    assert_same(g, o, "after failed sorts")
    assert g.batchDelete(sd, dd)["applied"] == o.delete(sd, dd)
    assert g.batchInsert(si, di)["applied"] == o.insert(si, di)
    assert_same(g, o, "after the same batches with the sort restored")
    g.close()


def test_delete_all_then_compact_returns_every_block():
    V = 3000
    g = make(V, False, pool=6 * V)
    s, d = O.gen_rmat(O.stream_key(4, O.STREAM_RMAT), 12, 30000)
    s, d = s % V, d % V
    g.bulkBuild(s, d)
    off, cols = g.export_csr()
    src = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    st = g.batchDelete(src, cols)
    assert st["applied"] == len(cols)
    assert g.num_edges() == 0
    n = g.compaction(0.0)
    assert n == int((np.diff(off) > 0).sum())
    free, total = g.pool_stats()
    assert free == total
    off2, cols2 = g.export_csr()
    assert off2[-1] == 0
    g.close()


def test_hub_beyond_reference_degree_limit():
    V = 600_001
    g = make(V, True, pool=V + 20000)
    o = O.OracleGraph(V, True)
    d = np.arange(1, V, dtype=np.uint32)
    s = np.zeros(V - 1, np.uint32)
    assert g.bulkBuild(s, d)["applied"] == V - 1
    o.bulk_build(s, d)
    rng = np.random.default_rng(0)
    dd = rng.choice(d, 50000, replace=False).astype(np.uint32)
    assert g.batchDelete(np.zeros(50000, np.uint32), dd)["applied"] == 50000
    o.delete(np.zeros(50000, np.uint32), dd)
    g.compaction(0.05)
    assert_same(g, o)
    offs, out = g.neighbors([0])
    assert offs[-1] == V - 1 - 50000
    g.close()


@pytest.mark.parametrize("block_slots", [32, 64, 128])
def test_hub_batch_segments_of_many_keys(block_slots):
    # one source with 65..4000 keys in a batch: its slots are searched through
    # 64 fence keys, including keys past the last fence and absent keys
    V = 20000
    rng = np.random.default_rng(block_slots)
    nb = rng.choice(np.arange(1, V, dtype=np.uint32), 12000, replace=False)
    g = make(V, True, pool=8 * V, block_slots=block_slots)
    o = O.OracleGraph(V, True)
    g.bulkBuild(np.zeros(len(nb), np.uint32), nb)
    o.bulk_build(np.zeros(len(nb), np.uint32), nb)
    live = np.sort(nb)
    for k in (65, 127, 1000, 4000):
        pick = rng.choice(live, k, replace=False)
        pick = np.concatenate([pick, live[-3:], rng.integers(1, V, 50).astype(np.uint32)])
        src = np.zeros(len(pick), np.uint32)
        assert g.batchDelete(src, pick)["applied"] == o.delete(src, pick)
        assert_same(g, o, f"delete of {k} keys")
        live = o.export_csr()[1]
        back = rng.choice(pick, len(pick) // 2, replace=False)
        src = np.zeros(len(back), np.uint32)
        assert g.batchInsert(src, back)["applied"] == o.insert(src, back)
        assert_same(g, o, f"re-insert after {k}")
        live = o.export_csr()[1]


def test_neighbors_query_matches_oracle():
    V = 4000
    g = make(V, False, pool=8 * V)
    o = O.OracleGraph(V, False)
    s, d = O.gen_rmat(O.stream_key(5, O.STREAM_RMAT), 12, 40000)
    s, d = s % V, d % V
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    q = np.random.default_rng(1).integers(0, V, 1000).astype(np.uint32)
    offs, out = g.neighbors(q)
    eo, ec = o.export_csr()
    for i, v in enumerate(q):
        got = out[offs[i]:offs[i + 1]]
        assert np.array_equal(got, ec[eo[v]:eo[v + 1]])
    g.close()


def test_neighbors_and_export_over_hub_items_with_tombstones():
    # hubs span many 1024-slot emission items; tombstones (no compaction)
    # make their items take output ranges from the per-query cursor; a
    # vertex is queried more than once
    V = 50000
    rng = np.random.default_rng(7)
    g = make(V, True, pool=16 * V)
    o = O.OracleGraph(V, True)
    h0 = rng.choice(np.arange(1, V, dtype=np.uint32), 30000, replace=False)
    h1 = rng.choice(np.arange(2, V, dtype=np.uint32), 3000, replace=False)
    rs = rng.integers(2, V, 60000).astype(np.uint32)
    rd = rng.integers(0, V, 60000).astype(np.uint32)
    s = np.concatenate([np.zeros(len(h0), np.uint32), np.ones(len(h1), np.uint32), rs])
    d = np.concatenate([h0, h1, rd])
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    d0 = rng.choice(h0, 5000, replace=False)
    d1 = rng.choice(h1, 500, replace=False)
    ds = np.concatenate([np.zeros(len(d0), np.uint32), np.ones(len(d1), np.uint32)])
    dd = np.concatenate([d0, d1])
    assert g.batchDelete(ds, dd)["applied"] == o.delete(ds, dd)
    # more inserts after the tombstones: items past a sorted prefix
    is_ = np.concatenate([np.zeros(2000, np.uint32), rng.integers(2, V, 5000).astype(np.uint32)])
    id_ = rng.integers(0, V, len(is_)).astype(np.uint32)
    assert g.batchInsert(is_, id_)["applied"] == o.insert(is_, id_)
    assert_same(g, o, "export with tombstones")
    q = np.concatenate([[0, 1, 0], rng.integers(0, V, 500), [1]]).astype(np.uint32)
    offs, out = g.neighbors(q)
    eo, ec = o.export_csr()
    for i, v in enumerate(q):
        assert np.array_equal(out[offs[i]:offs[i + 1]], ec[eo[v]:eo[v + 1]]), f"query {i} vertex {v}"
    g.close()


def test_device_generators_match_oracle():
    from gv_amd.hip import DeviceArray
    g = make(1 << 14, True)
    key = O.stream_key(1, O.STREAM_RMAT)
    ds, dd = G.gen_rmat(g, key, 14, 100000)
    hs, hd = O.gen_rmat(key, 14, 100000)
    assert np.array_equal(ds.numpy(), hs) and np.array_equal(dd.numpy(), hd)
    key = O.stream_key(1, O.STREAM_INSERT)
    ds, dd = G.gen_uniform(g, key, 12345, 50000)
    hs, hd = O.gen_uniform(key, 12345, 50000)
    assert np.array_equal(ds.numpy(), hs) and np.array_equal(dd.numpy(), hd)
    o = O.OracleGraph(1 << 14, True)
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 14, 50000)
    o.bulk_build(s, d)
    off, cols = o.export_csr()
    doff, dcols = DeviceArray.from_numpy(off), DeviceArray.from_numpy(cols)
    key = O.stream_key(1, O.STREAM_DELETE)
    ds, dd = G.gen_sample(g, key, 1 << 14, doff, dcols, 30000)
    hs, hd = O.gen_sample(key, 1 << 14, off, cols, 30000)
    assert np.array_equal(ds.numpy(), hs) and np.array_equal(dd.numpy(), hd)
    g.close()


def test_device_resident_batches_and_device_export():
    V = 1 << 12
    g = make(V, False, pool=16 * V)
    o = O.OracleGraph(V, False)
    ds, dd = G.gen_rmat(g, O.stream_key(9, O.STREAM_RMAT), 12, 50000)
    g.bulkBuild(ds, dd)
    o.bulk_build(ds.numpy(), dd.numpy())
    assert_same(g, o)
    off, cols, E = G.export_csr_device(g)
    assert E == o.num_edges()
    eo, ec = o.export_csr()
    assert np.array_equal(off.numpy(), eo) and np.array_equal(cols.numpy(E), ec)
    g.close()


def test_graph_from_matrix_market_file(tmp_path):
    """A symmetric MatrixMarket file (one triangle stored) loaded into an
    undirected graph gives the oracle's symmetric graph of the same pairs."""
    V = 2000
    s, d = O.gen_rmat(O.stream_key(81, O.STREAM_RMAT), 11, 20000)
    s, d = s % V, d % V
    lo, hi = np.minimum(s, d), np.maximum(s, d)
    keep = lo != hi
    lines = ["%%MatrixMarket matrix coordinate pattern symmetric", f"{V} {V} {int(keep.sum())}"]
    lines += [f"{h + 1} {l + 1}" for l, h in zip(lo[keep], hi[keep])]
    p = tmp_path / "g.mtx"
    p.write_text("\n".join(lines) + "\n")
    Vr, ms, md, sym = G.read_matrix_market(str(p))
    assert Vr == V and sym
    g = make(V, False)
    g.bulkBuild(ms, md)
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    assert_same(g, o)
    g.close()


def test_pool_capped_to_32bit_slot_indices():
    """A pool asked for more blocks than 32-bit slot indices address is capped
    at 2^32 / B - 1 blocks; blocks are taken from the top of the free stack,
    so the first inserts land on the highest slot indices, which must still
    read back exactly."""
    V = 1024
    cfg = G.default_config(block_slots=128)
    cfg.pool_blocks = 1 << 30
    g = G.GraphVine(V, directed=False, config=cfg)
    free, total = g.pool_stats()
    assert total == (1 << 32) // 128 - 1 and free == total
    o = O.OracleGraph(V, False)
    s, d = O.gen_uniform(O.stream_key(41, 1), V, 20000)
    g.batchInsert(s, d)
    o.insert(s, d)
    g.batchDelete(s[:5000], d[:5000])
    o.delete(s[:5000], d[:5000])
    g.compaction(0.0)
    assert_same(g, o, "top of the slot index range")
    g.close()
