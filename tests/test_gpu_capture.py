"""Captured update pipelines (gv_capture_*), the device Chung-Lu generator and
live tuning changes.  Every replay of a captured pipeline is checked against
the oracle applying the same batches in order, with eager calls and
synchronisations between replays."""
import json
import os

import numpy as np
import pytest

import gv_amd as G
from gv_amd.hip import DeviceArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
GOLD = np.uint64(0x9E3779B97F4A7C15)


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


def digest(s, d):
    k = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
    return int(np.bitwise_xor.reduce(k * GOLD))


def test_chung_lu_device_matches_fixture_and_oracle():
    g = make(1 << 10, False)
    with open(os.path.join(GOLDEN, "generators.json")) as f:
        gold = json.load(f)
    for case in gold["chung_lu"]:
        key = O.stream_key(case["seed"], O.STREAM_CHUNG_LU)
        ds, dd = G.gen_chung_lu(g, key, case["V"], case["n"], case["gamma"], case["avg_deg"],
                                case["max_deg"])
        s, d = ds.numpy(), dd.numpy()
        assert s[:8].tolist() == case["src_head"] and d[:8].tolist() == case["dst_head"]
        assert digest(s, d) == case["xor_digest"]
    V, n = 300_000, 1_000_000
    key = O.stream_key(9, O.STREAM_CHUNG_LU)
    ds, dd = G.gen_chung_lu(g, key, V, n, 2.1, 24.0, 200_000.0)
    hs, hd = O.gen_chung_lu(key, V, 2.1, 24.0, 200_000.0, n)
    assert np.array_equal(ds.numpy(), hs) and np.array_equal(dd.numpy(), hd)
    g.close()


SIZES = [500, 1500, 200, 3000, 800, 1000]
KINDS = [G.GV_OP_INSERT, G.GV_OP_DELETE, G.GV_OP_INSERT, G.GV_OP_INSERT, G.GV_OP_DELETE,
         G.GV_OP_INSERT]


def fill_group(o, V, seed):
    """Batches of the fixed shape (SIZES, KINDS): inserts uniform, deletes
    sampled from the oracle graph as it evolves; applied to the oracle in
    order."""
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


@pytest.mark.parametrize("directed", [False, True])
def test_captured_group_replays_with_refilled_inputs(directed):
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(31, O.STREAM_RMAT), 12, 16 * V)
    g = make(V, directed)
    o = O.OracleGraph(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    off = np.concatenate([[0], np.cumsum(SIZES)]).astype(np.uint64)
    n = int(off[-1])
    dsrc, ddst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    g.reserve(n, len(SIZES))
    with g.capture() as cap:
        g.applyBatches(dsrc, ddst, off, KINDS, compact_fraction=0.25, stats=False)
    c = cap.result
    for rep in range(3):
        hs, hd = fill_group(o, V, 500 + rep)
        g.sync()
        dsrc.upload(hs)
        ddst.upload(hd)
        c.launch()
        g.sync()
        assert_same(g, o, f"replay {rep}")
        # eager work between replays
        es, ed = O.gen_uniform(O.stream_key(600 + rep, 1), V, 700)
        g.batchInsert(es, ed)
        o.insert(es, ed)
        assert_same(g, o, f"eager after replay {rep}")
    c.close()
    g.close()


QCAP, QB = 6000, 8
REDEL = 2  # a delete batch drawn from the previous batch's pairs (test-only kind)


def fill_queue(o, V, seed, plan):
    """Host images of a device queue of capacity (QCAP pairs, QB batches)
    holding the batches of plan [(size, kind)], applied to the oracle in
    order.  Entries past the last batch hold out-of-range IDs (they must be
    ignored) and the unused trailing batches are empty."""
    srcs, dsts, kinds = [], [], []
    prev = (np.zeros(0, np.uint32), np.zeros(0, np.uint32))
    for b, (n, k) in enumerate(plan):
        if k == G.GV_OP_INSERT:
            s, d = O.gen_uniform(O.stream_key(seed, 30_000 + b), V, n)
            o.insert(s, d)
        elif k == G.GV_OP_DELETE:
            off, cols = o.export_csr()
            s, d = O.gen_sample(O.stream_key(seed, 40_000 + b), V, off, cols, n)
            o.delete(s, d)
        else:
            s, d = prev[0][:n].copy(), prev[1][:n].copy()
            o.delete(s, d)
            k = G.GV_OP_DELETE
        srcs.append(s)
        dsts.append(d)
        kinds.append(k)
        prev = (s, d)
    used = sum(len(x) for x in srcs)
    assert used <= QCAP and len(plan) <= QB
    hs = np.full(QCAP, 0xFFFFFFFF, np.uint32)
    hd = np.full(QCAP, 0xFFFFFFFF, np.uint32)
    if used:
        hs[:used] = np.concatenate(srcs)
        hd[:used] = np.concatenate(dsts)
    off = np.full(QB + 1, used, np.uint64)
    off[: len(plan) + 1] = np.concatenate([[0], np.cumsum([n for n, _ in plan])])
    kin = np.full(QB, G.GV_OP_INSERT, np.int32)
    kin[: len(plan)] = kinds
    return hs, hd, off, kin


I_, D_ = G.GV_OP_INSERT, G.GV_OP_DELETE
QUEUE_PLANS = [
    [(1500, I_), (800, D_), (0, I_), (1200, I_), (600, REDEL), (900, D_), (400, I_), (500, D_)],
    [(2000, I_), (700, REDEL), (300, D_)],
    [(400, D_)],
    [],
    [(3000, I_), (3000, REDEL)],
]


@pytest.mark.parametrize("directed", [False, True])
def test_device_queue_group_capture_replays_uneven_batches(directed):
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(36, O.STREAM_RMAT), 12, 16 * V)
    g = make(V, directed)
    o = O.OracleGraph(V, directed)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    qs, qd = DeviceArray(QCAP, np.uint32), DeviceArray(QCAP, np.uint32)
    qo, qk = DeviceArray(QB + 1, np.uint64), DeviceArray(QB, np.int32)
    g.reserve(QCAP, QB)
    # an eager call first
    hs, hd, off, kin = fill_queue(o, V, 800, QUEUE_PLANS[0])
    for arr, h in ((qs, hs), (qd, hd), (qo, off), (qk, kin)):
        arr.upload(h)
    g.applyBatchesDevice(qs, qd, qo, qk, compact_fraction=0.25)
    g.sync()
    assert_same(g, o, "eager device queue")
    with g.capture() as cap:
        g.applyBatchesDevice(qs, qd, qo, qk, compact_fraction=0.25)
    c = cap.result
    for rep, plan in enumerate(QUEUE_PLANS):
        hs, hd, off, kin = fill_queue(o, V, 810 + rep, plan)
        g.sync()
        for arr, h in ((qs, hs), (qd, hd), (qo, off), (qk, kin)):
            arr.upload(h)
        if rep == 1:
            st = c.launch(stats=True)
            assert st["live_edges"] == o.num_edges()
        else:
            c.launch()
        g.sync()
        assert_same(g, o, f"replay {rep} ({len(plan)} batches)")
        es, ed = O.gen_uniform(O.stream_key(820 + rep, 1), V, 500)
        g.batchInsert(es, ed)
        o.insert(es, ed)
        assert_same(g, o, f"eager after replay {rep}")
    c.close()
    g.close()


def test_device_queue_group_errors_leave_graph_unchanged():
    V = 1 << 11
    s, d = O.gen_rmat(O.stream_key(37, O.STREAM_RMAT), 11, 16 * V)
    g = make(V, False)
    o = O.OracleGraph(V, False)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    qs, qd = DeviceArray(QCAP, np.uint32), DeviceArray(QCAP, np.uint32)
    qo, qk = DeviceArray(QB + 1, np.uint64), DeviceArray(QB, np.int32)
    g.reserve(QCAP, QB)
    hs, _, _, _ = fill_queue(O.OracleGraph(V, False), V, 830, [(1000, I_)])
    hd = hs[::-1].copy()
    hd[:1000] = (hs[:1000] + 1) % V
    good_off = np.full(QB + 1, 1000, np.uint64)
    good_off[0] = 0
    good_kin = np.zeros(QB, np.int32)
    cases = [
        ("decreasing offsets", hs, hd, np.array([0, 600, 300] + [1000] * (QB - 2), np.uint64),
         good_kin, G.GV_E_INVALID),
        ("first offset not 0", hs, hd, np.array([5] + [1000] * QB, np.uint64), good_kin,
         G.GV_E_INVALID),
        ("past capacity", hs, hd, np.array([0] + [QCAP + 1] * QB, np.uint64), good_kin,
         G.GV_E_INVALID),
        ("bad kind", hs, hd, good_off, np.array([7] + [0] * (QB - 1), np.int32),
         G.GV_E_INVALID),
        ("vertex out of range", np.where(np.arange(QCAP) == 10, V + 3, hs).astype(np.uint32),
         hd, good_off, good_kin, G.GV_E_RANGE),
    ]
    for name, cs, cd, co, ck, code in cases:
        for arr, h in ((qs, cs), (qd, cd), (qo, co), (qk, ck)):
            arr.upload(h)
        g.applyBatchesDevice(qs, qd, qo, qk, compact_fraction=0.25)
        with pytest.raises(G.GraphVineError) as ei:
            g.sync()
        assert ei.value.status == code, name
        assert_same(g, o, name)
    # the handle applies a valid queue afterwards
    for arr, h in ((qs, hs), (qd, hd), (qo, good_off), (qk, good_kin)):
        arr.upload(h)
    g.applyBatchesDevice(qs, qd, qo, qk)
    g.sync()
    o.insert(hs[:1000], hd[:1000])
    assert_same(g, o, "valid queue after errors")
    g.close()


def test_captured_plain_sequence_with_stats():
    V = 1 << 11
    s, d = O.gen_rmat(O.stream_key(32, O.STREAM_RMAT), 11, 16 * V)
    g = make(V, False)
    o = O.OracleGraph(V, False)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    n = 2000
    del_s, del_d = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    ins_s, ins_d = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    g.reserve(n)
    with g.capture() as cap:
        g.batchDelete(del_s, del_d, stats=False)
        g.batchInsert(ins_s, ins_d, stats=False)
        G._check(G.lib().gv_compact(g.handle, 0.25, None))
    c = cap.result
    for rep in range(3):
        off, cols = o.export_csr()
        hs, hd = O.gen_sample(O.stream_key(700 + rep, 3), V, off, cols, n)
        o.delete(hs, hd)
        is_, id_ = O.gen_uniform(O.stream_key(700 + rep, 2), V, n)
        o.insert(is_, id_)
        g.sync()
        del_s.upload(hs)
        del_d.upload(hd)
        ins_s.upload(is_)
        ins_d.upload(id_)
        st = c.launch(stats=True)
        assert st["requested"] == 2 * n
        assert st["live_edges"] == o.num_edges()
        assert_same(g, o, f"replay {rep}")
    c.close()
    g.close()


def test_capture_rules():
    V = 1024
    g = make(V, False)
    n = 100
    dsrc, ddst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    hs, hd = O.gen_uniform(O.stream_key(33, 1), V, n)
    dsrc.upload(hs)
    ddst.upload(hd)
    g.reserve(n, 1)
    # a call that cannot be recorded fails, and so does the capture
    with pytest.raises(G.GraphVineError):
        with g.capture():
            g.num_edges()
    # host inputs cannot be recorded
    with pytest.raises(G.GraphVineError):
        with g.capture():
            g.batchInsert(hs, hd, stats=False)
    # a group larger than the reservation is refused inside a capture
    big_n = 50_000
    bs, bd = DeviceArray(big_n, np.uint32), DeviceArray(big_n, np.uint32)
    with pytest.raises(G.GraphVineError):
        with g.capture():
            g.batchInsert(bs, bd, stats=False)
    # the handle still works eagerly after refused captures
    o = O.OracleGraph(V, False)
    g.batchInsert(hs, hd)
    o.insert(hs, hd)
    assert_same(g, o)
    # a valid capture works until a later call reallocates its buffers
    with g.capture() as cap:
        g.batchDelete(dsrc, ddst, stats=False)
    c = cap.result
    c.launch()
    o.delete(hs, hd)
    assert_same(g, o)
    ws, wd = O.gen_uniform(O.stream_key(34, 1), V, big_n)
    g.batchInsert(ws, wd)
    o.insert(ws, wd)
    with pytest.raises(G.GraphVineError):
        c.launch()
    assert_same(g, o)
    c.close()
    g.close()


def test_set_tuning_keeps_pagerank():
    V = 1 << 13
    s, d = O.gen_rmat(O.stream_key(35, O.STREAM_RMAT), 13, 16 * V)
    g = make(V, False)
    g.bulkBuild(s, d)
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    for tm, gm, wm, layout in [(1, 1, 64, 1), (32, 512, 100_000, 1), (8, 96, 4096, 4),
                               (1, 1, 64, 4), (32, 512, 100_000, 4), (8, 96, 4096, 5)]:
        g.set_tuning(pr_thread_max=tm, pr_group_max=gm, pr_wave_max=wm, pr_layout=layout)
        cfg = g.get_config()
        assert (cfg.pr_thread_max, cfg.pr_group_max, cfg.pr_wave_max, cfg.pr_layout) == \
            (tm, gm, wm, layout)
        x, it, _ = G.pagerank(g, epsilon=1e-7)
        xo, _, _ = o.pagerank(fixed_iters=it)
        assert np.abs(x - xo).sum() <= 1e-6
    with pytest.raises(G.GraphVineError):
        g.set_tuning(pr_thread_max=100, pr_group_max=10)
    for bad in (0, 2, 3, 6):
        with pytest.raises(G.GraphVineError):
            g.set_tuning(pr_layout=bad)
    for r in (0.0, 2.5):
        g.set_tuning(bc_bu_ratio=r)
        assert g.get_config().bc_bu_ratio == r
    for bad in (-1.0, float("inf"), float("nan")):
        with pytest.raises(G.GraphVineError):
            g.set_tuning(bc_bu_ratio=bad)
    with pytest.raises(ValueError):
        g.set_tuning(block_slots=32)
    g.close()


def test_device_queue_with_more_batches_than_staged_offsets():
    """A queue of 3000 batch slots (its offsets searched in global memory, not
    staged) with a few non-empty batches among empty ones."""
    V = 1 << 11
    s, d = O.gen_rmat(O.stream_key(38, O.STREAM_RMAT), 11, 16 * V)
    g = make(V, False)
    o = O.OracleGraph(V, False)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    nb, cap = 3000, 8000
    sizes = np.zeros(nb, np.int64)
    kinds = np.full(nb, G.GV_OP_INSERT, np.int32)
    hs = np.full(cap, 0xFFFFFFFF, np.uint32)
    hd = np.full(cap, 0xFFFFFFFF, np.uint32)
    at = 0
    for j, b in enumerate((3, 700, 701, 1999, 2047, 2048, 2500, 2999)):
        n = 500 + 100 * j
        if j % 3 == 2:
            off, cols = o.export_csr()
            bs, bd = O.gen_sample(O.stream_key(38, 100 + j), V, off, cols, n)
            o.delete(bs, bd)
            kinds[b] = G.GV_OP_DELETE
        else:
            bs, bd = O.gen_uniform(O.stream_key(38, 200 + j), V, n)
            o.insert(bs, bd)
        sizes[b] = n
        hs[at:at + n], hd[at:at + n] = bs, bd
        at += n
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    qs, qd = DeviceArray.from_numpy(hs), DeviceArray.from_numpy(hd)
    qo, qk = DeviceArray.from_numpy(offsets), DeviceArray.from_numpy(kinds)
    g.applyBatchesDevice(qs, qd, qo, qk, compact_fraction=0.25)
    g.sync()
    assert_same(g, o, "3000-slot queue")
    g.close()
