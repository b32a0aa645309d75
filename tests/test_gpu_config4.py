"""The batching workload at a size the oracle finishes in seconds: an
undirected Chung-Lu power-law graph (V = 2^17, degree exponent 2.1, average
degree 24, seed 4) whose degree cap is scaled from 200k at 8M vertices to
20000 here, so hub rows still span dozens of 256-slot items and exceed the
4096 in-neighbours above which PageRank rows are summed in chunks.

  * a group of 48 uneven batches (sizes log-uniform in [64, 40000], 70%
    inserts drawn from the same Chung-Lu weights, 30% deletes sampled from
    the graph as it evolves) applied by gv_batch_apply, and by a captured
    replay of the same pipeline (twice, inputs refilled), leaves exactly the
    oracle's sorted neighbour sets after applying the batches in order;
  * 64 personalized PageRank columns (one of them seeded at the largest hub)
    within L1 1e-6 per column of the FP64 reference at the device's
    iteration count (the OpenMP form of the oracle's rule, equal to the
    oracle bit for bit), stopping within one iteration of it;
  * BC from 64 sources (two sweeps of 32) within relative 1e-5 of the oracle.
"""
import numpy as np
import pytest

import gv_amd as G
from gv_amd.hip import DeviceArray
from oracle import oracle as O

pytestmark = pytest.mark.gpu

V = 1 << 17
GAMMA, AVG, CAP = 2.1, 24.0, 20000.0
SEED = 4
NB = 48
PR_L1 = 1e-6
BC_REL = 1e-5


def group_shape(seed):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(64), np.log(40000), NB)).astype(np.int64)
    kinds = [G.GV_OP_INSERT if r < 0.7 else G.GV_OP_DELETE for r in rng.random(NB)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return sizes, kinds, offs


def fill_group(o, sizes, kinds, key_base):
    """Batches of the given shape, applied to the oracle in order: inserts
    from the Chung-Lu weights, deletes sampled from the current graph."""
    srcs, dsts = [], []
    for b, (n, k) in enumerate(zip(sizes, kinds)):
        if k == G.GV_OP_INSERT:
            s, d = O.gen_chung_lu(O.stream_key(SEED, key_base + b), V, GAMMA, AVG, CAP, int(n))
            o.insert(s, d)
        else:
            off, cols = o.export_csr()
            s, d = O.gen_sample(O.stream_key(SEED, key_base + 1000 + b), V, off, cols, int(n))
            o.delete(s, d)
        srcs.append(s)
        dsts.append(d)
    return np.concatenate(srcs), np.concatenate(dsts)


@pytest.fixture(scope="module")
def workload():
    """Host side, computed once: base edges, two groups of one shape and the
    oracle's CSR after the base build and after each group."""
    s, d = O.gen_chung_lu(O.stream_key(SEED, O.STREAM_CHUNG_LU), V, GAMMA, AVG, CAP,
                          int(V * AVG / 2))
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    sizes, kinds, offs = group_shape(SEED)
    groups, csrs = [], []
    for gi in range(2):
        gs, gd = fill_group(o, sizes, kinds, 6000 + 2000 * gi)
        groups.append((gs, gd))
        csrs.append(o.export_csr())
    return {"base": (s, d), "offs": offs, "kinds": kinds, "groups": groups, "csrs": csrs}


def make():
    return G.GraphVine(V, directed=False)


def assert_csr(g, csr, where):
    off, cols = g.export_csr()
    assert np.array_equal(off, csr[0]), f"degrees differ {where}"
    assert np.array_equal(cols, csr[1]), f"neighbour sets differ {where}"


def test_workload_has_hub_rows_past_the_chunk_size(workload):
    off, _ = workload["csrs"][0]
    deg = np.diff(off.astype(np.int64))
    assert deg.max() > 4096 * 2 and (deg > 4096).sum() >= 10
    assert deg.max() > 32 * 256
    sizes = np.diff(workload["offs"].astype(np.int64))
    assert sizes.min() >= 64 and sizes.max() > 10000
    assert 0 < workload["kinds"].count(G.GV_OP_DELETE) < NB


def test_uneven_group_batched_and_captured_equal_the_oracle(workload):
    s, d = workload["base"]
    offs, kinds = workload["offs"], workload["kinds"]
    # gv_batch_apply, host inputs, compaction at 25% tombstones
    g = make()
    g.bulkBuild(s, d)
    for gi, (gs, gd) in enumerate(workload["groups"]):
        g.applyBatches(gs, gd, offs, kinds, compact_fraction=0.25)
        assert_csr(g, workload["csrs"][gi], f"batched group {gi}")
    g.close()
    # the same pipeline captured once, replayed with refilled device inputs
    g = make()
    g.bulkBuild(s, d)
    n = int(offs[-1])
    dsrc, ddst = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    g.reserve(n, NB)
    with g.capture() as cap:
        g.applyBatches(dsrc, ddst, offs, kinds, compact_fraction=0.25, stats=False)
    c = cap.result
    for gi, (gs, gd) in enumerate(workload["groups"]):
        g.sync()
        dsrc.upload(gs)
        ddst.upload(gd)
        c.launch()
        g.sync()
        assert_csr(g, workload["csrs"][gi], f"captured replay {gi}")
    c.close()
    g.close()


def graph_after_first_group(workload):
    s, d = workload["base"]
    g = make()
    g.bulkBuild(s, d)
    gs, gd = workload["groups"][0]
    g.applyBatches(gs, gd, workload["offs"], workload["kinds"], compact_fraction=0.25)
    return g


def test_personalized_pagerank_64_columns(workload):
    g = graph_after_first_group(workload)
    off, cols = workload["csrs"][0]
    deg = np.diff(off.astype(np.int64))
    rng = np.random.default_rng(44)
    seeds = np.concatenate([[int(deg.argmax())],
                            rng.choice(V, 63, replace=False)]).astype(np.uint32)
    X, it, r = G.pagerank_personalized(g, seeds, 0.85, 1e-6, 100)
    g.close()
    Xo, ito, _ = O.omp_pagerank_personalized(V, off, cols, seeds, 0.85, 1e-6, 100)
    assert abs(it - ito) <= 1, (it, ito)
    if it != ito:
        Xo, _, _ = O.omp_pagerank_personalized(V, off, cols, seeds, 0.85, 1e-6, 100,
                                               fixed_iters=it)
    l1 = np.abs(X.astype(np.float64) - Xo).sum(axis=0)
    assert l1.max() <= PR_L1, (l1.max(), int(l1.argmax()))


def test_bc_64_sources(workload):
    g = graph_after_first_group(workload)
    off, cols = workload["csrs"][0]
    src = np.random.default_rng(45).choice(V, 64, replace=False).astype(np.uint32)
    b = G.bc(g, src).astype(np.float64)
    g.close()
    bo = O.omp_bc_dir(V, off, cols, src)
    err = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
    assert err.max() <= BC_REL, err.max()
