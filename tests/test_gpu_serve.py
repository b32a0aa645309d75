"""Request loop (gv_server_*): every request served in the mix -- batched
with its neighbours in arrival order -- must give the answer it gives when
served alone on the graph that the requests before it define, and the final
graph must be the oracle's after applying the accepted updates in order."""
import numpy as np
import pytest

import gv_amd as G
from gv_amd import serve as S
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


PR_TOL = 1e-6  # PageRank L1 bar, against the oracle at equal iterations and mix vs alone


def oracle_from_csr(V, directed, off, cols):
    g = O.OracleGraph(V, directed)
    g.bulk_build(np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64)), cols)
    return g


def pagerank_starts(reqs, expect):
    """Start vector of each refresh in the mix: a refresh warm-starts from the
    result of the previous serving step that ran PageRank (1/V for the first);
    requests served by one step share its run."""
    starts, prev, prev_step, cur = {}, None, None, None
    order = sorted((r.response["step"], i) for i, (r, e) in enumerate(zip(reqs, expect))
                   if e[0] == "pr")
    for step, i in order:
        if step != prev_step:
            prev, prev_step = cur, step
        starts[i] = prev
        cur = reqs[i].vector().astype(np.float64)
    return starts


def make_trace(o, V, seed, n_req=120, burst=20):
    """Requests with expected answers computed on the oracle at each request's
    point in the arrival order."""
    rng = np.random.default_rng(seed)
    reqs, expect, arrivals = [], [], []
    for i in range(n_req):
        arrivals.append((i // burst) * 30_000.0)  # bursts of equal arrival time
        u = rng.random()
        if u < 0.55:
            n = int(rng.integers(10, 2000))
            s, d = O.gen_uniform(O.stream_key(seed, 100 + i), V, n)
            o.insert(s, d)
            reqs.append(S.insert(s, d))
            expect.append(("ins", s, d))
        elif u < 0.80:
            off, cols = o.export_csr()
            n = int(rng.integers(10, 800))
            s, d = O.gen_sample(O.stream_key(seed, 200 + i), V, off, cols, n)
            o.delete(s, d)
            reqs.append(S.delete(s, d))
            expect.append(("del", s, d))
        elif u < 0.90:
            vs = rng.integers(0, V, 64).astype(np.uint32)
            off, cols = o.export_csr()
            lists = [cols[off[v]:off[v + 1]].copy() for v in vs]
            reqs.append(S.neighbors(vs, out_cap=sum(len(x) for x in lists) + 16))
            expect.append(("nbr", vs, lists))
        elif u < 0.95:
            # the graph at this point; the start vector (the previous
            # refresh's result) is known only after the replay
            reqs.append(S.pagerank(V))
            expect.append(("pr",) + o.export_csr())
        else:
            src = int(rng.integers(0, V))
            reqs.append(S.bc(src, V))
            expect.append(("bc", src, o.bc(np.array([src], np.uint32))))
    return reqs, expect, np.array(arrivals)


@pytest.mark.parametrize("directed", [False, True])
def test_mix_equals_alone(directed):
    V = 1 << 12
    s, d = O.gen_rmat(O.stream_key(41, O.STREAM_RMAT), 12, 16 * V)
    g, ga = make(V, directed), make(V, directed)
    o = O.OracleGraph(V, directed)
    for h in (g, ga):
        h.bulkBuild(s, d)
    o.bulk_build(s, d)
    reqs, expect, arr = make_trace(o, V, 42 + directed)
    with S.Server(g) as srv:
        resp = srv.replay(reqs, arr)
        st = srv.stats()
    assert all(r["status"] == 0 for r in resp), [r["status"] for r in resp]
    assert max(r["step_size"] for r in resp) > 1  # batching happened
    assert st["requests_done"] == len(reqs) and st["steps"] < len(reqs)
    for r in resp:
        assert r["done_us"] >= r["start_us"] >= r["arrival_us"] - 1.0
    # in the mix vs the oracle at the same point of the arrival order; a
    # refresh against the oracle from the same start at equal iterations
    starts = pagerank_starts(reqs, expect)
    assert sum(e[0] == "pr" for e in expect) >= 2
    for i, (req, exp) in enumerate(zip(reqs, expect)):
        if exp[0] == "nbr":
            offs, out = req.neighbours()
            for k, lst in enumerate(exp[2]):
                assert np.array_equal(out[offs[k]:offs[k + 1]], lst)
        elif exp[0] == "pr":
            og = oracle_from_csr(V, directed, exp[1], exp[2])
            it = req.response["iters"]
            xo, _, _ = og.pagerank(fixed_iters=it, x0=starts[i])
            assert np.abs(req.vector().astype(np.float64) - xo).sum() <= PR_TOL
        elif exp[0] == "bc":
            b, bo = req.vector().astype(np.float64), exp[2]
            assert (np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)).max() <= 1e-5
    assert_same(g, o, "after the mix")
    # alone: each request on its own, in arrival order, on a second graph; a
    # PageRank refresh alone warm-starts from the previous refresh alone
    prev_alone = None
    for req, exp in zip(reqs, expect):
        if exp[0] == "ins":
            ga.batchInsert(exp[1], exp[2])
        elif exp[0] == "del":
            ga.batchDelete(exp[1], exp[2])
        elif exp[0] == "nbr":
            offs_a, out_a = ga.neighbors(exp[1])
            offs, out = req.neighbours()
            assert np.array_equal(offs_a, offs) and np.array_equal(out_a, out)
        elif exp[0] == "pr":
            xa, _, _ = G.pagerank(ga, init=prev_alone)
            prev_alone = xa
            assert np.abs(xa.astype(np.float64) - req.vector()).sum() <= PR_TOL
        else:
            ba = G.bc(ga, np.array([exp[1]], np.uint32)).astype(np.float64)
            b = req.vector().astype(np.float64)
            assert (np.abs(b - ba) / np.maximum(np.abs(ba), 1.0)).max() <= 1e-5
    assert_same(ga, o, "alone")
    g.close()
    ga.close()


def test_admission_refuses_what_the_pool_cannot_hold():
    V = 1000
    g = make(V, False, pool=1200, block_slots=32)
    o = O.OracleGraph(V, False)
    s, d = O.gen_uniform(O.stream_key(43, 1), V, 2000)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    r1 = O.gen_uniform(O.stream_key(43, 2), V, 100)
    r2 = O.gen_uniform(O.stream_key(43, 3), V, 20000)  # ~1000 more blocks: cannot fit
    off, cols = o.export_csr()
    r3 = O.gen_sample(O.stream_key(43, 4), V, off, cols, 50)
    r4 = O.gen_uniform(O.stream_key(43, 5), V, 50)
    reqs = [S.insert(*r1), S.insert(*r2), S.delete(*r3), S.insert(*r4)]
    with S.Server(g) as srv:
        resp = srv.replay(reqs, np.zeros(4))
        st = srv.stats()
    assert [r["status"] for r in resp] == [0, G.GV_E_POOL, 0, 0]
    assert st["refused"] == 1 and st["update_pairs_applied"] == 200
    o.insert(*r1)
    o.delete(*r3)
    o.insert(*r4)
    assert_same(g, o, "after refusal")
    g.close()


def test_submit_wait_and_drain():
    V = 2048
    g = make(V, False)
    o = O.OracleGraph(V, False)
    s, d = O.gen_rmat(O.stream_key(44, O.STREAM_RMAT), 11, 8 * V)
    g.bulkBuild(s, d)
    o.bulk_build(s, d)
    srv = S.Server(g, max_step_requests=3)
    tickets, reqs = [], []
    for i in range(10):
        a, b = O.gen_uniform(O.stream_key(45, i), V, 300)
        o.insert(a, b)
        reqs.append(S.insert(a, b))
        tickets.append(srv.submit(reqs[-1]))
    vs = np.arange(0, V, 97, dtype=np.uint32)
    q = S.neighbors(vs, out_cap=1 << 16)
    tq = srv.submit(q)
    r = srv.wait(tq)
    assert r["status"] == 0
    eo, ec = o.export_csr()
    offs, out = q.neighbours()
    for k, v in enumerate(vs):
        assert np.array_equal(out[offs[k]:offs[k + 1]], ec[eo[v]:eo[v + 1]])
    srv.drain()
    assert all(srv.wait(t)["step_size"] <= 3 for t in tickets)
    # a neighbour request whose output buffer is too small reports the size
    small = S.neighbors(np.array([0], np.uint32), out_cap=0)
    rs = srv.wait(srv.submit(small))
    assert rs["status"] == G.GV_E_INVALID and rs["out_len"] == eo[1] - eo[0]
    srv.close()
    assert_same(g, o)
    g.close()
