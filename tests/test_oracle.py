"""Pins the CPU oracle (oracle/gv_oracle.c) without the reference binary.

The reference ships no tests, fixtures or golden vectors, and it is CUDA-only
(not buildable here), so every oracle function is pinned by a published
known-answer value, a closed form, brute force on small inputs, or a library
routine (networkx), chosen so a plausible slip (dropped term, wrong index,
off-by-one level, swapped operands) fails at least one check.
"""
import itertools
import json
import math
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

MASK = (1 << 64) - 1


def py_fmix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def py_rand(key, i):
    return py_fmix((key + (i + 1) * 0x9E3779B97F4A7C15) & MASK)


# ------------------------------------------------------------- generators --

def test_splitmix64_known_answers():
    # Published SplitMix64 outputs for state 0 (Steele, Lea & Flood 2014;
    # the reference sequence used to seed xoshiro generators).
    assert O.lib().orc_rand64(0, 0) == 0xE220A8397B1DCDAF
    assert O.lib().orc_rand64(0, 1) == 0x6E789E6AA1B965F4
    assert O.lib().orc_rand64(0, 2) == 0x06C45D188009454F


def test_rmat_matches_independent_python():
    key = O.stream_key(7, O.STREAM_RMAT)
    src, dst = O.gen_rmat(key, 10, 500)
    ta, tab, tabc = O.thresholds()
    for e in range(500):
        u = v = 0
        for lvl in range(10):
            r = py_rand(key, e * 10 + lvl) >> 32
            bit = 1 << (9 - lvl)
            if r < ta:
                pass
            elif r < tab:
                v |= bit
            elif r < tabc:
                u |= bit
            else:
                u |= bit
                v |= bit
        assert (src[e], dst[e]) == (u, v)


def test_rmat_quadrant_frequencies():
    # a/b/c/d = 0.57/0.19/0.19/0.05 -> top-bit quadrant frequencies
    src, dst = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 12, 200000)
    top_u = src >> 11
    top_v = dst >> 11
    f = [np.mean((top_u == a) & (top_v == b)) for a, b in [(0, 0), (0, 1), (1, 0), (1, 1)]]
    assert abs(f[0] - 0.57) < 0.005 and abs(f[1] - 0.19) < 0.005
    assert abs(f[2] - 0.19) < 0.005 and abs(f[3] - 0.05) < 0.005


def test_uniform_and_sample_match_python():
    key = O.stream_key(3, O.STREAM_INSERT)
    V = 1000
    src, dst = O.gen_uniform(key, V, 300)
    for i in range(300):
        r = py_rand(key, i)
        u = ((r >> 32) * V) >> 32
        v = ((r & 0xFFFFFFFF) * V) >> 32
        if u == v:
            v = 0 if v + 1 == V else v + 1
        assert (src[i], dst[i]) == (u, v)
        assert src[i] != dst[i]
    off = np.array([0, 2, 2, 5], np.uint64)
    cols = np.array([1, 2, 0, 1, 2], np.uint32)
    s, d = O.gen_sample(key, 3, off, cols, 50)
    for i in range(50):
        idx = py_rand(key, i) % 5
        u = int(np.searchsorted(off, idx, side="right") - 1)
        assert (s[i], d[i]) == (u, cols[idx])


def test_generator_golden_digest():
    with open(os.path.join(GOLDEN, "generators.json")) as f:
        gold = json.load(f)
    for case in gold["rmat"]:
        s, d = O.gen_rmat(O.stream_key(case["seed"], O.STREAM_RMAT), case["scale"], case["n"])
        assert s[:8].tolist() == case["src_head"] and d[:8].tolist() == case["dst_head"]
        k = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
        assert int(np.bitwise_xor.reduce(k * np.uint64(0x9E3779B97F4A7C15))) == case["xor_digest"]
    for case in gold["chung_lu"]:
        s, d = O.gen_chung_lu(O.stream_key(case["seed"], O.STREAM_CHUNG_LU), case["V"],
                              case["gamma"], case["avg_deg"], case["max_deg"], case["n"])
        assert s[:8].tolist() == case["src_head"] and d[:8].tolist() == case["dst_head"]
        k = (s.astype(np.uint64) << np.uint64(32)) | d.astype(np.uint64)
        assert int(np.bitwise_xor.reduce(k * np.uint64(0x9E3779B97F4A7C15))) == case["xor_digest"]


def test_chung_lu_degree_law():
    """Endpoint frequencies follow the capped power-law weights: the heaviest
    vertex's share of 2n draws matches w_0 / sum(w) within sampling error."""
    V, n, gamma, avg, mx = 20000, 200000, 2.1, 24.0, 1e9
    s, d = O.gen_chung_lu(O.stream_key(7, O.STREAM_CHUNG_LU), V, gamma, avg, mx, n)
    w = (np.arange(V, dtype=np.float64) + 1.0) ** (-1.0 / (gamma - 1.0))
    p0 = w[0] / w.sum()
    c0 = int((s == 0).sum() + (d == 0).sum())
    assert abs(c0 - 2 * n * p0) < 5 * np.sqrt(2 * n * p0)
    assert s.max() < V and d.max() < V


# ------------------------------------------------------------------ store --

def brute_apply(adj, src, dst, directed, delete):
    changed = 0
    for u, v in zip(src.tolist(), dst.tolist()):
        if u == v:
            continue
        pairs = [(u, v)] if directed else [(u, v), (v, u)]
        for a, b in pairs:
            if delete and b in adj[a]:
                adj[a].discard(b)
                changed += 1
            elif not delete and b not in adj[a]:
                adj[a].add(b)
                changed += 1
    return changed


def csr_of(adj, V):
    off = [0]
    cols = []
    for u in range(V):
        cols += sorted(adj[u])
        off.append(len(cols))
    return np.array(off, np.uint64), np.array(cols, np.uint32)


@pytest.mark.parametrize("directed", [True, False])
def test_store_matches_brute_force_sets(directed):
    rng = np.random.default_rng(5)
    V = 60
    g = O.OracleGraph(V, directed)
    adj = [set() for _ in range(V)]
    s = rng.integers(0, V, 400).astype(np.uint32)
    d = rng.integers(0, V, 400).astype(np.uint32)
    assert g.bulk_build(s, d) == brute_apply(adj, s, d, directed, False)
    for step in range(30):
        s = rng.integers(0, V, 50).astype(np.uint32)
        d = rng.integers(0, V, 50).astype(np.uint32)
        if step % 3 == 2:  # deletes of existing edges plus misses
            off, cols = csr_of(adj, V)
            if len(cols):
                idx = rng.integers(0, len(cols), 40)
                s[:40] = np.searchsorted(off, idx, side="right") - 1
                d[:40] = cols[idx]
            assert g.delete(s, d) == brute_apply(adj, s, d, directed, True)
        else:
            assert g.insert(s, d) == brute_apply(adj, s, d, directed, False)
        off, cols = g.export_csr()
        eo, ec = csr_of(adj, V)
        assert np.array_equal(off, eo) and np.array_equal(cols, ec)
        assert np.array_equal(g.degrees(), np.diff(eo).astype(np.uint32))


def test_store_special_cases():
    g = O.OracleGraph(5, True)
    assert g.bulk_build([0, 0, 0, 1], [1, 1, 0, 2]) == 2  # dup + self-loop dropped
    assert g.insert([0], [1]) == 0  # existing edge: no-op
    assert g.delete([3], [4]) == 0  # missing edge: no-op
    assert g.delete([0, 1], [1, 2]) == 2
    assert g.num_edges() == 0
    with pytest.raises(ValueError):
        g.insert([0], [5])
    assert g.num_edges() == 0
    u = O.OracleGraph(4, False)
    assert u.bulk_build([0, 1], [1, 0]) == 2  # both directions, one undirected edge
    assert u.degrees().tolist() == [1, 1, 0, 0]


def test_store_hub_over_reference_limit():
    # The reference caps a vertex at 4096 blocks x 128 = 524,288 edges
    # (constants.cuh:10); the contract has no cap.
    V = 600_001
    g = O.OracleGraph(V, True)
    d = np.arange(1, V, dtype=np.uint32)
    assert g.bulk_build(np.zeros(V - 1, np.uint32), d) == V - 1
    assert g.degrees()[0] == V - 1


# -------------------------------------------------------------- PageRank --

def test_pagerank_cycle_is_uniform_every_iteration():
    V = 7
    g = O.OracleGraph(V, True)
    g.bulk_build(np.arange(V), (np.arange(V) + 1) % V)
    for k in range(1, 5):
        x, it, r = g.pagerank(fixed_iters=k)
        assert it == k and np.allclose(x, 1.0 / V, rtol=0, atol=1e-15) and r < 1e-15


def test_pagerank_start_vector():
    """x0 = 1/V reproduces the default start bit for bit; on a directed cycle
    (v -> v+1) from an arbitrary start the iterate has the closed form
    x_t[v] = (1-a)/V * sum_{k<t} a^k + a^t * x0[v-t mod V], which pins both the
    start and the direction of the shift; and a start at the fixed point stops
    after one update."""
    a = 0.85
    g = O.OracleGraph(40, False)
    s, d = O.gen_rmat(O.stream_key(5, O.STREAM_RMAT), 6, 300)
    g.bulk_build(s % 40, d % 40)
    x_def, it_def, r_def = g.pagerank(eps=1e-9, max_iter=500)
    x_one, it_one, r_one = g.pagerank(eps=1e-9, max_iter=500, x0=np.full(40, 1.0 / 40))
    assert it_one == it_def and r_one == r_def and np.array_equal(x_one, x_def)
    V = 9
    c = O.OracleGraph(V, True)
    c.bulk_build(np.arange(V), (np.arange(V) + 1) % V)
    x0 = np.random.default_rng(3).random(V)
    for t in range(1, 12):
        x, it, _ = c.pagerank(alpha=a, fixed_iters=t, x0=x0)
        want = (1 - a) / V * sum(a ** k for k in range(t)) + a ** t * np.roll(x0, t)
        assert it == t and np.allclose(x, want, rtol=0, atol=1e-15)
    xf, _, _ = g.pagerank(eps=1e-15, max_iter=5000)
    x2, it2, r2 = g.pagerank(eps=1e-9, max_iter=500, x0=xf)
    assert it2 == 1 and r2 < 1e-9 and np.abs(x2 - xf).max() < 1e-15
    with pytest.raises(ValueError):
        g.pagerank(x0=np.ones(3))


def test_pagerank_two_vertex_dangling_dropped():
    a = 0.85
    g = O.OracleGraph(2, True)
    g.bulk_build([0], [1])
    xa, xb = 0.5, 0.5
    for k in range(1, 6):
        xa, xb = (1 - a) / 2, (1 - a) / 2 + a * xa
        x, it, _ = g.pagerank(alpha=a, fixed_iters=k)
        assert abs(x[0] - xa) < 1e-15 and abs(x[1] - xb) < 1e-15
    # dangling mass dropped: sum < 1
    assert x.sum() < 1.0


def test_pagerank_star_fixed_point():
    # undirected star, centre 0 with m leaves: x_c = (1-a)/V + a*m*x_l,
    # x_l = (1-a)/V + a*x_c/m
    a, m = 0.85, 6
    V = m + 1
    g = O.OracleGraph(V, False)
    g.bulk_build(np.zeros(m), np.arange(1, V))
    x, it, r = g.pagerank(alpha=a, eps=1e-14, max_iter=10000)
    b = (1 - a) / V
    xl = (b + a * b / m) / (1 - a * a)
    xc = b + a * m * xl
    assert abs(x[0] - xc) < 1e-12 and np.allclose(x[1:], xl, atol=1e-12)


def test_pagerank_stopping_rule_and_max_iter():
    g = O.OracleGraph(50, True)
    s, d = O.gen_rmat(O.stream_key(2, O.STREAM_RMAT), 6, 400)
    g.bulk_build(s % 50, d % 50)
    x, it, r = g.pagerank(eps=1e-6, max_iter=1000)
    assert r < 1e-6
    _, _, r_prev = g.pagerank(fixed_iters=it - 1)
    assert r_prev >= 1e-6  # first update below eps stops the loop
    _, it3, _ = g.pagerank(eps=1e-30, max_iter=3)
    assert it3 == 3


def test_pagerank_matches_networkx_on_dangling_free_graph():
    nx = pytest.importorskip("networkx")
    V = 200
    g = O.OracleGraph(V, False)
    s, d = O.gen_rmat(O.stream_key(4, O.STREAM_RMAT), 8, 3000)
    s, d = s % V, d % V
    # ring keeps every vertex non-dangling
    s = np.concatenate([s, np.arange(V, dtype=np.uint32)])
    d = np.concatenate([d, ((np.arange(V) + 1) % V).astype(np.uint32)])
    g.bulk_build(s, d)
    x, _, _ = g.pagerank(eps=1e-13, max_iter=1000)
    G = nx.Graph()
    G.add_nodes_from(range(V))
    G.add_edges_from((int(a), int(b)) for a, b in zip(s, d) if a != b)
    ref = nx.pagerank(G, alpha=0.85, tol=1e-14, max_iter=1000)
    assert max(abs(x[v] - ref[v]) for v in range(V)) < 1e-10


# ------------------------------------------------------- triangle counting --

def brute_tc(adj, V):
    t = [0] * V
    for u in range(V):
        for v in range(u + 1, V):
            for w in range(v + 1, V):
                if v in adj[u] and w in adj[v] and u in adj[w]:
                    t[u] += 1
    return t


@pytest.mark.parametrize("n", [3, 5, 9])
def test_tc_complete_graph_closed_form(n):
    g = O.OracleGraph(n, False)
    e = list(itertools.combinations(range(n), 2))
    g.bulk_build([a for a, _ in e], [b for _, b in e])
    pv, tot = g.triangles()
    assert tot == math.comb(n, 3)
    assert pv.tolist() == [math.comb(n - 1 - i, 2) for i in range(n)]


def test_tc_wheel_and_bipartite():
    m = 8
    g = O.OracleGraph(m + 1, False)
    s = [0] * m + list(range(1, m + 1))
    d = list(range(1, m + 1)) + [1 + (i % m) for i in range(1, m + 1)]
    g.bulk_build(s, d)
    pv, tot = g.triangles()
    assert tot == m and pv[0] == m
    b = O.OracleGraph(10, False)
    b.bulk_build([i for i in range(5) for _ in range(5)], [j for _ in range(5) for j in range(5, 10)])
    assert b.triangles()[1] == 0


def test_tc_brute_force_directed_and_undirected():
    rng = np.random.default_rng(9)
    for directed in (True, False):
        V = 25
        g = O.OracleGraph(V, directed)
        s = rng.integers(0, V, 150).astype(np.uint32)
        d = rng.integers(0, V, 150).astype(np.uint32)
        g.bulk_build(s, d)
        adj = [set() for _ in range(V)]
        brute_apply(adj, s, d, directed, False)
        pv, tot = g.triangles()
        assert pv.tolist() == brute_tc(adj, V) and tot == sum(brute_tc(adj, V))


def test_tc_total_matches_networkx():
    nx = pytest.importorskip("networkx")
    V = 300
    s, d = O.gen_rmat(O.stream_key(6, O.STREAM_RMAT), 9, 4000)
    s, d = s % V, d % V
    g = O.OracleGraph(V, False)
    g.bulk_build(s, d)
    G = nx.Graph()
    G.add_edges_from((int(a), int(b)) for a, b in zip(s, d) if a != b)
    assert g.triangles()[1] == sum(nx.triangles(G).values()) // 3


# -------------------------------------------------------------------- BC --

def test_bc_path_star_complete_closed_forms():
    n = 9
    p = O.OracleGraph(n, False)
    p.bulk_build(np.arange(n - 1), np.arange(1, n))
    bcv = p.bc(np.arange(n))
    assert np.allclose(bcv, [2 * i * (n - 1 - i) for i in range(n)])
    st = O.OracleGraph(n, False)
    st.bulk_build(np.zeros(n - 1), np.arange(1, n))
    bcs = st.bc(np.arange(n))
    assert abs(bcs[0] - (n - 1) * (n - 2)) < 1e-9 and np.allclose(bcs[1:], 0)
    k = O.OracleGraph(6, False)
    e = list(itertools.combinations(range(6), 2))
    k.bulk_build([a for a, _ in e], [b for _, b in e])
    assert np.allclose(k.bc(np.arange(6)), 0)


def test_bc_matches_networkx_directed_all_sources():
    nx = pytest.importorskip("networkx")
    V = 120
    s, d = O.gen_rmat(O.stream_key(8, O.STREAM_RMAT), 7, 900)
    s, d = s % V, d % V
    g = O.OracleGraph(V, True)
    g.bulk_build(s, d)
    G = nx.DiGraph()
    G.add_nodes_from(range(V))
    G.add_edges_from((int(a), int(b)) for a, b in zip(s, d) if a != b)
    ref = nx.betweenness_centrality(G, normalized=False)
    bcv = g.bc(np.arange(V))
    assert max(abs(bcv[v] - ref[v]) for v in range(V)) < 1e-9


def brute_bc(adj, V, sources):
    """Enumerate all shortest paths (BFS layering + path counting by DFS)."""
    out = [0.0] * V
    for s in sources:
        dist = {s: 0}
        q = [s]
        for u in q:
            for w in adj[u]:
                if w not in dist:
                    dist[w] = dist[u] + 1
                    q.append(w)
        for t in dist:
            if t == s:
                continue
            paths = []

            def walk(u, path):
                if u == t:
                    paths.append(list(path))
                    return
                for w in adj[u]:
                    if dist.get(w) == dist[u] + 1 and dist[w] <= dist[t]:
                        path.append(w)
                        walk(w, path)
                        path.pop()

            walk(s, [s])
            for pth in paths:
                for v in pth[1:-1]:
                    out[v] += 1.0 / len(paths)
    return out


def test_bc_subset_brute_force():
    rng = np.random.default_rng(2)
    V = 18
    s = rng.integers(0, V, 45).astype(np.uint32)
    d = rng.integers(0, V, 45).astype(np.uint32)
    g = O.OracleGraph(V, True)
    g.bulk_build(s, d)
    adj = [set() for _ in range(V)]
    brute_apply(adj, s, d, True, False)
    src = [0, 3, 7, 11]
    assert np.allclose(g.bc(src), brute_bc(adj, V, src), atol=1e-12)


# ------------------------------------------------------ OpenMP baseline --

def test_omp_baseline_matches_oracle():
    V = 2000
    s, d = O.gen_rmat(O.stream_key(1, O.STREAM_RMAT), 11, 20000)
    g = O.OracleGraph(V, False)
    g.bulk_build(s % V, d % V)
    off, cols = g.export_csr()
    for b in range(3):
        si, di = O.gen_uniform(O.stream_key(1, 100 + b), V, 1000)
        g.insert(si, di)
        off, cols = O.omp_apply_batch(V, False, off, cols, si, di, delete=False)
        eo, ec = g.export_csr()
        assert np.array_equal(off, eo) and np.array_equal(cols, ec)
        sd, dd = O.gen_sample(O.stream_key(1, 200 + b), V, off, cols, 1000)
        g.delete(sd, dd)
        off, cols = O.omp_apply_batch(V, False, off, cols, sd, dd, delete=True)
        eo, ec = g.export_csr()
        assert np.array_equal(off, eo) and np.array_equal(cols, ec)
    x, it = O.omp_pagerank_sym(V, off, cols)
    xo, ito, _ = g.pagerank()
    assert it == ito and np.abs(x - xo).sum() < 1e-12


def test_omp_pagerank_start_vector_and_fixed_iterations():
    V = 1500
    s, d = O.gen_rmat(O.stream_key(7, O.STREAM_RMAT), 11, 12000)
    g = O.OracleGraph(V, False)
    g.bulk_build(s % V, d % V)
    off, cols = g.export_csr()
    x0 = np.random.default_rng(7).random(V) / V
    for fixed in (1, 6):
        x, it, r = O.omp_pagerank_sym(V, off, cols, fixed_iters=fixed, x0=x0, residual=True)
        xo, ito, ro = g.pagerank(fixed_iters=fixed, x0=x0)
        assert it == ito == fixed and np.abs(x - xo).max() <= 1e-15
        assert abs(r - ro) <= 1e-12 * ro
    x, it, r = O.omp_pagerank_sym(V, off, cols, eps=1e-9, max_iter=300, x0=x0, residual=True)
    xo, ito, ro = g.pagerank(eps=1e-9, max_iter=300, x0=x0)
    assert it == ito and r < 1e-9 and np.abs(x - xo).max() <= 1e-15


@pytest.mark.parametrize("seed,V,n", [(8, 300, 3000), (9, 2048, 40000), (10, 40, 700)])
def test_omp_triangle_count_matches_oracle(seed, V, n):
    """The OpenMP triangle count (full-size reference of the
    benchmark's triangle parity) equals the sequential oracle per vertex --
    itself pinned by closed forms and brute force above -- on R-MAT graphs
    with hubs (orientation by degree puts hubs last), and on K_n."""
    s, d = O.gen_rmat(O.stream_key(seed, O.STREAM_RMAT), 12, n)
    g = O.OracleGraph(V, False)
    g.bulk_build(s % V, d % V)
    off, cols = g.export_csr()
    pv, tot = O.omp_triangle_count(V, off, cols)
    pvo, toto = g.triangles()
    assert tot == toto > 0 and np.array_equal(pv, pvo)
    _, tot2 = O.omp_triangle_count(V, off, cols, per_vertex=False)
    assert tot2 == tot
    k = 9
    kn = O.OracleGraph(k, False)
    iu, ju = np.triu_indices(k, 1)
    kn.bulk_build(iu.astype(np.uint32), ju.astype(np.uint32))
    off, cols = kn.export_csr()
    pv, tot = O.omp_triangle_count(k, off, cols)
    assert tot == k * (k - 1) * (k - 2) // 6
    assert pv.tolist() == [(k - 1 - i) * (k - 2 - i) // 2 for i in range(k)]


@pytest.mark.parametrize("directed", [True, False])
def test_omp_fp64_references_match_the_oracle(directed):
    """The OpenMP FP64 PageRank and BC (the references of the precision
    checks) give the sequential oracle's values on the same graph."""
    V = 1 << 11
    s, d = O.gen_rmat(O.stream_key(61, O.STREAM_RMAT), 11, 8 * V)
    o = O.OracleGraph(V, directed)
    o.bulk_build(s, d)
    off, cols = o.export_csr()
    x, it, r = O.omp_pagerank_dir(V, off, cols, eps=1e-10, max_iter=500)
    xo, ito, ro = o.pagerank(eps=1e-10, max_iter=500)
    assert it == ito and abs(r - ro) <= 1e-12 * max(r, 1e-30) + 1e-18
    assert np.abs(x - xo).max() <= 1e-15
    x5, _, _ = O.omp_pagerank_dir(V, off, cols, fixed_iters=5)
    xo5, _, _ = o.pagerank(fixed_iters=5)
    assert np.abs(x5 - xo5).max() <= 1e-15
    src = np.random.default_rng(61).choice(V, 9, replace=False).astype(np.uint32)
    b, bo = O.omp_bc_dir(V, off, cols, src), o.bc(src)
    assert np.allclose(b, bo, rtol=1e-12, atol=1e-9)
    ioff, icols = O.omp_transpose(V, off, cols)
    assert ioff[-1] == off[-1]
    assert np.array_equal(np.diff(ioff), np.bincount(cols, minlength=V))


@pytest.mark.parametrize("directed", [True, False])
def test_omp_personalized_pagerank_equals_the_oracle_bitwise(directed):
    """The OpenMP personalized PageRank (the full-size reference of the
    64-column block) runs the oracle's per-column operations in the oracle's
    order, so both its stopping iteration and every value equal the
    oracle's exactly: on a graph with dangling vertices and repeated seeds,
    converging and at a fixed iteration count."""
    V = 1500
    s, d = O.gen_rmat(O.stream_key(62, O.STREAM_RMAT), 11, 6 * V)
    o = O.OracleGraph(V, directed)
    o.bulk_build(s % V, d % V)
    off, cols = o.export_csr()
    assert directed is False or (np.diff(off) == 0).any()
    seeds = np.array([0, 7, 7, 1499, 3, 100, 42], np.uint32)
    X, it, r = O.omp_pagerank_personalized(V, off, cols, seeds, 0.85, 1e-6, 200)
    Xo, ito, ro = o.pagerank_personalized(seeds, 0.85, 1e-6, 200)
    assert it == ito and r == ro and 1 < it < 200
    assert np.array_equal(X, Xo)
    X3, it3, r3 = O.omp_pagerank_personalized(V, off, cols, seeds, fixed_iters=3)
    Xo3, ito3, ro3 = o.pagerank_personalized(seeds, fixed_iters=3)
    assert it3 == ito3 == 3 and r3 == ro3 and np.array_equal(X3, Xo3)
