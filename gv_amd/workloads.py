"""Benchmark drivers for the store's larger workloads (the reference's driver
is main.cu:20-134, whose benchmark switch :76-128 is dead code).

Each run_* function builds its synthetic input on the device, times the
workload with HIP events on the graph's stream and returns one JSON-ready
dict.  Nothing here checks results against a host restatement; the parity
tests do that at sizes a CPU finishes in seconds.
"""
import ctypes
import time

import numpy as np

STREAM_RMAT = 1
STREAM_CHUNG_LU = 4


def _events():
    from .hip import Event
    return Event(), Event()


def _timed(g, fn):
    """Run fn() (which enqueues work on g's stream) between two events and
    return (result, elapsed ms)."""
    e0, e1 = _events()
    s = g.stream()
    e0.record(s)
    r = fn()
    e1.record(s)
    e1.synchronize()
    return r, e0.elapsed_ms(e1)


# ------------------------------------------------------------- Chung-Lu --
def chung_lu_stream(V, n_batches, seed=4):
    """Sizes (log-uniform in [1e3, 1e6]) and kinds (insert with probability
    0.7) of the Chung-Lu update stream, from a seeded host generator."""
    from . import GV_OP_DELETE, GV_OP_INSERT
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(1e3), np.log(1e6), n_batches)).astype(np.int64)
    kinds = [GV_OP_INSERT if r < 0.7 else GV_OP_DELETE for r in rng.random(n_batches)]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    return sizes, kinds, offs


def run_chunglu(args):
    """Chung-Lu power-law graph (gamma 2.1, average degree 24, degree cap
    200k, undirected, seed 4) and a stream of uneven batches (log-uniform
    sizes in [1k, 1M], 70% inserts drawn from the same weights, 30% deletes
    sampled from the built graph).  The stream runs three ways on fresh
    copies of the same graph:
      plain:    one gv_batch_insert / gv_batch_delete per batch, compaction
                (25% tombstones) after every delete batch;
      batched:  one gv_batch_apply over all batches + one compaction;
      captured: the batched pipeline recorded once (gv_capture_*) and
                replayed by one graph launch;
      queue:    gv_batch_apply_device recorded once over a device queue and
                replayed over four successive fills of that queue with
                uneven batch counts (the fills are device-to-device copies
                on the graph's stream, timed with the replays).
    The four final graphs must be identical.  Then 64 personalized PageRank
    columns and BC from 256 sources (32 per pass) run on the result."""
    from . import (GV_MEM_DEVICE, GV_OP_INSERT, GraphVine, _check, default_config,
                   export_csr_device, gen_chung_lu, lib, stream_key)
    from .autotune import overrides, tune
    from .hip import DeviceArray

    V = int(args.cl_vertices)
    seed = 4
    gamma, avg, mx = 2.1, 24.0, 200_000.0
    n_graph = int(V * avg / 2)
    nb = int(args.cl_batches)
    sizes, kinds, offs = chung_lu_stream(V, nb, seed)
    total = int(offs[-1])
    t_setup = time.time()

    g = GraphVine(V, directed=False)
    gsrc, gdst = gen_chung_lu(g, stream_key(seed, STREAM_CHUNG_LU), V, n_graph, gamma, avg, mx)
    g.bulkBuild(gsrc, gdst, stats=False)
    off, cols, E0 = export_csr_device(g)
    cs, cd = DeviceArray(total, np.uint32), DeviceArray(total, np.uint32)
    for b in range(nb):
        o, n = int(offs[b]), int(sizes[b])
        vs, vd = cs.view(o, n), cd.view(o, n)
        if kinds[b] == GV_OP_INSERT:
            _check(lib().gv_gen_chung_lu(g.handle, stream_key(seed, 6000 + b), V, gamma, avg, mx,
                                         n, vs.ptr, vd.ptr))
        else:
            _check(lib().gv_gen_sample(g.handle, stream_key(seed, 7000 + b), V, off.ptr,
                                       cols.ptr, n, vs.ptr, vd.ptr))
    g.sync()
    off.free()
    cols.free()
    g.close()
    setup_s = time.time() - t_setup

    def build(h):
        h.bulkBuild(gsrc, gdst, stats=False)

    def batched(h):
        h.applyBatches(cs, cd, offs, kinds, compact_fraction=0.25, stats=False)

    tuned = None
    cfg_kw = {}
    if getattr(args, "autotune", 1):
        # BC timed from 32 sources (one sweep), PageRank on the updated graph
        bsrc = np.random.default_rng(seed + 200).choice(V, 32, replace=False)
        tuned = tune(V, False, build, batched, bc_sources=bsrc)
        cfg_kw = overrides(tuned["best"])

    def fresh():
        h = GraphVine(V, directed=False, config=default_config(**cfg_kw))
        build(h)
        h.sync()
        return h

    forms = {}
    # plain
    h = fresh()

    def plain():
        for b in range(nb):
            o, n = int(offs[b]), int(sizes[b])
            vs, vd = cs.view(o, n), cd.view(o, n)
            if kinds[b] == GV_OP_INSERT:
                _check(lib().gv_batch_insert(h.handle, vs.ptr, vd.ptr, n, GV_MEM_DEVICE, None))
            else:
                _check(lib().gv_batch_delete(h.handle, vs.ptr, vd.ptr, n, GV_MEM_DEVICE, None))
                _check(lib().gv_compact(h.handle, 0.25, None))
    _, ms = _timed(h, plain)
    h.sync()
    forms["plain"] = {"ms": ms, "edge_updates_per_s": total / (ms / 1e3)}
    ref_off, ref_cols = h.export_csr()
    E_end = h.num_edges()
    h.close()
    # batched
    h = fresh()
    _, ms = _timed(h, lambda: batched(h))
    h.sync()
    forms["batched"] = {"ms": ms, "edge_updates_per_s": total / (ms / 1e3)}
    o2, c2 = h.export_csr()
    eq_batched = bool(np.array_equal(o2, ref_off) and np.array_equal(c2, ref_cols))
    del o2, c2
    h.close()
    # captured
    h = fresh()
    h.reserve(total, nb)
    with h.capture() as cap:
        batched(h)
    c = cap.result
    _, ms = _timed(h, c.launch)
    h.sync()
    forms["captured"] = {"ms": ms, "edge_updates_per_s": total / (ms / 1e3)}
    o3, c3 = h.export_csr()
    eq_captured = bool(np.array_equal(o3, ref_off) and np.array_equal(c3, ref_cols))
    del o3, c3
    c.close()
    h.close()
    # device queue: fills of uneven batch counts, one capture for all
    cuts = sorted({0, nb} | {max(1, min(nb - 1, int(round(nb * f)))) for f in (0.15, 0.45, 0.7)})
    fills = list(zip(cuts[:-1], cuts[1:]))
    qb = max(b1 - b0 for b0, b1 in fills)
    qp = max(int(offs[b1] - offs[b0]) for b0, b1 in fills)
    fill_dev = []
    for b0, b1 in fills:
        fo = np.full(qb + 1, offs[b1] - offs[b0], np.uint64)
        fo[: b1 - b0 + 1] = offs[b0:b1 + 1] - offs[b0]
        fk = np.full(qb, GV_OP_INSERT, np.int32)
        fk[: b1 - b0] = kinds[b0:b1]
        fill_dev.append((cs.view(int(offs[b0]), int(offs[b1] - offs[b0])),
                         cd.view(int(offs[b0]), int(offs[b1] - offs[b0])),
                         DeviceArray.from_numpy(fo), DeviceArray.from_numpy(fk)))
    qs, qd = DeviceArray(qp, np.uint32), DeviceArray(qp, np.uint32)
    qo, qk = DeviceArray(qb + 1, np.uint64), DeviceArray(qb, np.int32)
    h = fresh()
    h.reserve(qp, qb)
    with h.capture() as cap:
        h.applyBatchesDevice(qs, qd, qo, qk, compact_fraction=0.25)
    c = cap.result
    stream = h.stream()

    def drain():
        for vs, vd, fo, fk in fill_dev:
            qs.copy_from_async(vs, stream)
            qd.copy_from_async(vd, stream)
            qo.copy_from_async(fo, stream)
            qk.copy_from_async(fk, stream)
            c.launch()
    _, ms = _timed(h, drain)
    h.sync()
    forms["queue"] = {"ms": ms, "edge_updates_per_s": total / (ms / 1e3),
                      "fills": [b1 - b0 for b0, b1 in fills], "queue_batches": qb,
                      "queue_pairs": qp}
    o4, c4 = h.export_csr()
    eq_queue = bool(np.array_equal(o4, ref_off) and np.array_equal(c4, ref_cols))
    del o4, c4, ref_off, ref_cols
    c.close()
    for a in (qs, qd, qo, qk):
        a.free()
    for _, _, fo, fk in fill_dev:
        fo.free()
        fk.free()

    # analytics on the updated graph
    rng = np.random.default_rng(seed + 100)
    E = h.num_edges()
    seeds = DeviceArray.from_numpy(rng.choice(V, 64, replace=False).astype(np.uint32))
    X = DeviceArray(V * 64, np.float32)
    it, res = ctypes.c_int(), ctypes.c_double()
    _, ppr_ms = _timed(h, lambda: _check(lib().gv_pagerank_personalized(
        h.handle, 0.85, 1e-6, 100, seeds.ptr, 64, X.ptr, GV_MEM_DEVICE, ctypes.byref(it),
        ctypes.byref(res))))
    X.free()
    srcs = DeviceArray.from_numpy(rng.choice(V, 256, replace=False).astype(np.uint32))
    bco = DeviceArray(V, np.float32)
    # warm-up of the same kind (one batch of 32 sources): the source-batch
    # workspace is allocated once and kept with the handle
    _check(lib().gv_bc(h.handle, srcs.ptr, 32, bco.ptr, GV_MEM_DEVICE))
    h.sync()
    _, bc_ms = _timed(h, lambda: _check(lib().gv_bc(h.handle, srcs.ptr, 256, bco.ptr,
                                                      GV_MEM_DEVICE)))
    bco.free()
    h.close()

    cap_rate = forms["captured"]["edge_updates_per_s"]
    return {
        "metric": "edge updates/s (Chung-Lu stream of uneven batches, captured batched form)",
        "value": cap_rate,
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": forms["captured"]["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (updates), f32 (personalized PageRank), f64 sigma + f32 delta (BC)",
        "data": "synthetic (device-generated Chung-Lu graph and update batches)",
        "config": {
            "workload": f"Chung-Lu V={V}, gamma {gamma}, avg degree {avg}, cap {mx:.0f}, "
                        f"undirected, seed {seed}; {nb} batches log-uniform in [1k, 1M], "
                        "70% insert / 30% delete; compaction at 25% tombstones",
            "vertices": V,
            "directed_edges_initial": int(E0),
            "directed_edges_final": int(E_end),
            "updates_total": total,
            "block_slots": cfg_kw.get("block_slots", 64),
            "tuned": cfg_kw or None,
        },
        "forms": forms,
        "speedup_batched_vs_plain": forms["batched"]["edge_updates_per_s"]
        / forms["plain"]["edge_updates_per_s"],
        "speedup_captured_vs_plain": cap_rate / forms["plain"]["edge_updates_per_s"],
        "final_graphs_equal": {"batched": eq_batched, "captured": eq_captured,
                               "queue": eq_queue},
        "personalized_pagerank_64": {"ms": ppr_ms, "iters": int(it.value),
                                     "max_residual": float(res.value),
                                     "gteps_columns": 64.0 * E * it.value / (ppr_ms / 1e3) / 1e9},
        "bc_256_sources": {"ms": bc_ms, "batches_of": 32},
        "autotune": tuned,
        "setup_s": round(setup_s, 2),
    }


# --------------------------------------------------------------- serving --
def serve_trace(V, seconds, seed=5, update_rate=2000.0, query_rate=100.0,
                query_mix=(0.995, 0.002, 0.003)):
    """Arrival times and request shapes of the serving trace: updates arrive
    as a Poisson stream (exponential gaps, mean 1/update_rate) with sizes
    log-uniform in [100, 100k] (70% inserts); queries arrive as a second
    Poisson stream split into neighbour lookups (1k vertices), PageRank
    refreshes and single-source BC by query_mix.  Returns a list of
    (arrival_us, kind, size_or_source) sorted by arrival."""
    rng = np.random.default_rng(seed)
    ev = []
    t = 0.0
    while True:
        t += rng.exponential(1e6 / update_rate)
        if t >= seconds * 1e6:
            break
        n = int(np.exp(rng.uniform(np.log(100), np.log(100_000))))
        ev.append((t, "insert" if rng.random() < 0.7 else "delete", n))
    t = 0.0
    while True:
        t += rng.exponential(1e6 / query_rate)
        if t >= seconds * 1e6:
            break
        u = rng.random()
        if u < query_mix[0]:
            ev.append((t, "neighbors", 1000))
        elif u < query_mix[0] + query_mix[1]:
            ev.append((t, "pagerank", 0))
        else:
            ev.append((t, "bc", int(rng.integers(0, V))))
    ev.sort(key=lambda e: e[0])
    return ev


def _pct(x, q):
    return float(np.percentile(np.asarray(x, dtype=np.float64), q)) if len(x) else None


def run_serve(args):
    """Serving form: directed R-MAT (default scale 26, edge factor 16,
    seed 5), device-resident with its in-edge mirror; the trace of
    serve_trace() replayed open-loop through the request loop.  Reports
    p50/p99 latency (arrival -> device work done) overall and per kind, and
    the sustained rate of applied edge updates."""
    from . import GraphVine, export_csr_device, stream_key
    from . import serve as S
    from .hip import DeviceArray

    scale = int(getattr(args, "serve_scale", 26))
    if scale > 24:
        # A run of this workload at scale 26 made a GPU memory fault whose
        # cause is not found yet; until it is, only scales whose runs pass
        # (24, as in the default bench's serving leg) are accepted.
        raise RuntimeError(f"--workload serve at scale {scale} is disabled (an unresolved GPU "
                           "memory fault at scale 26); use --serve-scale 24 or less")
    seconds = float(getattr(args, "serve_seconds", 10.0))
    seed = 5
    V = 1 << scale
    t_setup = time.time()
    g = GraphVine(V, directed=True)
    build_rmat_chunked(g, scale, seed)
    E0 = g.num_edges()
    mix = tuple(getattr(args, "serve_query_mix", (0.995, 0.002, 0.003)))
    ev = serve_trace(V, seconds, seed, query_mix=mix)
    upd = [(i, e) for i, e in enumerate(ev) if e[1] in ("insert", "delete")]
    total = sum(e[2] for _, e in upd)
    us, ud = DeviceArray(max(total, 1), np.uint32), DeviceArray(max(total, 1), np.uint32)
    off, cols, _ = export_csr_device(g)
    from . import _check, lib
    at = 0
    where = {}
    for j, (i, e) in enumerate(upd):
        n = e[2]
        vs, vd = us.view(at, n), ud.view(at, n)
        if e[1] == "insert":
            _check(lib().gv_gen_uniform(g.handle, stream_key(seed, 10_000_000 + j), V, n,
                                        vs.ptr, vd.ptr))
        else:
            _check(lib().gv_gen_sample(g.handle, stream_key(seed, 20_000_000 + j), V, off.ptr,
                                       cols.ptr, n, vs.ptr, vd.ptr))
        where[i] = (at, n)
        at += n
    g.sync()
    off.free()
    cols.free()
    rng = np.random.default_rng(seed + 1)
    reqs, arr = [], []
    for i, (t, kind, x) in enumerate(ev):
        if kind in ("insert", "delete"):
            a, n = where[i]
            f = S.insert if kind == "insert" else S.delete
            reqs.append(f(us.view(a, n), ud.view(a, n)))
        elif kind == "neighbors":
            reqs.append(S.neighbors(rng.integers(0, V, x).astype(np.uint32)))
        elif kind == "pagerank":
            reqs.append(S.pagerank())
        else:
            reqs.append(S.bc(x))
        arr.append(t)
    setup_s = time.time() - t_setup
    with S.Server(g) as srv:
        # one warm-up refresh so the first PageRank of the trace is incremental
        w = S.pagerank()
        srv.wait(srv.submit(w))
        t0 = time.time()
        resp = srv.replay(reqs, arr)
        wall = time.time() - t0
        st = srv.stats()
    E_end = g.num_edges()
    g.close()
    lat, svc = {}, {}
    step_ms = {}
    for (t, kind, _), r in zip(ev, resp):
        lat.setdefault(kind, []).append((r["done_us"] - r["arrival_us"]) / 1e3)
        svc.setdefault(kind, []).append((r["done_us"] - r["start_us"]) / 1e3)
        step_ms[r["step"]] = (r["done_us"] - r["start_us"]) / 1e3
    all_ms = [x for v in lat.values() for x in v]
    first_arrival = min(r["arrival_us"] for r in resp)
    last_done = max(r["done_us"] for r in resp)
    span_s = (last_done - first_arrival) / 1e6
    pr_it = [r["iters"] for (t, k, _), r in zip(ev, resp) if k == "pagerank"]
    return {
        "metric": "request latency p99 (serving: Poisson update arrivals + interleaved queries)",
        "value": _pct(all_ms, 99),
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(reqs),
        "warmup": 1,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (updates, lookups), f32 + f64 residual (PageRank), f64 sigma (BC)",
        "data": "synthetic (device-generated R-MAT, update batches; host-generated arrivals)",
        "config": {
            "workload": f"directed R-MAT scale {scale}, edge factor 16, seed {seed}, in-edge "
                        f"mirror; {seconds:.0f} s trace: updates Poisson 2000/s, sizes "
                        "log-uniform [100, 100k], 70% insert / 30% delete; queries Poisson "
                        f"100/s: {100 * mix[0]:.1f}% 1-hop lookups of 1000 vertices, "
                        f"{100 * mix[1]:.1f}% PageRank refresh (warm start, L1 1e-6), "
                        f"{100 * mix[2]:.1f}% single-source BC",
            "vertices": V,
            "directed_edges_initial": int(E0),
            "directed_edges_final": int(E_end),
        },
        "latency_ms": {k: {"n": len(v), "p50": _pct(v, 50), "p99": _pct(v, 99),
                           "max": max(v)} for k, v in lat.items()},
        "service_ms": {k: {"p50": _pct(v, 50), "p99": _pct(v, 99)} for k, v in svc.items()},
        "device_busy_fraction": sum(step_ms.values()) / 1e3 / span_s,
        "p50_ms": _pct(all_ms, 50),
        "p99_ms": _pct(all_ms, 99),
        "sustained_edge_updates_per_s": st["update_pairs_applied"] / span_s,
        "offered_edge_updates": int(total),
        "trace_span_s": span_s,
        "replay_wall_s": wall,
        "steps_served": int(st["steps"]),
        "refused": int(st["refused"]),
        "pagerank_refresh_iters_mean": float(np.mean(pr_it)) if pr_it else None,
        "setup_s": round(setup_s, 2),
    }


# ------------------------------------------------------------- precision --
def build_rmat_chunked(g, scale, seed, edge_factor=16, chunk=1 << 27):
    """Bulk-build edges 0 .. edge_factor*2^scale-1 of the R-MAT stream of
    `seed` into g in chunks (the staging buffers stay chunk-sized).  Output i
    of a stream keyed k is mix(k + (i+1)*GOLD), so shifting the key by
    c0*scale*GOLD continues the stream at edge c0."""
    from . import gen_rmat, stream_key
    n_raw = edge_factor << scale
    for c0 in range(0, n_raw, chunk):
        n = min(chunk, n_raw - c0)
        key = (stream_key(seed, STREAM_RMAT) + c0 * scale * 0x9E3779B97F4A7C15) % (1 << 64)
        src, dst = gen_rmat(g, key, scale, n)
        (g.bulkBuild if c0 == 0 else g.batchInsert)(src, dst, stats=False)
        g.sync()
        src.free()
        dst.free()


def run_precision(args):
    """Precision form, throughput side: directed R-MAT (default scale 24,
    seed 3).  PageRank with per-vertex Kahan sums and an FP64 residual to L1
    1e-8 (at most 300 updates), then plain FP32 sums for the same number of
    updates; BC from 64 sources (FP64 sigma, FP32 compensated
    dependencies).  The errors against FP64 references are measured by
    tests/precision_report.py."""
    from . import GV_MEM_DEVICE, GraphVine, _check, lib, pagerank
    from .hip import DeviceArray

    scale = int(getattr(args, "prec_scale", 24))
    V = 1 << scale
    seed = 3
    t0 = time.time()
    g = GraphVine(V, directed=True)
    build_rmat_chunked(g, scale, seed)
    E = g.num_edges()
    setup_s = time.time() - t0
    for comp in (True, False):
        pagerank(g, epsilon=1e-8, max_iter=2, compensated=comp)  # warm-up
    xc, itc, rc, stc = pagerank(g, epsilon=1e-8, max_iter=300, compensated=True, stats=True)
    xp, itp, rp, stp = pagerank(g, epsilon=0.0, max_iter=itc, compensated=False, stats=True)
    comp_ms, plain_ms = stc["kernel_ms"] / itc, stp["kernel_ms"] / itp
    rng = np.random.default_rng(seed)
    srcs = DeviceArray.from_numpy(rng.choice(V, 64, replace=False).astype(np.uint32))
    out = DeviceArray(V, np.float32)
    _, bc_ms = _timed(g, lambda: _check(lib().gv_bc(g.handle, srcs.ptr, 64, out.ptr,
                                                     GV_MEM_DEVICE)))
    g.close()
    gteps = E / (comp_ms / 1e3) / 1e9
    return {
        "metric": "PageRank GTEPS (compensated FP32 sums, FP64 residual)",
        "value": gteps,
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": itc,
        "warmup": 1,
        "ms_per_step": comp_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 + Kahan (PageRank), f64 residual; f64 sigma + f32 Kahan delta (BC)",
        "data": "synthetic (device-generated R-MAT)",
        "config": {"workload": f"directed R-MAT scale {scale}, edge factor 16, seed {seed}; "
                               "PageRank to L1 1e-8 (max 300 updates), BC from 64 sources",
                   "vertices": V, "edges": int(E)},
        "pagerank": {"updates": itc, "final_residual": rc,
                     "compensated_ms_per_update": comp_ms, "plain_ms_per_update": plain_ms,
                     "plain_gteps": E / (plain_ms / 1e3) / 1e9,
                     "compensated_cost_vs_plain": comp_ms / plain_ms},
        "bc_64_sources": {"ms": bc_ms, "per_source_ms": bc_ms / 64},
        "setup_s": round(setup_s, 2),
    }
