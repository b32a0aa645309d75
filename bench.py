"""Benchmark: plain-form dynamic graph updates on undirected R-MAT scale 22.

Workload (one GPU): undirected R-MAT scale 22 (4,194,304 vertices), edge
factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 2, symmetrised (~134M directed
slots).  One step = one update batch of 1,000,000 undirected edges: 500,000
deletes sampled from existing edges, then 500,000 uniform-random inserts,
then compaction of vertices whose tombstones exceed 25% of their slots, each
issued as its own launch sequence (the plain form).  After every timed batch
PageRank (FP32, alpha 0.85, L1 residual 1e-6) and an exact triangle count run
on the updated graph and are timed separately.

All inputs are generated on the device before the timed region (synthetic,
counter-based SplitMix64 streams).  value = edge updates/s over the K timed
batches (updates = undirected edge updates requested).

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
COPY_CALIB_BYTES = 256 << 20


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--scale", type=int, default=22)
    p.add_argument("--edge-factor", type=int, default=16)
    p.add_argument("--batch", type=int, default=1_000_000)
    p.add_argument("--seed", type=int, default=2)
    p.add_argument("--analytics-every", type=int, default=1,
                   help="run PageRank + TC after every k-th timed batch (0 = never)")
    p.add_argument("--no-tc", action="store_true")
    p.add_argument("--no-parity", action="store_true",
                   help="skip the full-size neighbour-set comparison with the oracle")
    p.add_argument("--cpu-baseline", type=int, default=1, help="1 = time the CPU baseline")
    p.add_argument("--cpu-batches", type=int, default=2)
    p.add_argument("--batched", type=int, default=1, help="1 = also time the batched form")
    p.add_argument("--captured", type=int, default=1,
                   help="1 = also time the captured update -> compact -> PageRank pipeline")
    p.add_argument("--pr-layout", type=int, default=5,
                   help="PageRank layout: 5 = degree-relabelled snapshot, hot stream + "
                        "propagation-blocked cold stream (default), 1 = the same snapshot as "
                        "one flat stream, 4 = the same snapshot row by row")
    p.add_argument("--bc-bu-ratio", type=float, default=None,
                   help="BC: bottom-up when a level's candidate in-slots are below this "
                        "multiple of its frontier's out-slots (default: the library's)")
    p.add_argument("--workload", default="rmat22",
                   choices=["rmat22", "chunglu", "precision", "serve"],
                   help="rmat22: plain vs batched update stream on undirected R-MAT (default); "
                        "chunglu: 256 uneven batches on a Chung-Lu graph (plain, batched, "
                        "captured), 64-column personalized PageRank, BC from 256 sources; "
                        "precision: compensated vs plain PageRank and BC throughput on directed "
                        "R-MAT scale 24; serve: request loop with Poisson arrivals")
    p.add_argument("--autotune", type=int, default=1,
                   help="chunglu: 1 = sweep block size and PageRank bins first")
    p.add_argument("--cl-vertices", type=int, default=8_000_000)
    p.add_argument("--cl-batches", type=int, default=256)
    p.add_argument("--prec-scale", type=int, default=24)
    p.add_argument("--serve-scale", type=int, default=24,
                   help="serve workload R-MAT scale (26, the configuration's size, is disabled "
                        "until a GPU memory fault seen there is resolved)")
    p.add_argument("--serve-seconds", type=float, default=20.0)
    p.add_argument("--summarize-pmc", nargs=4,
                   metavar=("FETCH_DB", "WRITE_DB", "BENCH_JSON", "OUT_JSON"),
                   help="turn two rocprofv3 --pmc result databases (FETCH_SIZE, WRITE_SIZE) of "
                        "this bench, run with the same flags as given here, into the "
                        "per-iteration HBM traffic summary of the PageRank iteration, "
                        "calibrated on the bench's 256 MiB copy (BENCH_JSON: the JSON line "
                        "that run printed, for its final edge count), and exit")
    p.add_argument("--serve-leg", type=int, default=1,
                   help="rmat22: 1 = also run a bounded serving leg (request loop, p50/p99)")
    p.add_argument("--serve-leg-scale", type=int, default=24)
    p.add_argument("--serve-leg-seconds", type=float, default=8.0)
    return p.parse_args()


def skey(seed, stream):
    """Stream key of the counter-based generators: mix(mix(seed) ^ stream*C)."""
    M = (1 << 64) - 1

    def fmix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    return fmix(fmix(seed) ^ ((stream * 0xD6E8FEB86659FD93) & M))


def main():
    args = parse()
    if args.summarize_pmc:
        summarize_pmc(args, *args.summarize_pmc)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported by this build "
              "(the multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "chunglu":
        from gv_amd.workloads import run_chunglu
        out = run_chunglu(args)
    elif args.workload == "precision":
        from gv_amd.workloads import run_precision
        out = run_precision(args)
    elif args.workload == "serve":
        from gv_amd.workloads import run_serve
        out = run_serve(args)
    else:
        out = run_rmat(args)
    print(json.dumps(out), flush=True)


def run_rmat(args):
    import gv_amd as G
    from gv_amd.hip import DeviceArray, Event, synchronize

    STREAM_RMAT, STREAM_INS, STREAM_DEL = 1, 2, 3

    V = 1 << args.scale
    n_raw = args.edge_factor * V
    half = args.batch // 2
    t_setup = time.time()
    kw = {} if args.bc_bu_ratio is None else {"bc_bu_ratio": args.bc_bu_ratio}
    g = G.GraphVine(V, directed=False, pr_layout=args.pr_layout, **kw)
    stream = g.stream()
    src, dst = G.gen_rmat(g, skey(args.seed, STREAM_RMAT), args.scale, n_raw)
    st = g.bulkBuild(src, dst)
    src.free()
    dst.free()
    E0 = st["live_edges"]
    off, cols, E = G.export_csr_device(g)
    nb = args.warmup + args.steps
    batches = []
    for b in range(nb):
        ds, dd = G.gen_sample(g, skey(args.seed, STREAM_DEL * 1000 + b), V, off, cols, half)
        is_, id_ = G.gen_uniform(g, skey(args.seed, STREAM_INS * 1000 + b), V, half)
        batches.append((ds, dd, is_, id_))
    off.free()
    cols.free()
    synchronize()
    setup_s = time.time() - t_setup

    pr_out = DeviceArray(V, np.float32)
    # degrees and batch sources before the timed batches, for the update
    # path's algorithmic bytes (touched adjacency read by the probes)
    deg0 = None

    def run_batch(b):
        ds, dd, is_, id_ = batches[b]
        g.batchDelete(ds, dd, stats=False)
        g.batchInsert(is_, id_, stats=False)
        G.lib().gv_compact(g.handle, 0.25, None)

    for b in range(args.warmup):
        run_batch(b)
    g.sync()
    synchronize()
    deg0 = g.degrees().astype(np.int64)

    # Plain device-to-device copy of a known size (the runtime's copy kernel):
    # the calibration point for HBM byte counters collected with --pmc.
    calib = DeviceArray(COPY_CALIB_BYTES // 4, np.uint32)
    calib2 = DeviceArray(COPY_CALIB_BYTES // 4, np.uint32)
    from gv_amd.hip import D2D, hip
    hip().hipMemcpy(calib2.ptr, calib.ptr, COPY_CALIB_BYTES, D2D)
    synchronize()
    calib.free()
    calib2.free()

    upd_ms = []
    pr_ms, pr_kernel_ms, pr_iters, tc_ms, tc_total = [], [], [], [], []
    ev0, ev1 = Event(), Event()
    for k in range(args.steps):
        b = args.warmup + k
        ev0.record(stream)
        run_batch(b)
        ev1.record(stream)
        ev1.synchronize()
        g.sync()  # raises if the device rejected the batch
        upd_ms.append(ev0.elapsed_ms(ev1))
        if args.analytics_every and (k + 1) % args.analytics_every == 0:
            ev0.record(stream)
            prs = G.pagerank_device(g, pr_out, 0.85, 1e-6, 100)
            ev1.record(stream)
            ev1.synchronize()
            pr_ms.append(ev0.elapsed_ms(ev1))
            pr_kernel_ms.append(prs["kernel_ms"])
            pr_iters.append(prs["iters"])
            pr_edges = prs["edges"]
            if not args.no_tc:
                ev0.record(stream)
                _, tot = G.static_tc(g, per_vertex=False)
                ev1.record(stream)
                ev1.synchronize()
                tc_ms.append(ev0.elapsed_ms(ev1))
                tc_total.append(tot)
    E_end = g.num_edges()
    tot_upd_ms = float(sum(upd_ms))
    value = args.batch * args.steps / (tot_upd_ms / 1e3)

    out = {
        "metric": "edge updates/s (plain form, 1M-edge batches) + PageRank GTEPS",
        "value": value,
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tot_upd_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32 (updates), f32+f64 residual (PageRank), u64 (TC)",
        "data": "synthetic (device-generated R-MAT and update batches)",
        "config": {
            "workload": f"plain update stream on undirected R-MAT scale {args.scale}, "
                        f"edge factor {args.edge_factor}, seed {args.seed}; "
                        f"{args.batch} edges/batch (50% delete sampled, 50% uniform insert), "
                        "compaction at 25% tombstones",
            "vertices": V,
            "directed_edges_initial": int(E0),
            "directed_edges_final": int(E_end),
            "batch_edges": args.batch,
        },
        "update_ms_per_batch": [round(x, 4) for x in upd_ms],
        "setup_s": round(setup_s, 2),
    }
    if pr_ms:
        it = float(np.mean(pr_iters))
        k_ms = float(np.mean(pr_kernel_ms))
        per_iter_ms = k_ms / it
        bytes_per_iter = 8.0 * pr_edges + 16.0 * V
        achieved = bytes_per_iter / (per_iter_ms / 1e3) / 1e9
        out["pagerank"] = {
            "iters_mean": it,
            "call_ms_mean": float(np.mean(pr_ms)),
            "kernel_ms_per_iter": per_iter_ms,
            "gteps": pr_edges / (per_iter_ms / 1e3) / 1e9,
            "gteps_whole_call": pr_edges * it / (float(np.mean(pr_ms)) / 1e3) / 1e9,
        }
        traffic, tsrc = load_traffic(args, E_end)
        out["roofline"] = {
            "kernel": ("PageRank iteration, propagation-blocked layout (k_pb_tiles + k_pb_rows "
                       "+ k_pb_expand)" if args.pr_layout == 5 else
                       "PageRank flat-stream iteration (k_pf_tiles + k_pf_rows + k_pf_resid)"
                       if args.pr_layout == 1 else f"PageRank iteration, layout {args.pr_layout}")
                      + ", timed with HIP events on the graph's stream around the counted "
                        "iterations (launch gaps included)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": tsrc,
            "algorithmic_bytes_per_launch": bytes_per_iter,
        }
    if tc_ms:
        out["triangle_count"] = {"ms_mean": float(np.mean(tc_ms)), "total_last": int(tc_total[-1])}
    out["rooflines"] = other_rooflines(g, G, V, args, batches, deg0, upd_ms, tc_ms)
    if pr_ms:
        it = float(np.mean(pr_iters))
        out["rooflines"]["pagerank_whole_call"] = _roof(
            it * (8.0 * pr_edges + 16.0 * V), float(np.mean(pr_ms)),
            "whole PageRank call (snapshot of the live in-adjacency + iterations + host "
            "waits), HIP events on the graph's stream; bytes = iterations x (8 E + 16 V)")

    # the plain run's final graph (sorted CSR), the reference for every
    # full-size check below and for the batched form's final graph
    final_csr = start_csr = None
    if not args.no_parity:
        par, final_csr, start_csr = full_size_parity(g, G, V, args, batches, n_raw,
                                                     pr_out if pr_ms else None,
                                                     pr_iters[-1] if pr_ms else 0)
        out["parity_full_size"] = par
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, V, args, batches, final_csr, start_csr)
    del start_csr
    out["h2d_inclusive"] = h2d_inclusive(g, G, V, args)
    if not args.no_tc:
        out["tc_update"] = tc_update(g, G, V, args)
    pr_final = pr_out.numpy() if pr_ms else None
    g.close()
    if args.batched:
        out["batched"] = run_batched(args, G, skey, V, n_raw, batches, E_end, value, final_csr)
    if args.captured and pr_ms:
        out["captured_pipeline"] = run_captured(args, G, skey, V, n_raw, batches, final_csr,
                                                pr_final, upd_ms, pr_ms)
    del final_csr
    if args.serve_leg:
        out["serving"] = serve_leg(args)
    return out


def serve_leg(args):
    """Bounded serving leg of the default run: the request loop on directed
    R-MAT (scale --serve-leg-scale, default 24) over a --serve-leg-seconds
    trace of Poisson update arrivals (2000/s, sizes log-uniform [100, 100k],
    70/30) and queries (100/s) with a heavy share that exercises the
    analytics: 97.5% 1-hop lookups of 1000 vertices, 1% warm-started PageRank
    refreshes, 1.5% single-source BC.  p50/p99 latency per request kind."""
    from gv_amd.workloads import run_serve

    class A:
        serve_scale = args.serve_leg_scale
        serve_seconds = args.serve_leg_seconds
        serve_query_mix = (0.975, 0.01, 0.015)
    r = run_serve(A)
    keep = ("config", "latency_ms", "service_ms", "p50_ms", "p99_ms",
            "sustained_edge_updates_per_s", "offered_edge_updates", "trace_span_s",
            "steps_served", "refused", "pagerank_refresh_iters_mean", "device_busy_fraction",
            "setup_s")
    return {k: r[k] for k in keep if k in r}


def _roof(nbytes, ms, what):
    gbs = nbytes / (ms / 1e3) / 1e9
    return {"what": what, "algorithmic_bytes": nbytes, "ms": ms, "achieved": gbs,
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}


def other_rooflines(g, G, V, args, batches, deg0, upd_ms, tc_ms):
    """Algorithmic-byte rooflines of the update path, triangle counting and
    BC on the same graph (secondary to the PageRank line):
      update batch: 12 k + sum over touched sources of 4 d(u) + 16 |T| for
        each half (k directed entries), with d from the degrees before the
        timed batches;
      TC: sum over oriented edges (u, v) of 4 (d+(u) + d+(v));
      BC: 24 E + 24 V per source, 16 sources."""
    out = {}
    nb = 0
    for b in range(args.warmup, args.warmup + len(upd_ms)):
        for s_arr, d_arr in (batches[b][0:2], batches[b][2:4]):
            su, sd = s_arr.numpy().astype(np.int64), d_arr.numpy().astype(np.int64)
            src = np.concatenate([su, sd])  # undirected: both directions
            t = np.unique(src)
            nb += 12 * len(src) + 4 * int(deg0[t].sum()) + 16 * len(t)
    out["update_batch"] = _roof(nb / len(upd_ms), float(np.mean(upd_ms)),
                                "plain update batch (delete half, insert half, compaction)")
    off, cols = g.export_csr()
    deg = np.diff(off).astype(np.int64)
    if tc_ms:
        u = np.repeat(np.arange(V, dtype=np.int64), deg)
        w = cols.astype(np.int64)
        keep = (deg[w] > deg[u]) | ((deg[w] == deg[u]) & (w > u))
        u, w = u[keep], w[keep]
        dplus = np.bincount(u, minlength=V).astype(np.int64)
        nbytes = 4 * int((dplus[u] + dplus[w]).sum())
        out["triangle_count"] = _roof(nbytes, float(tc_ms[-1]), "exact triangle count, last graph")
        del u, w, keep
    E = int(off[-1])
    from gv_amd.hip import DeviceArray, Event
    srcs = np.random.default_rng(args.seed).integers(0, V, 16).astype(np.uint32)
    s_dev = DeviceArray.from_numpy(srcs)
    bc_out = DeviceArray(V, np.float32)
    L = G.lib()
    # warm-up call of the same kind: the batched workspace is allocated once
    G._check(L.gv_bc(g.handle, s_dev.ptr, len(srcs), bc_out.ptr, G.GV_MEM_DEVICE))
    stream = g.stream()
    ev0, ev1 = Event(), Event()
    ev0.record(stream)
    G._check(L.gv_bc(g.handle, s_dev.ptr, len(srcs), bc_out.ptr, G.GV_MEM_DEVICE))
    ev1.record(stream)
    ev1.synchronize()
    bc_ms = ev0.elapsed_ms(ev1)
    s_dev.free()
    bc_out.free()
    out["bc_16_sources"] = _roof(16 * (24 * E + 24 * V), bc_ms,
                                 "Brandes BC from 16 sources (one batch of the 32-source "
                                 "kernels), HIP events on the graph's stream around the call "
                                 "(host round trips between levels included)")
    # sorted CSR export of the whole graph into device memory
    off_d, cols_d = DeviceArray(V + 1, np.uint64), DeviceArray(max(E, 1), np.uint32)
    G._check(L.gv_export_csr(g.handle, off_d.ptr, cols_d.ptr, G.GV_MEM_DEVICE))
    t0 = time.perf_counter()
    G._check(L.gv_export_csr(g.handle, off_d.ptr, cols_d.ptr, G.GV_MEM_DEVICE))
    ex_ms = (time.perf_counter() - t0) * 1e3
    off_d.free()
    cols_d.free()
    out["export_csr"] = {"ms": ex_ms, "directed_edges": E,
                         "what": "gv_export_csr of the whole graph into device memory (item-"
                                 "balanced key emission, radix sort of the keys, columns), host "
                                 "wall time of the call including its workspace allocations"}
    return out


def full_size_parity(g, G, V, args, batches, n_raw, pr_x, pr_iters):
    """Parity gate on the benchmark graph itself, after the timed batches.
    The sequential oracle is built from the host R-MAT generator (the same
    stream as the device one) and given the same warm-up and timed batches;
    degrees and every sorted neighbour set must equal the device store's.
    On that CSR the FP64 OpenMP PageRank (pinned to the oracle) runs as many
    updates as the device's last PageRank call did, and the OpenMP triangle
    count (pinned to the oracle per vertex) is compared with a device count
    per vertex and in total.  Returns (summary, sorted CSR of the final
    graph, sorted CSR of the graph the timed batches start from)."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    s, d = O.gen_rmat(skey(args.seed, 1), args.scale, n_raw)
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    del s, d
    start_csr = None
    for b in range(args.warmup + args.steps):
        if b == args.warmup:
            start_csr = o.export_csr()  # the graph the timed batches start from
        ds, dd, is_, id_ = batches[b]
        o.delete(ds.numpy(), dd.numpy())
        o.insert(is_.numpy(), id_.numpy())
    goff, gcols = g.export_csr()
    ooff, ocols = o.export_csr()
    o.close()
    deg_eq = bool(np.array_equal(goff, ooff))
    sets_eq = deg_eq and bool(np.array_equal(gcols, ocols))
    del goff, gcols
    res = {"degrees_equal": deg_eq, "neighbour_sets_equal": sets_eq,
           "directed_edges": int(ooff[-1]),
           "what": "sequential oracle built from the same R-MAT stream and the same "
                   f"{args.warmup + args.steps} batches, compared with the device store's "
                   "sorted export after the timed batches; PageRank and triangle counts of "
                   "that graph against the OpenMP FP64 / exact references on the oracle's CSR"}
    if pr_x is not None:
        x = pr_x.numpy().astype(np.float64)
        xo, ito, _ = O.omp_pagerank_sym(V, ooff, ocols, fixed_iters=pr_iters, residual=True)
        err = np.abs(x - xo)
        res["pagerank"] = {"iters": int(pr_iters), "l1": float(err.sum()),
                           "linf": float(err.max()), "l1_bound": 1e-6,
                           "within_bound": bool(err.sum() <= 1e-6),
                           "reference": "oracle/gv_omp.c omp_pagerank_sym (FP64, x0 = 1/V) at "
                                        "the device's iteration count"}
    if not args.no_tc:
        pv, tot = G.static_tc(g, per_vertex=True)
        t1 = time.perf_counter()
        pvo, toto = O.omp_triangle_count(V, ooff, ocols)
        res["triangles"] = {"total": int(tot), "total_reference": int(toto),
                            "per_vertex_equal": bool(np.array_equal(pv, pvo)),
                            "equal": bool(tot == toto and np.array_equal(pv, pvo)),
                            "reference_s": round(time.perf_counter() - t1, 2),
                            "reference": "oracle/gv_omp.c omp_triangle_count (exact, min-ID "
                                         "per vertex)"}
    res["check_s"] = round(time.perf_counter() - t0, 1)
    return res, (ooff, ocols), start_csr


def h2d_inclusive(g, G, V, args):
    """The plain form with the batch handed over in host memory (the boundary's
    GV_MEM_HOST path: each call copies its pairs host->device on the graph's
    stream), host-timed per batch around delete half + insert half +
    compaction.  Fresh batches of the same kind (deletes sampled from the
    current graph, uniform inserts), generated on the device and copied to
    host arrays before timing."""
    STREAM_INS, STREAM_DEL = 2, 3
    nb = 3
    half = args.batch // 2
    off, cols, _ = G.export_csr_device(g)
    host = []
    for b in range(nb):
        ds, dd = G.gen_sample(g, skey(args.seed, STREAM_DEL * 1000 + 900 + b), V, off, cols, half)
        is_, id_ = G.gen_uniform(g, skey(args.seed, STREAM_INS * 1000 + 900 + b), V, half)
        host.append((ds.numpy(), dd.numpy(), is_.numpy(), id_.numpy()))
        for a in (ds, dd, is_, id_):
            a.free()
    off.free()
    cols.free()
    g.sync()
    ms = []
    for ds, dd, is_, id_ in host:
        t0 = time.perf_counter()
        g.batchDelete(ds, dd, stats=False)
        g.batchInsert(is_, id_, stats=False)
        G.lib().gv_compact(g.handle, 0.25, None)
        g.sync()
        ms.append((time.perf_counter() - t0) * 1e3)
    per = float(np.mean(ms))
    return {"value": args.batch / (per / 1e3), "unit": "edge updates/s", "ms_per_batch": per,
            "batches": nb,
            "what": "fresh 1M-edge batches of the same kind passed as host arrays through the "
                    "C ABI (GV_MEM_HOST, pageable memory), host-timed; deletes sampled from the "
                    "graph as it stands after the timed batches"}


def tc_update(g, G, V, args):
    """Exact triangle-count update (gv_triangle_count_update) over one more
    batch of the same kind (a delete half sampled from the current graph, then
    a uniform insert half), per-vertex counts kept on the device, against the
    full recount; the updated counts are checked against a recount after."""
    from gv_amd.hip import DeviceArray, Event
    STREAM_INS, STREAM_DEL = 2, 3
    half = args.batch // 2
    off, cols, _ = G.export_csr_device(g)
    ds, dd = G.gen_sample(g, skey(args.seed, STREAM_DEL * 1000 + 999), V, off, cols, half)
    is_, id_ = G.gen_uniform(g, skey(args.seed, STREAM_INS * 1000 + 999), V, half)
    off.free()
    cols.free()
    pv = DeviceArray(V, np.uint64)
    tot = ctypes.c_uint64()
    L = G.lib()
    stream = g.stream()
    ev0, ev1 = Event(), Event()
    ev0.record(stream)
    G._check(L.gv_triangle_count(g.handle, pv.ptr, G.GV_MEM_DEVICE, ctypes.byref(tot)))
    ev1.record(stream)
    ev1.synchronize()
    res = {"full_recount_ms": ev0.elapsed_ms(ev1)}
    for name, s_arr, d_arr, kind in (("delete", ds, dd, G.GV_OP_DELETE),
                                     ("insert", is_, id_, G.GV_OP_INSERT)):
        ev0.record(stream)
        G._check(L.gv_triangle_count_update(g.handle, kind, s_arr.ptr, d_arr.ptr, half,
                                            G.GV_MEM_DEVICE, pv.ptr, G.GV_MEM_DEVICE,
                                            ctypes.byref(tot), None))
        ev1.record(stream)
        ev1.synchronize()
        res[f"{name}_half_ms"] = ev0.elapsed_ms(ev1)
    # the updated counts against the OpenMP exact count of the graph after
    # the batch (its sorted export), per vertex and in total
    from oracle import oracle as O
    goff, gcols = g.export_csr()
    pvo, toto = O.omp_triangle_count(V, goff, gcols)
    del goff, gcols
    res["exact"] = bool(tot.value == toto and np.array_equal(pv.numpy(), pvo))
    res["reference"] = "oracle/gv_omp.c omp_triangle_count on the graph after the batch"
    res["speedup_vs_recount"] = res["full_recount_ms"] / (res["delete_half_ms"]
                                                          + res["insert_half_ms"])
    res["what"] = (f"one {args.batch}-edge batch (delete half, insert half), each applied with "
                   "its triangle-count update; times include the update itself and the "
                   "host round trips of the call")
    for a in (ds, dd, is_, id_, pv):
        a.free()
    return res


def run_batched(args, G, skey, V, n_raw, batches, E_plain_end, plain_value, final_csr=None):
    """Batched form: the same K timed batches (each a delete half then an
    insert half) applied as one group by gv_batch_apply, followed by one
    compaction at 25% tombstones, on a fresh graph brought to the same state
    as the plain run's starting point."""
    from gv_amd.hip import D2D, DeviceArray, Event, hip, synchronize
    g2 = G.GraphVine(V, directed=False)
    src, dst = G.gen_rmat(g2, skey(args.seed, 1), args.scale, n_raw)
    g2.bulkBuild(src, dst)
    src.free()
    dst.free()
    for b in range(args.warmup):
        ds, dd, is_, id_ = batches[b]
        g2.batchDelete(ds, dd, stats=False)
        g2.batchInsert(is_, id_, stats=False)
        G.lib().gv_compact(g2.handle, 0.25, None)
    half = args.batch // 2
    parts = []
    for k in range(args.steps):
        ds, dd, is_, id_ = batches[args.warmup + k]
        parts += [(ds, dd, G.GV_OP_DELETE), (is_, id_, G.GV_OP_INSERT)]
    n = half * len(parts)
    cs, cd = DeviceArray(n, np.uint32), DeviceArray(n, np.uint32)
    offsets = [0]
    for i, (s, d, _) in enumerate(parts):
        o = i * half * 4
        hip().hipMemcpy(ctypes.c_void_p(cs.ptr.value + o), s.ptr, half * 4, D2D)
        hip().hipMemcpy(ctypes.c_void_p(cd.ptr.value + o), d.ptr, half * 4, D2D)
        offsets.append(offsets[-1] + half)
    kinds = [p[2] for p in parts]
    # staging sized once for the group, as the plain form's is by its warm-up
    # batches (a batching caller reserves once; captured pipelines must)
    g2.reserve(n, len(parts))
    g2.sync()
    synchronize()
    stream = g2.stream()
    ev0, ev1 = Event(), Event()
    ev0.record(stream)
    g2.applyBatches(cs, cd, offsets, kinds, compact_fraction=0.25, stats=False)
    ev1.record(stream)
    ev1.synchronize()
    g2.sync()
    ms = ev0.elapsed_ms(ev1)
    E2 = g2.num_edges()
    sets_eq = None
    if final_csr is not None:
        boff, bcols = g2.export_csr()
        sets_eq = bool(np.array_equal(boff, final_csr[0]) and np.array_equal(bcols, final_csr[1]))
        del boff, bcols
    g2.close()
    val = args.batch * args.steps / (ms / 1e3)
    return {
        "value": val,
        "unit": "edge updates/s",
        "ms_total": ms,
        "batches_in_group": len(parts),
        "speedup_vs_plain": val / plain_value,
        "final_edges_equal_plain": bool(E2 == E_plain_end),
        "final_neighbour_sets_equal_plain": sets_eq,
        "form": "one gv_batch_apply over the same timed batches (delete half, insert half "
                "each) + one compaction at 25% tombstones",
    }


def run_captured(args, G, skey, V, n_raw, batches, final_csr, pr_final, upd_ms, pr_ms):
    """Captured pipeline: one HIP graph recording a batch's delete half,
    insert half, compaction at 25% tombstones and a PageRank call (alpha
    0.85, L1 1e-6, up to 100 updates), replayed once per timed batch on a
    fresh graph brought to the plain run's starting state; the batch is
    copied into the capture's input buffers before each replay (outside the
    timed region).  The final graph must equal the plain run's, and the last
    replay's PageRank the plain run's last call (same graph), bit for bit."""
    from gv_amd.hip import D2D, DeviceArray, Event, hip, synchronize
    g3 = G.GraphVine(V, directed=False)
    src, dst = G.gen_rmat(g3, skey(args.seed, 1), args.scale, n_raw)
    g3.bulkBuild(src, dst)
    src.free()
    dst.free()
    for b in range(args.warmup):
        ds, dd, is_, id_ = batches[b]
        g3.batchDelete(ds, dd, stats=False)
        g3.batchInsert(is_, id_, stats=False)
        G.lib().gv_compact(g3.handle, 0.25, None)
    half = args.batch // 2
    bufs = [DeviceArray(half, np.uint32) for _ in range(4)]
    out = DeviceArray(V, np.float32)
    g3.reserve(half)
    # PageRank buffers for the live edges now plus every timed batch inserting
    g3.reserve_pagerank(g3.num_edges() + args.batch * args.steps)
    with g3.capture() as cap:
        g3.batchDelete(bufs[0], bufs[1], stats=False)
        g3.batchInsert(bufs[2], bufs[3], stats=False)
        G._check(G.lib().gv_compact(g3.handle, 0.25, None))
        G.pagerank_device_run(g3, out, alpha=0.85, epsilon=1e-6, max_iter=100, stats=False)
    c = cap.result
    stream = g3.stream()
    ev0, ev1 = Event(), Event()
    ms, iters = [], []
    for k in range(args.steps):
        for buf, arr in zip(bufs, batches[args.warmup + k]):
            hip().hipMemcpy(buf.ptr, arr.ptr, half * 4, D2D)
        synchronize()
        ev0.record(stream)
        c.launch()
        ev1.record(stream)
        ev1.synchronize()
        ms.append(ev0.elapsed_ms(ev1))
        iters.append(g3.pagerank_last()["iters"])
    sets_eq = None
    if final_csr is not None:
        coff, ccols = g3.export_csr()
        sets_eq = bool(np.array_equal(coff, final_csr[0]) and np.array_equal(ccols, final_csr[1]))
        del coff, ccols
    pr_eq = bool(np.array_equal(out.numpy(), pr_final)) if pr_final is not None else None
    c.close()
    g3.close()
    eager = [u + p for u, p in zip(upd_ms, pr_ms)]
    return {
        "ms_per_batch_mean": float(np.mean(ms)),
        "eager_ms_per_batch_mean": float(np.mean(eager)),
        "speedup_vs_eager": float(np.mean(eager)) / float(np.mean(ms)),
        "pagerank_iters_mean": float(np.mean(iters)),
        "replays": len(ms),
        "final_neighbour_sets_equal_plain": sets_eq,
        "last_pagerank_bitwise_equal_plain": pr_eq,
        "form": "one HIP graph per batch: delete half, insert half, compaction at 25% "
                "tombstones, PageRank (snapshot and layout sized on the device, 100 recorded "
                "iterations that exit after convergence); eager = the plain run's update + "
                "PageRank call times of the same batches",
    }


def _median_timed(fn, runs=5, warmup=1, budget_s=None):
    """Median wall time (s) of `runs` calls of fn after `warmup` untimed ones;
    with budget_s, stops early once the timed calls exceed it (at least one).
    Returns (median_s, timed_runs)."""
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
        if budget_s is not None and sum(ts) > budget_s:
            break
    return float(np.median(ts)), len(ts)


def cpu_baseline(g, V, args, batches, final_csr=None, start_csr=None):
    """CPU baselines on this host, on the same graph and the same batches:
      updates: the OpenMP CSR baseline (sorted CSR rebuilt per batch by
        parallel sort-merge) from the graph the timed batches start from,
        over the first timed batches in order: one untimed, then five timed,
        median per batch (without the parity check's CSRs it starts from
        the device graph as it stands);
      analytics (OpenMP, all host threads, on the plain run's final graph):
        pull PageRank per iteration, exact triangle count, Brandes BC from
        the same 16 sources as the device's BC line, each the median of 5
        after 1 warm-up (the triangle count stops early past 30 s);
      sequential oracle (one core): the same update batches, median of 5
        after 1."""
    from oracle import oracle as O
    if final_csr is None:
        final_csr = g.export_csr()
    if start_csr is None:
        start_csr, first = final_csr, 0
    else:
        first = args.warmup
    nbat = min(6, len(batches) - first)
    state = {"off": start_csr[0], "cols": start_csr[1], "b": 0}
    host = [tuple(a.numpy() for a in batches[first + b]) for b in range(nbat)]

    def one_batch():
        sd, dd_h, si, di = host[state["b"] % nbat]
        state["b"] += 1
        off, cols = O.omp_apply_batch(V, False, state["off"], state["cols"], sd, dd_h,
                                      delete=True)
        state["off"], state["cols"] = O.omp_apply_batch(V, False, off, cols, si, di,
                                                        delete=False)

    t_upd, n_upd = _median_timed(one_batch, runs=max(nbat - 1, 1))
    del state
    out = {
        "value": args.batch / t_upd,
        "unit": "edge updates/s",
        "cores": O.omp_threads(),
        "kind": "port",
        "sample": f"1M-edge batches of the same stream (delete half, insert half) applied to "
                  f"the scale-{args.scale} graph by OpenMP sorted-CSR rebuild (parallel radix "
                  f"sort + merge): 1 untimed, then the median of {n_upd}",
        "cpu_model": _cpu_model(),
    }
    off, cols = final_csr
    it = 3
    t_pr, n_pr = _median_timed(lambda: O.omp_pagerank_sym(V, off, cols, fixed_iters=it))
    E = int(off[-1])
    an = {"pagerank_ms_per_iter": t_pr / it * 1e3,
          "pagerank_gteps": E / (t_pr / it) / 1e9,
          "pagerank_runs": n_pr}
    if not args.no_tc:
        t_tc, n_tc = _median_timed(lambda: O.omp_triangle_count(V, off, cols, per_vertex=False),
                                   budget_s=30.0)
        an["triangle_count_ms"] = t_tc * 1e3
        an["triangle_count_runs"] = n_tc
    srcs = np.random.default_rng(args.seed).integers(0, V, 16).astype(np.uint32)
    t_bc, n_bc = _median_timed(lambda: O.omp_bc_dir(V, off, cols, srcs), budget_s=30.0)
    an["bc_16_sources_ms"] = t_bc * 1e3
    an["bc_runs"] = n_bc
    an["what"] = ("OpenMP CSR analytics on the plain run's final graph, all host threads: pull "
                  "PageRank (FP64), exact triangle count (degree-oriented lists), Brandes BC "
                  "from the 16 sources of the device's BC line; median of the runs after 1 "
                  "warm-up")
    out["analytics"] = an
    out["sequential_oracle"] = cpu_sequential(V, args, batches, start_csr[0], start_csr[1], first)
    return out


def cpu_sequential(V, args, batches, off, cols, first):
    """The host-sequential oracle (one core: sorted adjacency arrays, batch
    radix-sorted and merged per touched vertex) loaded with the graph the
    timed batches start from (not timed), then the first timed batches in
    order: 1 untimed, the median of 5."""
    from oracle import oracle as O
    u = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    keep = u < cols
    o = O.OracleGraph(V, False)
    o.bulk_build(u[keep], cols[keep])
    del u, keep
    nbat = min(6, len(batches) - first)
    host = [tuple(a.numpy() for a in batches[first + b]) for b in range(nbat)]
    state = {"b": 0}

    def one_batch():
        sd, dd_h, si, di = host[state["b"] % nbat]
        state["b"] += 1
        o.delete(sd, dd_h)
        o.insert(si, di)

    t, n = _median_timed(one_batch, runs=max(nbat - 1, 1))
    o.close()
    return {"value": args.batch / t, "unit": "edge updates/s", "cores": 1, "kind": "port",
            "sample": f"1M-edge batches of the same stream (delete half, insert half) on the "
                      f"scale-{args.scale} graph, oracle/gv_oracle.c on one core: 1 untimed, "
                      f"then the median of {n}"}


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pagerank_pb_rmat22_traffic.json")
PR_ITER_KERNELS = ("k_pf_tiles", "k_pf_rows", "k_pf_resid", "k_pb_tiles", "k_pb_rows",
                   "k_pb_expand")


def _traffic_key(args, edges_final):
    return {"scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed,
            "pr_layout": args.pr_layout, "steps": args.steps, "warmup": args.warmup,
            "batch": args.batch, "directed_edges_final": int(edges_final)}


def load_traffic(args, edges_final):
    """Measured HBM bytes per PageRank iteration (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, calibrated) for this workload, if a summary made by
    --summarize-pmc of a run of the same shape is committed: same graph,
    layout, warm-up and timed batch counts, batch size and final edge count
    (the graph the last PageRank call ran on).  Anything else is refused."""
    try:
        with open(TRAFFIC_JSON) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, None
    if d.get("workload") != _traffic_key(args, edges_final):
        return None, None
    return d["traffic_bytes_per_iteration"], os.path.relpath(TRAFFIC_JSON, ROOT)


def summarize_pmc(args, fetch_db, write_db, bench_json, out_json):
    import sqlite3
    with open(bench_json) as f:
        line = [x for x in f if x.lstrip().startswith("{")][-1]
    edges_final = json.loads(line)["config"]["directed_edges_final"]

    def per_dispatch(db, counter):
        c = sqlite3.connect(db)
        rows = c.execute("select kernel_name, value from counters_collection "
                         "where counter_name = ?", (counter,)).fetchall()
        c.close()
        return rows

    res = {}
    for db, counter in ((fetch_db, "FETCH_SIZE"), (write_db, "WRITE_SIZE")):
        rows = per_dispatch(db, counter)
        copies = [v for k, v in rows if "copyBuffer" in k]
        if not copies:
            raise SystemExit(f"{db}: no copy dispatch to calibrate {counter} on")
        # the largest copy is the bench's 256 MiB device-to-device calibration copy
        factor = (COPY_CALIB_BYTES / 1024.0) / max(copies)
        kern = {}
        for name in PR_ITER_KERNELS:
            vals = [v for k, v in rows if name in k]
            # launches issued after convergence (the host polls the done flag
            # every few iterations) exit at once; they are not iterations
            vals = [v for v in vals if v >= 0.01 * max(vals)] if vals else vals
            if vals:
                kern[name] = {"dispatches": len(vals), "kib_mean": sum(vals) / len(vals),
                              "bytes_calibrated": sum(vals) / len(vals) * 1024.0 * factor}
        res[counter] = {"calibration_factor": factor, "copy_kib": max(copies), "kernels": kern}
    total = sum(k["bytes_calibrated"] for c in res.values() for k in c["kernels"].values())
    out = {
        "workload": _traffic_key(args, edges_final),
        "traffic_bytes_per_iteration": total,
        "counters": res,
        "made_by": "bench.py [the profiled run's flags] --summarize-pmc FETCH_DB WRITE_DB "
                   "BENCH_JSON OUT_JSON",
    }
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"traffic_bytes_per_iteration": total}))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
