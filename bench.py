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
    p.add_argument("--pr-layout", type=int, default=1,
                   help="PageRank layout: 1 = degree-relabelled CSR snapshot (default), "
                        "3 = snapshot in vertex order, 2 = pull over the store, "
                        "0 = XCD-segmented snapshot")
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
    p.add_argument("--serve-scale", type=int, default=26)
    p.add_argument("--serve-seconds", type=float, default=20.0)
    p.add_argument("--summarize-pmc", nargs=3, metavar=("FETCH_DB", "WRITE_DB", "OUT_JSON"),
                   help="turn two rocprofv3 --pmc result databases (FETCH_SIZE, WRITE_SIZE) of "
                        "this bench into the per-iteration HBM traffic summary of the PageRank "
                        "iteration, calibrated on the bench's 256 MiB copy, and exit")
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
    g = G.GraphVine(V, directed=False, pr_layout=args.pr_layout)
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
        traffic, tsrc = load_traffic(args)
        out["roofline"] = {
            "kernel": "PageRank flat-stream iteration (k_pf_tiles + k_pf_rows + k_pf_resid), "
                      "timed with HIP events on the graph's stream around the counted "
                      "iterations (launch gaps included)"
                      if args.pr_layout == 1 else f"PageRank iteration, layout {args.pr_layout}",
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

    if not args.no_parity:
        out["parity_full_size"] = full_size_parity(g, V, args, batches, n_raw)
    out["h2d_inclusive"] = h2d_inclusive(g, G, V, args)
    if args.cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(g, V, args, batches)
    if not args.no_tc:
        out["tc_update"] = tc_update(g, G, V, args)
    g.close()
    if args.batched:
        out["batched"] = run_batched(args, G, skey, V, n_raw, batches, E_end, value)
    return out


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
    return out


def full_size_parity(g, V, args, batches, n_raw):
    """Parity gate on the benchmark graph itself: the oracle is built from the
    host R-MAT generator (the same stream as the device one) and given the
    same warm-up and timed batches; degrees and every sorted neighbour set
    must equal the device store's after the timed batches."""
    from oracle import oracle as O
    t0 = time.perf_counter()
    s, d = O.gen_rmat(skey(args.seed, 1), args.scale, n_raw)
    o = O.OracleGraph(V, False)
    o.bulk_build(s, d)
    del s, d
    for b in range(args.warmup + args.steps):
        ds, dd, is_, id_ = batches[b]
        o.delete(ds.numpy(), dd.numpy())
        o.insert(is_.numpy(), id_.numpy())
    goff, gcols = g.export_csr()
    ooff, ocols = o.export_csr()
    o.close()
    deg_eq = bool(np.array_equal(goff, ooff))
    sets_eq = deg_eq and bool(np.array_equal(gcols, ocols))
    return {"degrees_equal": deg_eq, "neighbour_sets_equal": sets_eq,
            "directed_edges": int(goff[-1]), "check_s": round(time.perf_counter() - t0, 1),
            "what": "sequential oracle built from the same R-MAT stream and the same "
                    f"{args.warmup + args.steps} batches, compared with the device store's "
                    "sorted export after the timed batches"}


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
    pv, pv2 = DeviceArray(V, np.uint64), DeviceArray(V, np.uint64)
    tot, tot2 = ctypes.c_uint64(), ctypes.c_uint64()
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
    G._check(L.gv_triangle_count(g.handle, pv2.ptr, G.GV_MEM_DEVICE, ctypes.byref(tot2)))
    res["exact"] = bool(tot.value == tot2.value and np.array_equal(pv.numpy(), pv2.numpy()))
    res["speedup_vs_recount"] = res["full_recount_ms"] / (res["delete_half_ms"]
                                                          + res["insert_half_ms"])
    res["what"] = (f"one {args.batch}-edge batch (delete half, insert half), each applied with "
                   "its triangle-count update; times include the update itself and the "
                   "host round trips of the call")
    for a in (ds, dd, is_, id_, pv, pv2):
        a.free()
    return res


def run_batched(args, G, skey, V, n_raw, batches, E_plain_end, plain_value):
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
    g2.close()
    val = args.batch * args.steps / (ms / 1e3)
    return {
        "value": val,
        "unit": "edge updates/s",
        "ms_total": ms,
        "batches_in_group": len(parts),
        "speedup_vs_plain": val / plain_value,
        "final_edges_equal_plain": bool(E2 == E_plain_end),
        "form": "one gv_batch_apply over the same timed batches (delete half, insert half "
                "each) + one compaction at 25% tombstones",
    }


def cpu_baseline(g, V, args, batches):
    """OpenMP CSR baseline (sorted CSR rebuilt per batch by parallel
    sort-merge) on the same graph and the first batches, timed on this host."""
    from oracle import oracle as O
    off, cols = g.export_csr()
    nbat = min(args.cpu_batches, len(batches))
    t_total = 0.0
    done = 0
    for b in range(nbat):
        ds, dd, is_, id_ = batches[b]
        sd, dd_h, si, di = ds.numpy(), dd.numpy(), is_.numpy(), id_.numpy()
        t0 = time.perf_counter()
        off, cols = O.omp_apply_batch(V, False, off, cols, sd, dd_h, delete=True)
        off, cols = O.omp_apply_batch(V, False, off, cols, si, di, delete=False)
        t_total += time.perf_counter() - t0
        done += args.batch
    out = {
        "value": done / t_total,
        "unit": "edge updates/s",
        "cores": O.omp_threads(),
        "kind": "port",
        "sample": f"{nbat} of the same 1M-edge batches applied to the same scale-{args.scale} "
                  "graph by OpenMP sorted-CSR rebuild (parallel radix sort + merge)",
        "cpu_model": _cpu_model(),
    }
    out["sequential_oracle"] = cpu_sequential(g, V, args, batches)
    return out


def cpu_sequential(g, V, args, batches):
    """The host-sequential oracle (one core: sorted adjacency arrays, batch
    radix-sorted and merged per touched vertex) loaded with the same graph
    (not timed), then one of the same 1M-edge batches timed."""
    from oracle import oracle as O
    off, cols = g.export_csr()
    u = np.repeat(np.arange(V, dtype=np.uint32), np.diff(off).astype(np.int64))
    keep = u < cols
    o = O.OracleGraph(V, False)
    o.bulk_build(u[keep], cols[keep])
    del u, keep, off, cols
    ds, dd, is_, id_ = batches[0]
    sd, dd_h, si, di = ds.numpy(), dd.numpy(), is_.numpy(), id_.numpy()
    t0 = time.perf_counter()
    o.delete(sd, dd_h)
    o.insert(si, di)
    t = time.perf_counter() - t0
    o.close()
    return {"value": args.batch / t, "unit": "edge updates/s", "cores": 1, "kind": "port",
            "sample": f"1 of the same 1M-edge batches (delete half, insert half) on the "
                      f"scale-{args.scale} graph, oracle/gv_oracle.c on one core"}


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "pagerank_flat_rmat22_traffic.json")
PR_ITER_KERNELS = ("k_pf_tiles", "k_pf_rows", "k_pf_resid")


def load_traffic(args):
    """Measured HBM bytes per PageRank iteration (rocprofv3 FETCH_SIZE +
    WRITE_SIZE, calibrated) for this workload, if a summary made by
    --summarize-pmc for the same graph and layout is committed."""
    try:
        with open(TRAFFIC_JSON) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None, None
    w = d.get("workload", {})
    if (w.get("scale"), w.get("edge_factor"), w.get("seed"), w.get("pr_layout")) != \
            (args.scale, args.edge_factor, args.seed, args.pr_layout):
        return None, None
    return d["traffic_bytes_per_iteration"], os.path.relpath(TRAFFIC_JSON, ROOT)


def summarize_pmc(args, fetch_db, write_db, out_json):
    import sqlite3

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
        "workload": {"scale": args.scale, "edge_factor": args.edge_factor, "seed": args.seed,
                     "pr_layout": args.pr_layout},
        "traffic_bytes_per_iteration": total,
        "counters": res,
        "made_by": "bench.py --summarize-pmc FETCH_DB WRITE_DB OUT_JSON",
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
