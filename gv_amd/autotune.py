"""Autotuner for the store's layout and the PageRank launch shapes.

The edge-block size (32, 64 or 128 four-byte slots) fixes the slot layout, so
each candidate needs its own graph: the tuner builds the caller's graph once
per block size, times the caller's update workload on it, then times PageRank
on the updated graph under each candidate set of degree-bin thresholds
(vertex in-degree <= thread_max -> one lane, <= group_max -> eight lanes,
<= wave_max -> one wavefront, larger -> wavefront chunks of wave_max slots),
switched live with gv_set_tuning.  The score of a (block size, bins) pair is
update time + PageRank time, both measured with HIP events on the handle's
stream; the smallest score wins.

The reference fixes these choices at compile time (128-slot blocks and one
thread per vertex, constants.cuh:4-17, pr_kernel.cu:3-32); there is no
counterpart to port.
"""
import numpy as np

from . import GraphVine, default_config, pagerank_device
from .hip import DeviceArray, Event

BLOCK_SLOTS = (32, 64, 128)
# (thread_max, group_max, wave_max)
PR_BINS = ((8, 96, 4096), (4, 64, 2048), (16, 256, 8192), (8, 128, 16384), (2, 32, 1024))


def tune(V, directed, build, updates, block_slots=BLOCK_SLOTS, pr_bins=PR_BINS, alpha=0.85,
         eps=1e-6, max_iter=100, log=None, **cfg_overrides):
    """Sweep and return {"best": {...}, "table": [...]}.

    build(g): bulk-builds the workload's graph into a fresh handle g.
    updates(g): issues the update workload on g (it may enqueue work without
    waiting; the tuner brackets it with events on g's stream).
    """
    table = []
    best = None
    for B in block_slots:
        cfg = default_config(block_slots=B, **cfg_overrides)
        g = GraphVine(V, directed=directed, config=cfg)
        try:
            build(g)
            g.sync()
            stream = g.stream()
            e0, e1 = Event(), Event()
            e0.record(stream)
            updates(g)
            e1.record(stream)
            e1.synchronize()
            g.sync()
            upd_ms = e0.elapsed_ms(e1)
            out = DeviceArray(V, np.float32)
            for tm, gm, wm in pr_bins:
                g.set_tuning(pr_thread_max=tm, pr_group_max=gm, pr_wave_max=wm, pr_layout=1)
                pagerank_device(g, out, alpha, eps, max_iter)  # settle allocations
                e0.record(stream)
                st = pagerank_device(g, out, alpha, eps, max_iter)
                e1.record(stream)
                e1.synchronize()
                pr_ms = e0.elapsed_ms(e1)
                row = {"block_slots": B, "pr_thread_max": tm, "pr_group_max": gm,
                       "pr_wave_max": wm, "update_ms": upd_ms, "pagerank_ms": pr_ms,
                       "pagerank_iters": int(st["iters"]),
                       "pagerank_ms_per_iter": pr_ms / max(int(st["iters"]), 1),
                       "score_ms": upd_ms + pr_ms}
                table.append(row)
                if log:
                    log(row)
                if best is None or row["score_ms"] < best["score_ms"]:
                    best = row
            out.free()
        finally:
            g.close()
    return {"best": best, "table": table}
