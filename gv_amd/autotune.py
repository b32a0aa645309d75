"""Autotuner over the knobs the default paths read.

Every knob below changes what runs, and the tuner measures each on the
caller's own workload:

  block_slots   edge-block size (32, 64 or 128 four-byte slots).  It fixes
                the slot layout, so each candidate gets its own graph: the
                tuner builds the caller's graph once per block size and times
                the caller's update workload on it.
  pr_layout     PageRank layout (5: hot stream + propagation-blocked cold
                stream over a degree-relabelled snapshot; 1: the same
                snapshot as one flat stream; 4: the snapshot row by row),
                timed on the updated graph, switched live with
                gv_set_tuning.  Layout 4 cuts rows into degree bins
                (in-degree <= thread_max: one lane, <= group_max: eight
                lanes, <= wave_max: a wavefront, larger: wavefront chunks of
                wave_max slots), so it is timed under every bin set; the
                other layouts do not read the bins.
  bc_bu_ratio   multi-source BC: a BFS level is expanded bottom-up when its
                candidates' in-slots are below this multiple of its
                frontier's out-slots (0: always top-down), timed on one BC
                call from the caller's sources.

Scores are device times from HIP events on the handle's stream.  For each
block size the best layout (with its bins) and the best BC ratio are picked
independently (they time different calls); the block size with the smallest
update + PageRank + BC total wins.

The reference fixes these choices at compile time (128-slot blocks and one
thread per vertex, constants.cuh:4-17, pr_kernel.cu:3-32; BC always top-down,
bc_kernel.cu:19-56); there is no counterpart to port.
"""
import numpy as np

from . import GV_MEM_DEVICE, GraphVine, _check, default_config, lib, pagerank_device
from .hip import DeviceArray, Event

BLOCK_SLOTS = (32, 64, 128)
PR_LAYOUTS = (5, 1, 4)
BINNED_LAYOUTS = (4,)
# (thread_max, group_max, wave_max)
PR_BINS = ((8, 96, 4096), (4, 64, 2048), (16, 256, 8192), (2, 32, 1024))
BC_RATIOS = (0.0, 1.0, 2.0, 3.0, 4.0, 8.0)


def _timed(g, fn):
    stream = g.stream()
    e0, e1 = Event(), Event()
    e0.record(stream)
    r = fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_ms(e1), r


def tune(V, directed, build, updates, block_slots=BLOCK_SLOTS, pr_layouts=PR_LAYOUTS,
         pr_bins=PR_BINS, bc_ratios=BC_RATIOS, bc_sources=None, alpha=0.85, eps=1e-6,
         max_iter=100, log=None, **cfg_overrides):
    """Sweep and return {"best": {...}, "table": [...]}.

    build(g): bulk-builds the workload's graph into a fresh handle g.
    updates(g): issues the update workload on g (it may enqueue work without
    waiting; the tuner brackets it with events on g's stream).
    bc_sources: BC sources (uint32 array; default 16 seeded ones); an empty
    array skips the BC sweep.
    Every row of the table is one timed call: kind "update", "pagerank"
    (layout and bins) or "bc" (ratio)."""
    if bc_sources is None:
        bc_sources = np.random.default_rng(4).choice(V, min(16, V), replace=False)
    srcs = np.ascontiguousarray(bc_sources, dtype=np.uint32)
    table = []
    best = None

    def add(row):
        table.append(row)
        if log:
            log(row)

    for B in block_slots:
        cfg = default_config(block_slots=B, **cfg_overrides)
        g = GraphVine(V, directed=directed, config=cfg)
        out = bc_out = s_dev = None
        try:
            build(g)
            g.sync()
            upd_ms, _ = _timed(g, lambda: updates(g))
            g.sync()
            add({"kind": "update", "block_slots": B, "ms": upd_ms})
            out = DeviceArray(V, np.float32)
            pr_best = None
            for layout in pr_layouts:
                for tm, gm, wm in (pr_bins if layout in BINNED_LAYOUTS else pr_bins[:1]):
                    g.set_tuning(pr_thread_max=tm, pr_group_max=gm, pr_wave_max=wm,
                                 pr_layout=layout)
                    pagerank_device(g, out, alpha, eps, max_iter)  # settle allocations
                    ms, st = _timed(g, lambda: pagerank_device(g, out, alpha, eps, max_iter))
                    row = {"kind": "pagerank", "block_slots": B, "pr_layout": layout,
                           "pr_thread_max": tm, "pr_group_max": gm, "pr_wave_max": wm,
                           "ms": ms, "iters": int(st["iters"]),
                           "ms_per_iter": ms / max(int(st["iters"]), 1)}
                    add(row)
                    if pr_best is None or ms < pr_best["ms"]:
                        pr_best = row
            bc_best = {"bc_bu_ratio": g.get_config().bc_bu_ratio, "ms": 0.0}
            if len(srcs) > 1:
                s_dev = DeviceArray.from_numpy(srcs)
                bc_out = DeviceArray(V, np.float32)
                bc_best = None
                for r in bc_ratios:
                    g.set_tuning(bc_bu_ratio=r)

                    def bc_call():
                        _check(lib().gv_bc(g.handle, s_dev.ptr, len(srcs), bc_out.ptr,
                                           GV_MEM_DEVICE))
                    bc_call()  # settle the workspace
                    ms, _ = _timed(g, bc_call)
                    row = {"kind": "bc", "block_slots": B, "bc_bu_ratio": r, "ms": ms,
                           "sources": len(srcs)}
                    add(row)
                    if bc_best is None or ms < bc_best["ms"]:
                        bc_best = row
            score = upd_ms + pr_best["ms"] + bc_best["ms"]
            cand = {"block_slots": B, "pr_layout": pr_best["pr_layout"],
                    "pr_thread_max": pr_best["pr_thread_max"],
                    "pr_group_max": pr_best["pr_group_max"],
                    "pr_wave_max": pr_best["pr_wave_max"],
                    "bc_bu_ratio": bc_best["bc_bu_ratio"], "update_ms": upd_ms,
                    "pagerank_ms": pr_best["ms"], "bc_ms": bc_best["ms"], "score_ms": score}
            if best is None or score < best["score_ms"]:
                best = cand
        finally:
            for buf in (out, bc_out, s_dev):
                if buf is not None:
                    buf.free()
            g.close()
    return {"best": best, "table": table}


def overrides(best):
    """The tuned choice as GraphVine config overrides (block_slots cannot change
    on a live handle, so a new handle is built with these)."""
    keys = ("block_slots", "pr_layout", "pr_thread_max", "pr_group_max", "pr_wave_max",
            "bc_bu_ratio")
    return {k: best[k] for k in keys}
