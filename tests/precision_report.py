"""Precision report of the FP32 analytics at full size (not a pytest module;
run directly): directed R-MAT scale 24 (16.8M vertices, edge factor 16,
seed 3).  PageRank with per-vertex Kahan sums and an FP64 residual is run to
L1 1e-8 (or 300 updates), and plain FP32 sums to the same number of updates;
BC from 64 random sources (FP64 sigma, FP32 compensated dependencies).
Errors are against the FP64 OpenMP references: L1 and L-infinity distance to
the FP64 iteration run for the same number of updates (rounding) and to the
FP64 fixed point (convergence); BC relative error per vertex.  Prints one
JSON object.

    python tests/precision_report.py [--scale 24] [--out FILE]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--sources", type=int, default=64)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import gv_amd as G
    from gv_amd.hip import DeviceArray
    from oracle import oracle as O

    def log(*a):
        print(*a, file=sys.stderr, flush=True)

    scale, V = args.scale, 1 << args.scale
    seed = 3
    t0 = time.time()
    g = G.GraphVine(V, directed=True)
    chunk = 1 << 27
    for c0 in range(0, 16 * V, chunk):
        n = min(chunk, 16 * V - c0)
        key = (G.stream_key(seed, 1) + c0 * scale * 0x9E3779B97F4A7C15) % (1 << 64)
        s, d = G.gen_rmat(g, key, scale, n)
        (g.bulkBuild if c0 == 0 else g.batchInsert)(s, d, stats=False)
        g.sync()
        s.free()
        d.free()
    E = g.num_edges()
    log(f"built V={V} E={E} in {time.time() - t0:.1f} s")
    out = {"graph": f"directed R-MAT scale {scale}, edge factor 16, seed {seed}",
           "vertices": V, "edges": int(E)}
    dev = DeviceArray(V, np.float32)
    res = {}
    for comp in (True, False):
        G.pagerank(g, epsilon=1e-8, max_iter=2, compensated=comp)  # warm-up
    xc, itc, rc, stc = G.pagerank(g, epsilon=1e-8, max_iter=300, compensated=True, stats=True)
    xp, itp, rp, stp = G.pagerank(g, epsilon=0.0, max_iter=itc, compensated=False, stats=True)
    log(f"GPU PageRank: {itc} updates, residual {rc:.3e}")
    src = np.random.default_rng(seed).choice(V, args.sources, replace=False).astype(np.uint32)
    t = time.time()
    b = G.bc(g, src).astype(np.float64)
    bc_s = time.time() - t
    off, cols = g.export_csr()
    g.close()
    log("exported; FP64 references ...")
    t = time.time()
    x_it, _, _ = O.omp_pagerank_dir(V, off, cols, fixed_iters=itc)
    x_fix, it_fix, r_fix = O.omp_pagerank_dir(V, off, cols, eps=1e-15, max_iter=3000)
    log(f"FP64 PageRank done in {time.time() - t:.1f} s ({it_fix} updates to the fixed point)")
    t = time.time()
    bo = O.omp_bc_dir(V, off, cols, src)
    cpu_bc_s = time.time() - t
    log(f"FP64 BC done in {cpu_bc_s:.1f} s")

    def errs(x, ref):
        e = np.abs(x.astype(np.float64) - ref)
        return {"L1": float(e.sum()), "Linf": float(e.max()),
                "Linf_rel_to_max": float(e.max() / ref.max())}

    out["pagerank"] = {
        "stop": "L1 residual < 1e-8 or 300 updates",
        "updates": itc,
        "final_residual": rc,
        "compensated": {"vs_fp64_same_updates": errs(xc, x_it), "vs_fp64_fixed_point":
                        errs(xc, x_fix), "kernel_ms_per_update": stc["kernel_ms"] / itc,
                        "gteps": E * itc / (stc["kernel_ms"] / 1e3) / 1e9},
        "plain_fp32": {"vs_fp64_same_updates": errs(xp, x_it), "vs_fp64_fixed_point":
                       errs(xp, x_fix), "kernel_ms_per_update": stp["kernel_ms"] / itp,
                       "gteps": E * itp / (stp["kernel_ms"] / 1e3) / 1e9},
        "fp64_fixed_point_updates": it_fix,
        "fp32_ulp_at_1_over_V": float(np.spacing(np.float32(1.0 / V))),
    }
    rel = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
    out["bc"] = {"sources": args.sources, "max_rel_error": float(rel.max()),
                 "L1": float(np.abs(b - bo).sum()), "Linf": float(np.abs(b - bo).max()),
                 "gpu_s": bc_s, "cpu_fp64_s": cpu_bc_s, "cpu_threads": O.omp_threads()}
    txt = json.dumps(out, indent=1)
    print(txt, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main()
