"""Precision of the FP32 analytics against FP64 references (the OpenMP
restatements, themselves checked against the sequential oracle):
  - PageRank: per-vertex Kahan sums + FP64 residual; the rounding error is the
    L1 distance to the FP64 iteration run for the same number of updates, the
    convergence error the distance to the FP64 fixed point;
  - BC: FP64 sigma, FP32 compensated dependencies, relative error <= 1e-5."""
import numpy as np
import pytest

import gv_amd as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def build(scale, directed, seed):
    V = 1 << scale
    s, d = O.gen_rmat(O.stream_key(seed, O.STREAM_RMAT), scale, 16 * V)
    cfg = G.default_config()
    cfg.pool_blocks = 8 * V + 4096
    g = G.GraphVine(V, directed=directed, config=cfg)
    g.bulkBuild(s, d)
    return g, V


@pytest.mark.parametrize("layout", [1, 4, 5])
def test_pagerank_rounding_and_convergence_error(layout):
    g, V = build(16, True, 71)
    g.set_tuning(pr_layout=layout)
    off, cols = g.export_csr()
    xc, itc, rc, st = G.pagerank(g, epsilon=1e-8, max_iter=300, stats=True)
    xp, itp, rp = G.pagerank(g, epsilon=0.0, max_iter=itc, compensated=False)
    assert itp == itc
    x_it, _, _ = O.omp_pagerank_dir(V, off, cols, fixed_iters=itc)
    x_fix, _, r_fix = O.omp_pagerank_dir(V, off, cols, eps=1e-15, max_iter=2000)
    err_c = np.abs(xc.astype(np.float64) - x_it)
    err_p = np.abs(xp.astype(np.float64) - x_it)
    # rounding: a few FP32 ulps per vertex with compensated sums (per-update
    # rounding compounds over the iterations, damped by alpha)
    ulp = np.spacing(np.float32(1.0 / V)).astype(np.float64)
    assert err_c.sum() <= 16 * V * ulp, (err_c.sum(), V * ulp)
    assert err_c.max() <= 1e-6 * x_it.max()
    assert err_c.sum() <= err_p.sum() * 1.05 + V * ulp
    # convergence: the stop rule bounds the distance to the fixed point
    if rc < 1e-8:
        assert np.abs(xc.astype(np.float64) - x_fix).sum() <= 1e-7
    g.close()


def test_bc_64_sources_fp64_reference():
    g, V = build(14, True, 72)
    off, cols = g.export_csr()
    src = np.random.default_rng(72).choice(V, 64, replace=False).astype(np.uint32)
    b = G.bc(g, src).astype(np.float64)
    bo = O.omp_bc_dir(V, off, cols, src)
    rel = np.abs(b - bo) / np.maximum(np.abs(bo), 1.0)
    assert rel.max() <= 1e-5, rel.max()
    g.close()
