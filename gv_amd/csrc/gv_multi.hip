// gv_multi.hip -- many-vector analytics: personalized PageRank as an FP32
// column block, and Brandes BC for a batch of sources at once.
//
// Personalized PageRank (the reference's static_pagerank iteration,
// pagerank.cu:17-59, with the teleport term concentrated on one seed per
// column): x_c' = (1-a) e_{s_c} + a * sum_{u in N_in(v)} x_c[u]/deg(u),
// dangling mass dropped, x_c starts at e_{s_c}; stop after the first update
// whose largest per-column L1 residual is below eps, or after max_iter.
// Layout: X[v*64 + c] (row of 64 FP32 columns = 256 B).  One wavefront per
// vertex, lane = column: every in-edge u reads the contiguous row
// CB[u*64 .. u*64+63], so the gathers are whole, coalesced 256-byte rows and
// the block is HBM-bound instead of gather-bound.  Kahan per lane.
//
// Multi-source BC (bc, bc.cu:3-88; Brandes as in the single-source path):
// sources are processed S = 32 at a time.  dist/sigma/delta are [V][S]
// arrays; a group of S lanes handles one frontier vertex, lane = source, and
// the frontier of a level is the union over the batch's sources (a vertex is
// queued once for every distinct depth at which some source reaches it).  Forward:
// CAS on dist[w][s], FP64 atomic sigma; backward: per-lane Kahan sums over
// successors, then the S dependencies are added in a fixed lane order to
// BC[v] (FP32 Kahan).
#include <vector>

#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {

constexpr int PPR_COLS = 64;
constexpr int PPR_THREADS = 256;
constexpr uint32_t PPR_GRID = 2048;
// Rows longer than PPR_HUB slots are summed in chunks of PPR_HUB by separate
// wavefronts and merged in chunk order: with one wavefront per row, the
// iteration waited for the wavefront holding the largest hub.
constexpr uint32_t PPR_HUB = 4096;
constexpr uint32_t PPR_FIN_GRID = 256;
// contribution rows (256 bytes each) requested per wavefront before the first
// is added
constexpr int PPR_ILP = 16;

struct PprArgs {
  Store in;
  const uint32_t *deg_out;
  const uint32_t *seeds;
  uint32_t ncols;
  float alpha;
  double eps;
  uint32_t max_iter;
  double *partials;  // [gridDim][64]
  PrCtl *ctl;
  double *col_res;   // [64] residual of the last update per column
};

__global__ void k_ppr_init(float *X, float *CB, const uint32_t *deg_out, const uint32_t *seeds,
                           uint32_t ncols, uint32_t V) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)V * PPR_COLS;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v = (uint32_t)(i / PPR_COLS), c = (uint32_t)(i % PPR_COLS);
    float x = (c < ncols && seeds[c] == v) ? 1.0f : 0.0f;
    uint32_t d = deg_out[v];
    X[i] = x;
    CB[i] = d ? x / (float)d : 0.f;
  }
}

__global__ __launch_bounds__(PPR_THREADS) void k_ppr_iter(PprArgs a, const float *X, const float *CB,
                                                          float *Xn, float *CBn) {
  if (a.ctl->done) return;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t V = a.in.V;
  const uint32_t seed = lane < a.ncols ? a.seeds[lane] : 0xFFFFFFFFu;
  double diff = 0.0;
  for (uint64_t v0 = (uint64_t)blockIdx.x * (PPR_THREADS / WAVE); v0 < V;
       v0 += (uint64_t)gridDim.x * (PPR_THREADS / WAVE)) {
    const uint64_t v = v0 + wid;
    if (v >= V) continue;  // uniform per wavefront
    const uint32_t us = a.in.used[v], tbl = a.in.tbl_off[v];
    if (us > PPR_HUB) continue;  // chunked: k_ppr_hub, k_ppr_hubfin
    float s = 0.f, c = 0.f;
    for (uint32_t base = 0; base < us; base += WAVE) {
      const uint32_t n = us - base < (uint32_t)WAVE ? us - base : (uint32_t)WAVE;
      uint32_t slot = TOMB;
      if (lane < n) slot = a.in.slots[slot_index(a.in.arena, tbl, base + lane, a.in.log2B, a.in.B)];
      for (uint32_t t = 0; t < n; t += PPR_ILP) {
        uint32_t u[PPR_ILP];
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k) u[k] = __shfl(slot, (int)((t + k) & (WAVE - 1)), WAVE);
        float g[PPR_ILP];
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k)
          g[k] = (t + k < n && u[k] != TOMB) ? CB[(uint64_t)u[k] * PPR_COLS + lane] : 0.f;
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k)
          if (t + k < n && u[k] != TOMB) kahan_add(s, c, g[k]);
      }
    }
    float xv = a.alpha * (s - c) + ((uint32_t)v == seed ? 1.0f - a.alpha : 0.0f);
    if (lane >= a.ncols) xv = 0.f;
    const uint64_t idx = v * PPR_COLS + lane;
    const uint32_t d = a.deg_out[v];
    diff += fabs((double)xv - (double)X[idx]);
    Xn[idx] = xv;
    CBn[idx] = d ? xv / (float)d : 0.f;
  }
  // per-column partial of this workgroup, waves added in a fixed order
  __shared__ double red[PPR_THREADS / WAVE][WAVE];
  red[wid][lane] = diff;
  __syncthreads();
  if (wid == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < PPR_THREADS / WAVE; ++w) t += red[w][lane];
    a.partials[(uint64_t)blockIdx.x * PPR_COLS + lane] = t;
  }
}

// Chunk counts of the rows longer than PPR_HUB slots (0 for the others).
__global__ void k_ppr_hubcnt(Store in, uint32_t *cnt) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < in.V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = in.used[v];
    cnt[v] = u > PPR_HUB ? (u + PPR_HUB - 1) / PPR_HUB : 0u;
  }
}

// iv[i] = row of chunk item i (coff = exclusive scan of the chunk counts).
__global__ void k_ppr_hubitems(const uint32_t *coff, uint32_t V, uint32_t *iv) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = coff[v], e = coff[v + 1]; i < e; ++i) iv[i] = (uint32_t)v;
}

// One wavefront per chunk item, lane = column: the Kahan sum of the chunk's
// contribution rows, in slot order.
__global__ __launch_bounds__(PPR_THREADS) void k_ppr_hub(PprArgs a, const float *CB,
                                                         const uint32_t *iv, const uint32_t *coff,
                                                         uint32_t n_items, float2 *part) {
  if (a.ctl->done) return;
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; it < n_items;
       it += nwaves) {
    const uint32_t v = iv[it];
    const uint32_t us = a.in.used[v], tbl = a.in.tbl_off[v];
    const uint32_t lo = ((uint32_t)it - coff[v]) * PPR_HUB;
    const uint32_t hi = lo + PPR_HUB < us ? lo + PPR_HUB : us;
    float s = 0.f, c = 0.f;
    for (uint32_t base = lo; base < hi; base += WAVE) {
      const uint32_t n = hi - base < (uint32_t)WAVE ? hi - base : (uint32_t)WAVE;
      uint32_t slot = TOMB;
      if (lane < n) slot = a.in.slots[slot_index(a.in.arena, tbl, base + lane, a.in.log2B, a.in.B)];
      for (uint32_t t = 0; t < n; t += PPR_ILP) {
        uint32_t u[PPR_ILP];
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k) u[k] = __shfl(slot, (int)((t + k) & (WAVE - 1)), WAVE);
        float g[PPR_ILP];
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k)
          g[k] = (t + k < n && u[k] != TOMB) ? CB[(uint64_t)u[k] * PPR_COLS + lane] : 0.f;
#pragma unroll
        for (int k = 0; k < PPR_ILP; ++k)
          if (t + k < n && u[k] != TOMB) kahan_add(s, c, g[k]);
      }
    }
    part[it * PPR_COLS + lane] = make_float2(s, c);
  }
}

// One wavefront per chunked row (its first chunk item): the chunk sums
// joined in chunk order (TwoSum), the update, and per-column residual
// partials in rows roff + blockIdx.x of a.partials.
__global__ __launch_bounds__(PPR_THREADS) void k_ppr_hubfin(PprArgs a, const float *X, float *Xn,
                                                            float *CBn, const uint32_t *iv,
                                                            const uint32_t *coff, uint32_t n_items,
                                                            const float2 *part, uint32_t roff) {
  if (a.ctl->done) return;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  double diff = 0.0;
  for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; it < n_items;
       it += nwaves) {
    const uint32_t v = iv[it];
    if ((uint32_t)it != coff[v]) continue;  // not the row's first chunk
    float s = 0.f, c = 0.f;
    for (uint32_t i = coff[v], e = coff[v + 1]; i < e; ++i) {
      const float2 q = part[(uint64_t)i * PPR_COLS + lane];
      const float t = s + q.x;
      const float bp = t - s;
      const float er = (s - (t - bp)) + (q.x - bp);
      s = t;
      c = c + q.y - er;
    }
    const uint32_t seed = lane < a.ncols ? a.seeds[lane] : 0xFFFFFFFFu;
    float xv = a.alpha * (s - c) + (v == seed ? 1.0f - a.alpha : 0.0f);
    if (lane >= a.ncols) xv = 0.f;
    const uint64_t idx = (uint64_t)v * PPR_COLS + lane;
    const uint32_t d = a.deg_out[v];
    diff += fabs((double)xv - (double)X[idx]);
    Xn[idx] = xv;
    CBn[idx] = d ? xv / (float)d : 0.f;
  }
  __shared__ double red[PPR_THREADS / WAVE][WAVE];
  red[wid][lane] = diff;
  __syncthreads();
  if (wid == 0) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < PPR_THREADS / WAVE; ++w) t += red[w][lane];
    a.partials[(uint64_t)(roff + blockIdx.x) * PPR_COLS + lane] = t;
  }
}

__global__ __launch_bounds__(PPR_COLS) void k_ppr_resid(PprArgs a, uint32_t nblocks) {
  if (a.ctl->done) return;
  const uint32_t c = threadIdx.x;
  double r = 0.0;
  for (uint32_t b = 0; b < nblocks; ++b) r += a.partials[(uint64_t)b * PPR_COLS + c];
  a.col_res[c] = r;
  double m = c < a.ncols ? r : 0.0;
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) m = fmax(m, __shfl_xor(m, d, WAVE));
  if (c == 0) {
    a.ctl->residual = m;
    unsigned it = a.ctl->iters + 1;
    a.ctl->iters = it;
    if (m < a.eps || it >= a.max_iter) a.ctl->done = 1;
  }
}

// Copy the first ncols columns of the 64-wide block out (row-major V x ncols).
__global__ void k_ppr_out(const float *X, uint32_t V, uint32_t ncols, float *out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)V * ncols;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = X[(i / ncols) * PPR_COLS + i % ncols];
}

gv_status ppr_run(const Store &in, const uint32_t *deg_out, float alpha, double eps, int max_iter,
                  const uint32_t *seeds_dev, uint32_t ncols, float *out_dev, int *iters,
                  double *max_residual, double *iter_ms, hipStream_t s) {
  const uint32_t V = in.V;
  if (iters) *iters = 0;
  if (max_residual) *max_residual = 0.0;
  if (!V || !ncols) return GV_OK;
  const size_t blk = (size_t)V * PPR_COLS * sizeof(float);
  float *X[2] = {nullptr, nullptr}, *CB[2] = {nullptr, nullptr};
  double *partials = nullptr, *col_res = nullptr;
  PrCtl *ctl = nullptr, *h_ctl = nullptr;
  uint32_t *hcnt = nullptr, *hoff = nullptr, *hiv = nullptr;
  float2 *hpart = nullptr;
  void *htmp = nullptr;
  std::vector<hipEvent_t> ev;
  gv_status rc = GV_OK;
  do {
    hipError_t e = hipSuccess;
    for (int i = 0; i < 2 && e == hipSuccess; ++i) {
      e = hipMalloc(&X[i], blk);
      if (e == hipSuccess) e = hipMalloc(&CB[i], blk);
    }
    if (e == hipSuccess)
      e = hipMalloc(&partials, (size_t)(PPR_GRID + PPR_FIN_GRID) * PPR_COLS * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&hcnt, ((size_t)V + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&hoff, ((size_t)V + 1) * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&htmp, scan_tmp_bytes((uint64_t)V + 1));
    if (e == hipSuccess) e = hipMalloc(&col_res, PPR_COLS * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&ctl, sizeof(PrCtl));
    if (e == hipSuccess) e = hipHostMalloc(&h_ctl, sizeof(PrCtl), hipHostMallocDefault);
    if (e != hipSuccess) {
      rc = hip_fail(e, "personalized PageRank alloc");
      break;
    }
    // chunk items of the rows longer than PPR_HUB slots
    k_ppr_hubcnt<<<grid_for(V), 256, 0, s>>>(in, hcnt);
    scan_u32(hcnt, hoff, V, nullptr, htmp, s);
    uint32_t n_items = 0;
    if ((e = hipMemcpyAsync(&n_items, hoff + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      rc = hip_fail(e, "personalized PageRank");
      break;
    }
    if (n_items) {
      if ((e = hipMalloc(&hiv, (size_t)n_items * sizeof(uint32_t))) != hipSuccess ||
          (e = hipMalloc(&hpart, (size_t)n_items * PPR_COLS * sizeof(float2))) != hipSuccess) {
        rc = hip_fail(e, "personalized PageRank alloc");
        break;
      }
      k_ppr_hubitems<<<grid_for(V), 256, 0, s>>>(hoff, V, hiv);
    }
    const uint32_t hub_grid = (uint32_t)std::min<uint64_t>(cdiv(n_items, PPR_THREADS / WAVE), 4096);
    const uint32_t fin_grid =
        (uint32_t)std::min<uint64_t>(cdiv(n_items, PPR_THREADS / WAVE), PPR_FIN_GRID);
    pr_ctl_reset(ctl, s);
    k_ppr_init<<<grid_for((uint64_t)V * PPR_COLS, 256, 16384), 256, 0, s>>>(X[0], CB[0], deg_out,
                                                                           seeds_dev, ncols, V);
    PprArgs a;
    a.in = in;
    a.deg_out = deg_out;
    a.seeds = seeds_dev;
    a.ncols = ncols;
    a.alpha = alpha;
    a.eps = eps;
    a.max_iter = (uint32_t)(max_iter > 0 ? max_iter : 0);
    a.partials = partials;
    a.ctl = ctl;
    a.col_res = col_res;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(cdiv(V, PPR_THREADS / WAVE), PPR_GRID);
    int launched = 0;
    for (; launched < max_iter && rc == GV_OK; ++launched) {
      const int t = launched;
      if (iter_ms) {
        hipEvent_t e0, e1;
        hipEventCreate(&e0);
        hipEventCreate(&e1);
        ev.push_back(e0);
        ev.push_back(e1);
        hipEventRecord(e0, s);
      }
      k_ppr_iter<<<grid, PPR_THREADS, 0, s>>>(a, X[t & 1], CB[t & 1], X[(t + 1) & 1],
                                              CB[(t + 1) & 1]);
      if (n_items) {
        k_ppr_hub<<<hub_grid, PPR_THREADS, 0, s>>>(a, CB[t & 1], hiv, hoff, n_items, hpart);
        k_ppr_hubfin<<<fin_grid, PPR_THREADS, 0, s>>>(a, X[t & 1], X[(t + 1) & 1],
                                                      CB[(t + 1) & 1], hiv, hoff, n_items, hpart,
                                                      grid);
      }
      k_ppr_resid<<<1, PPR_COLS, 0, s>>>(a, grid + fin_grid);
      if (iter_ms) hipEventRecord(ev.back(), s);
      if ((launched + 1) % 8 == 0) {
        if (hipMemcpyAsync(h_ctl, ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) {
          rc = hip_fail(hipGetLastError(), "personalized PageRank");
          break;
        }
        if (h_ctl->done) {
          ++launched;
          break;
        }
      }
    }
    if (rc) break;
    if ((e = hipMemcpyAsync(h_ctl, ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      rc = hip_fail(e, "personalized PageRank");
      break;
    }
    const uint32_t it = h_ctl->iters;
    k_ppr_out<<<grid_for((uint64_t)V * ncols, 256, 16384), 256, 0, s>>>(X[it & 1], V, ncols,
                                                                       out_dev);
    if ((e = hipStreamSynchronize(s)) != hipSuccess) {
      rc = hip_fail(e, "personalized PageRank");
      break;
    }
    if (iters) *iters = (int)it;
    if (max_residual) *max_residual = h_ctl->residual;
    if (iter_ms) {
      double tot = 0.0;
      for (uint32_t t = 0; t < it && 2 * t + 1 < ev.size(); ++t) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev[2 * t], ev[2 * t + 1]);
        tot += ms;
      }
      *iter_ms = tot;
    }
  } while (0);
  for (hipEvent_t e : ev) hipEventDestroy(e);
  for (int i = 0; i < 2; ++i) {
    hipFree(X[i]);
    hipFree(CB[i]);
  }
  hipFree(partials);
  hipFree(col_res);
  hipFree(ctl);
  hipFree(hcnt);
  hipFree(hoff);
  hipFree(hiv);
  hipFree(hpart);
  hipFree(htmp);
  if (h_ctl) hipHostFree(h_ctl);
  return rc;
}

// ================================================================ BC ==
// Brandes BC from a sweep of up to 32 sources at once (bit-parallel
// multi-source BFS, Then et al., VLDB 2014), one lane per source.
//
// Forward, level by level.  An expansion only marks claims: next[w] |= the
// sources that reach w at level d + 1, and sigma.  Opening level d + 1 then
// scans the vertices once: every vertex with a claim is appended to the
// level queue with its source mask (qmask) and one work item per BCM_CH
// slots of its adjacency, each workgroup reserving its range's entries with
// one atomic per counter (appends from the expansion itself would put
// millions of atomics on one cursor).  A level is expanded either top-down
// (the frontier's items push sigma along their edges with FP64 atomics) or
// bottom-up (Beamer, Asanovic, Patterson, SC 2012): the vertices still
// unreached by a source that is active at level d pull sigma from their
// level-d neighbours, with plain stores.  Both are launched for every level
// and the one the level's slot counts do not pick exits at once, so the
// level loop needs no host decision; the host checks for the end of the
// sweep once per BCM_GROUP levels.
//
// Backward, deepest level first: each level-d vertex sums, over its
// successors w (level d + 1 for that source), coef_w = (1 + delta_w) /
// sigma_w; delta_v = sigma_v * sum, and coef_v is stored for the level
// above.  Sources of the sweep with no level-d entry contribute nothing.
constexpr int BCM_S = 32;                      // sources per sweep (at most)
constexpr int BCM_ILP = 8;                     // neighbour rows requested before the first is used
constexpr uint32_t BCM_CH = 256;               // adjacency slots per work item
constexpr uint32_t BCM_MULTI = 0x80000000u;    // item flag: its vertex has several items
constexpr uint32_t BCM_NOLEV = 0xFFFFFFFFu;    // no level yet
constexpr int BCM_GROUP = 8;                   // forward levels per host check
constexpr int BCM_THREADS = 256;

// Level counters (device).  qb/ib/lb: first queue entry, item and
// multi-item entry of the level; pn: partial rows taken by the level's
// multi-item vertices; amask: sources with a level-d vertex; ncand:
// bottom-up items of the level's multi-item candidates; done: workgroups
// of the level's open kernel that have finished; fslots/cslots: adjacency
// slots of the level's vertices and of the bottom-up candidates.
struct BcmLevel {
  uint32_t qb, ib, lb, pn, amask, ncand, done, pad;
  unsigned long long fslots, cslots;
};

// Append cursors (queue entries in the low 32 bits of qi and items in the
// high 32 bits, so one atomic reserves both), the first empty level (the
// depth) and the overflow flag (a queue, item or multi-item buffer was too
// small: the sweep is rerun with larger buffers).
struct BcmCtl {
  unsigned long long qi;
  uint32_t lcur, depth, overflow, pad;
};

struct BcmArgs {
  Store out, in;       // adjacency (successors) and in-adjacency (predecessors)
  double *sigma;       // [V][RS] shortest-path counts
  float *coef;         // [V][RS] (1 + delta) / sigma of processed (vertex, source) pairs
  float *bc, *bcc;     // result and its compensation
  uint2 *sn;           // forward: (seen, next) masks; backward: (mask, level) of level d + 1
  uint2 *fl;           // forward: (mask, level) of each vertex's latest level
  uint32_t *queue, *qmask, *qpart;
  uint2 *items;        // {queue entry, chunk | BCM_MULTI}
  uint32_t *longl;     // queue entries of multi-item vertices
  uint2 *cand;         // bottom-up items {vertex, chunk | BCM_MULTI}
  float2 *part;        // [pn][RS] per-item partial sums of multi-item vertices
  BcmLevel *lv;
  BcmCtl *ctl;
  const uint32_t *src;
  uint32_t ns, V;
  uint32_t qcap, icap, lcap;
  float bu_ratio;      // bottom-up when cand slots < bu_ratio * frontier slots (0: never)
};

__device__ __forceinline__ uint32_t bcm_chunks(uint32_t used) {
  return used > BCM_CH ? (used + BCM_CH - 1) / BCM_CH : 1u;
}

__device__ __forceinline__ bool bcm_live(const BcmArgs &a) {
  return a.ctl->overflow == 0u && a.ctl->depth == BCM_NOLEV;
}

__device__ __forceinline__ bool bcm_bottom_up(const BcmArgs &a, const BcmLevel &L) {
  return a.bu_ratio > 0.f && (double)L.cslots < (double)a.bu_ratio * (double)L.fslots;
}

__device__ __forceinline__ uint32_t wave_or(uint32_t v) {
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) v |= __shfl_xor(v, d, WAVE);
  return v;
}

// Per sweep: sigma rows and masks cleared, level counters reset.
template <int RS>
__global__ void k_bcm_init(BcmArgs a, uint32_t lvcap) {
  const uint64_t ns = (uint64_t)a.V * RS;
  const uint64_t n = ns > lvcap ? ns : lvcap;
  const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += st) {
    if (i < ns) a.sigma[i] = 0.0;
    if (i < a.V) {
      a.sn[i] = make_uint2(0u, 0u);
      a.fl[i] = make_uint2(0u, BCM_NOLEV);
    }
    if (i < lvcap) a.lv[i] = BcmLevel{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (i == 0) *a.ctl = BcmCtl{0, 0, BCM_NOLEV, 0, 0};
  }
}

// The sweep's sources claim level 0.  One wavefront.
template <int RS>
__global__ void k_bcm_src(BcmArgs a) {
  const uint32_t s = lane_id();
  if (s < a.ns) {
    const uint32_t v = a.src[s];
    a.sigma[(uint64_t)v * RS + s] = 1.0;
    atomicOr(&a.sn[v].y, 1u << s);
  }
}

template <typename T>
__device__ __forceinline__ T block_or(T v) {
  __shared__ T wo[BCM_THREADS / WAVE];
#pragma unroll
  for (int d = WAVE / 2; d >= 1; d >>= 1) v |= __shfl_xor(v, d, WAVE);
  if (lane_id() == 0) wo[threadIdx.x / WAVE] = v;
  __syncthreads();
  T t = 0;
#pragma unroll
  for (int i = 0; i < BCM_THREADS / WAVE; ++i) t |= wo[i];
  __syncthreads();
  return t;
}

// Opens level d from the claims of the previous expansion (or of the
// sources): each vertex with a non-empty next mask m is appended to the
// level queue with its items, m joins its seen mask and is recorded per
// queue entry (qmask) and per vertex (fl).  amask, fslots and the bottom-up
// candidates of the level are collected on the way: a candidate has
// in-adjacency and an unreached source among those active at level d - 1
// (a superset of those active at level d, which the bottom-up kernel uses);
// candidates with more than BCM_CH slots are listed as items.  Each
// workgroup owns a contiguous vertex range: it counts, reserves with one
// atomic per counter, then writes.  The last workgroup to finish fixes the
// level's bounds and, if the level is empty, the depth.
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_open(BcmArgs a, uint32_t d) {
  __shared__ uint32_t s_live, s_last, s_q, s_i, s_l, s_p, s_c;
  if (threadIdx.x == 0)
    s_live = a.ctl->depth == BCM_NOLEV && a.ctl->overflow == 0u;  // read once per workgroup
  __syncthreads();
  if (!s_live) return;
  const uint32_t amp = d ? a.lv[d - 1].amask : (a.ns >= 32 ? 0xFFFFFFFFu : (1u << a.ns) - 1u);
  const bool bu = a.bu_ratio > 0.f;
  const uint32_t per = (uint32_t)(((uint64_t)a.V + gridDim.x - 1) / gridDim.x + BCM_THREADS - 1) /
                       BCM_THREADS * BCM_THREADS;
  const uint32_t v0 = (uint32_t)std::min<uint64_t>((uint64_t)blockIdx.x * per, a.V);
  const uint32_t v1 = (uint32_t)std::min<uint64_t>((uint64_t)v0 + per, a.V);
  // pass 1: counts
  uint32_t nq = 0, ni = 0, nl = 0, np = 0, nc = 0, am = 0;
  unsigned long long fs = 0, cs = 0;
  for (uint32_t v = v0 + threadIdx.x; v < v1; v += BCM_THREADS) {
    const uint2 x = a.sn[v];
    if (x.y) {
      const uint32_t used = a.out.used[v], k = bcm_chunks(used);
      ++nq;
      ni += k;
      if (k > 1u) {
        ++nl;
        np += k;
      }
      am |= x.y;
      fs += used;
    }
    if (bu) {
      const uint32_t iu = a.in.used[v];
      if (iu && (amp & ~(x.x | x.y))) {
        cs += iu;
        if (iu > BCM_CH) nc += bcm_chunks(iu);
      }
    }
  }
  nq = block_sum<uint32_t, BCM_THREADS>(nq);
  ni = block_sum<uint32_t, BCM_THREADS>(ni);
  nl = block_sum<uint32_t, BCM_THREADS>(nl);
  np = block_sum<uint32_t, BCM_THREADS>(np);
  nc = block_sum<uint32_t, BCM_THREADS>(nc);
  am = block_or(am);
  fs = block_sum<unsigned long long, BCM_THREADS>(fs);
  cs = block_sum<unsigned long long, BCM_THREADS>(cs);
  if (threadIdx.x == 0) {
    uint32_t ok = 1;
    if (nq) {
      const unsigned long long qi =
          atomicAdd(&a.ctl->qi, (unsigned long long)nq | ((unsigned long long)ni << 32));
      s_q = (uint32_t)qi;
      s_i = (uint32_t)(qi >> 32);
      if ((uint64_t)s_q + nq > a.qcap || (uint64_t)s_i + ni > a.icap) ok = 0;
      if (nl) {
        s_l = atomicAdd(&a.ctl->lcur, nl);
        s_p = atomicAdd(&a.lv[d].pn, np);
        if ((uint64_t)s_l + nl > a.lcap) ok = 0;
      }
      atomicOr(&a.lv[d].amask, am);
      atomicAdd(&a.lv[d].fslots, fs);
    }
    if (nc) s_c = atomicAdd(&a.lv[d].ncand, nc);
    if (cs) atomicAdd(&a.lv[d].cslots, cs);
    if (!ok) atomicOr(&a.ctl->overflow, 1u);
    s_live = ok;
  }
  __syncthreads();
  // pass 2: writes, positions from workgroup-local cursors in LDS
  if (s_live && (nq || nc)) {
    for (uint32_t b = v0; b < v1; b += BCM_THREADS) {
      const uint32_t v = b + threadIdx.x;
      uint2 x = make_uint2(0u, 0u);
      uint32_t k = 0, kc = 0;
      if (v < v1) {
        x = a.sn[v];
        if (x.y) k = bcm_chunks(a.out.used[v]);
        if (bu) {
          const uint32_t iu = a.in.used[v];
          if (iu > BCM_CH && (amp & ~(x.x | x.y))) kc = bcm_chunks(iu);
        }
      }
      const bool q = k != 0u, multi = k > 1u;
      // wave-aggregated reservations from the workgroup's cursors
      const unsigned long long mq = __ballot(q), mm = __ballot(multi);
      const uint32_t kinc = wave_incl_scan<uint32_t>(k);
      const uint32_t kminc = wave_incl_scan<uint32_t>(multi ? k : 0u);
      const uint32_t kcinc = wave_incl_scan<uint32_t>(kc);
      uint32_t bq = 0, bi = 0, bl = 0, bp = 0, bc = 0;
      if (lane_id() == WAVE - 1) {
        if (mq) {
          bq = atomicAdd(&s_q, (uint32_t)__popcll(mq));
          bi = atomicAdd(&s_i, kinc);
        }
        if (mm) {
          bl = atomicAdd(&s_l, (uint32_t)__popcll(mm));
          bp = atomicAdd(&s_p, kminc);
        }
        if (kcinc) bc = atomicAdd(&s_c, kcinc);
      }
      bq = __shfl(bq, WAVE - 1, WAVE);
      bi = __shfl(bi, WAVE - 1, WAVE);
      bl = __shfl(bl, WAVE - 1, WAVE);
      bp = __shfl(bp, WAVE - 1, WAVE);
      bc = __shfl(bc, WAVE - 1, WAVE);
      if (q) {
        const uint32_t j = bq + mask_lt_count(mq);
        const uint32_t i0 = bi + kinc - k;
        a.queue[j] = v;
        a.qmask[j] = x.y;
        a.qpart[j] = bp + kminc - (multi ? k : 0u);
        const uint32_t fm = multi ? BCM_MULTI : 0u;
        for (uint32_t c = 0; c < k; ++c) a.items[i0 + c] = make_uint2(j, c | fm);
        if (multi) a.longl[bl + mask_lt_count(mm)] = j;
        a.sn[v] = make_uint2(x.x | x.y, 0u);
        a.fl[v] = make_uint2(x.y, d);
      }
      for (uint32_t c = 0; c < kc; ++c) a.cand[bc + kcinc - kc + c] = make_uint2(v, c | BCM_MULTI);
    }
  }
  // The last workgroup fixes the level's bounds.  It reads only the append
  // counters, through atomics, and every workgroup's reservations returned
  // before its done increment was issued, so no fence is needed (the queue
  // writes reach the next kernels at the kernel boundary); a device-scope
  // fence per workgroup would write back the L2 each time.
  __syncthreads();
  if (threadIdx.x == 0) s_last = atomicAdd(&a.lv[d].done, 1u) == gridDim.x - 1;
  __syncthreads();
  if (s_last && threadIdx.x == 0) {
    const unsigned long long qi = atomicAdd(&a.ctl->qi, 0ull);
    a.lv[d + 1].qb = (uint32_t)qi;
    a.lv[d + 1].ib = (uint32_t)(qi >> 32);
    a.lv[d + 1].lb = atomicAdd(&a.ctl->lcur, 0u);
    if ((uint32_t)qi == a.lv[d].qb) a.ctl->depth = d;
  }
}

// Top-down expansion of level d, one wavefront per item, lanes over its
// slots: for a neighbour w, cand = F & ~seen[w] are exactly the sources for
// which v -> w is a shortest-path DAG edge, so sigma[w][s] += sigma[v][s]
// for each, and the sources join next[w].
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_td(BcmArgs a, uint32_t d) {
  __shared__ double s_sig[BCM_THREADS / WAVE][RS];
  if (!bcm_live(a)) return;
  const BcmLevel L = a.lv[d];
  if (bcm_bottom_up(a, L)) return;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t ib = L.ib, ie = a.lv[d + 1].ib;
  const uint32_t nw = gridDim.x * (BCM_THREADS / WAVE);
  for (uint32_t i = ib + blockIdx.x * (BCM_THREADS / WAVE) + wid; i < ie; i += nw) {
    const uint2 it = a.items[i];
    const uint32_t v = a.queue[it.x], F = a.qmask[it.x];
    const uint32_t us = a.out.used[v], tbl = a.out.tbl_off[v];
    const uint32_t c0 = (it.y & ~BCM_MULTI) * BCM_CH;
    const uint32_t c1 = c0 + BCM_CH < us ? c0 + BCM_CH : us;
    if (lane < (uint32_t)RS) s_sig[wid][lane] = a.sigma[(uint64_t)v * RS + lane];
    wave_lds_sync();
    for (uint32_t base = c0; base < c1; base += WAVE) {
      const uint32_t pos = base + lane;
      const uint32_t w =
          pos < c1 ? a.out.slots[slot_index(a.out.arena, tbl, pos, a.out.log2B, a.out.B)] : TOMB;
      if (w != TOMB) {
        const uint2 m = a.sn[w];
        const uint32_t cand = F & ~m.x;
        if (cand) {
          if (cand & ~m.y) atomicOr(&a.sn[w].y, cand);
          for (uint32_t b = cand; b; b &= b - 1) {
            const uint32_t s = __ffs(b) - 1;
            atomicAdd(&a.sigma[(uint64_t)w * RS + s], s_sig[wid][s]);
          }
        }
      }
    }
    wave_lds_sync();
  }
}

// Bottom-up expansion of level d, RS lanes (one per source) per candidate
// item: lane s sums sigma[v][s] over the item's in-neighbours v at level d
// for source s (fl[v] = (mask, d)), in neighbour order.  A vertex with one
// item owns its row and writes it; items of a multi-item vertex add
// atomically.
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_bu(BcmArgs a, uint32_t d) {
  if (!bcm_live(a)) return;
  const BcmLevel L = a.lv[d];
  if (!bcm_bottom_up(a, L)) return;
  const uint32_t lane = lane_id();
  const uint32_t s = lane % RS, sub = lane / RS;
  const uint32_t nl = L.ncand, n = nl + a.V;
  const uint32_t per_wave = WAVE / RS;
  const uint32_t ngroups = gridDim.x * (BCM_THREADS / WAVE) * per_wave;
  for (uint32_t g0 = (blockIdx.x * (BCM_THREADS / WAVE) + threadIdx.x / WAVE) * per_wave;
       g0 < n; g0 += ngroups) {
    const uint32_t gi = g0 + sub;
    bool have = false, multi = false;
    uint32_t w = 0, U = 0, c0 = 0, c1 = 0, tbl = 0;
    if (gi < nl) {  // an item of a listed multi-item candidate
      const uint2 it = a.cand[gi];
      w = it.x;
      multi = have = true;
      U = L.amask & ~a.sn[w].x;
      const uint32_t us = a.in.used[w];
      tbl = a.in.tbl_off[w];
      c0 = (it.y & ~BCM_MULTI) * BCM_CH;
      c1 = c0 + BCM_CH < us ? c0 + BCM_CH : us;
    } else if (gi < n) {  // a vertex: a candidate if it has one item's worth of slots
      w = gi - nl;
      const uint32_t us = a.in.used[w];
      if (us && us <= BCM_CH) {
        U = L.amask & ~a.sn[w].x;
        if (U) {
          have = true;
          tbl = a.in.tbl_off[w];
          c1 = us;
        }
      }
    }
    const uint32_t span = c1 - c0;
    uint32_t span_max = span;
#pragma unroll
    for (int x = RS; x < WAVE; x <<= 1) span_max = max(span_max, __shfl_xor(span_max, x, WAVE));
    double acc = 0.0;
    for (uint32_t base = 0; base < span_max; base += RS) {
      uint32_t v = TOMB, m = 0;
      if (base + s < span) {
        v = a.in.slots[slot_index(a.in.arena, tbl, c0 + base + s, a.in.log2B, a.in.B)];
        if (v != TOMB) {
          const uint2 f = a.fl[v];
          m = f.y == d ? (f.x & U) : 0u;
        }
      }
      for (uint32_t t0 = 0; t0 < (uint32_t)RS; t0 += BCM_ILP) {
        uint32_t vv[BCM_ILP];
        bool hit[BCM_ILP];
        double sg[BCM_ILP];
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) {
          vv[k] = __shfl(v, (int)(t0 + k), RS);
          hit[k] = (__shfl(m, (int)(t0 + k), RS) >> s) & 1u;
        }
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) sg[k] = hit[k] ? a.sigma[(uint64_t)vv[k] * RS + s] : 0.0;
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) acc += sg[k];
      }
    }
    const uint32_t got = (uint32_t)((__ballot(acc != 0.0) >> (sub * RS)) & ((RS == 32) ? 0xFFFFFFFFull : 0xFFFFull));
    if (have && got) {
      if (!multi) {
        if (acc != 0.0) a.sigma[(uint64_t)w * RS + s] = acc;
        if (s == 0) a.sn[w].y = got;
      } else {
        if (acc != 0.0) atomicAdd(&a.sigma[(uint64_t)w * RS + s], acc);
        if (s == 0) atomicOr(&a.sn[w].y, got);
      }
    }
  }
}

// Backward level masks: for the queue entries of level t, sn[w] = (mask, t).
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_lmask(BcmArgs a, uint32_t t) {
  const uint32_t qb = a.lv[t].qb, qe = a.lv[t + 1].qb;
  const uint32_t st = gridDim.x * blockDim.x;
  for (uint32_t j = qb + blockIdx.x * blockDim.x + threadIdx.x; j < qe; j += st)
    a.sn[a.queue[j]] = make_uint2(a.qmask[j], t);
}

__device__ __forceinline__ void bcm_merge(float &s, float &c, float s2, float c2) {
  float t = s + s2;
  float bp = t - s;
  float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 - e;
}

// Finishes vertex v for the RS source lanes of its group: delta = sigma_v *
// S (S = the compensated successor sum), coef for the level above, and the
// dependencies of the sources that are not v itself added to BC[v] in a
// fixed lane order.
template <int RS>
__device__ __forceinline__ void bcm_finish(const BcmArgs &a, bool have, uint32_t v, bool act,
                                           double sv, float S, uint32_t s) {
  float dl = 0.f;
  if (have && act) {
    dl = (float)(sv * (double)S);
    a.coef[(uint64_t)v * RS + s] = (float)((1.0 + (double)dl) / sv);
  }
  const uint32_t my_src = s < a.ns ? a.src[s] : 0xFFFFFFFFu;
  float bs = (have && act && v != my_src) ? dl : 0.f, bcomp = 0.f;
#pragma unroll
  for (int k = RS / 2; k >= 1; k >>= 1) {
    float s2 = __shfl_xor(bs, k, RS), c2 = __shfl_xor(bcomp, k, RS);
    bcm_merge(bs, bcomp, s2, c2);
  }
  if (have && s == 0) {
    float b = a.bc[v], bcv = a.bcc[v];
    kahan_add(b, bcv, bs - bcomp);
    a.bc[v] = b;
    a.bcc[v] = bcv;
  }
}

// Backward of level d, RS lanes per item: Kahan sum over the item's
// successors (sn[w] = (mask, d + 1)) of coef[w][s].  Single-item vertices
// finish at once; multi-item vertices leave (sum, compensation) partials.
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_bwd(BcmArgs a, uint32_t d, float2 *part) {
  const uint32_t lane = lane_id();
  const uint32_t s = lane % RS, sub = lane / RS;
  const uint32_t ib = a.lv[d].ib, n = a.lv[d + 1].ib - ib;
  const uint32_t per_wave = WAVE / RS;
  const uint32_t ngroups = gridDim.x * (BCM_THREADS / WAVE) * per_wave;
  for (uint32_t g0 = (blockIdx.x * (BCM_THREADS / WAVE) + threadIdx.x / WAVE) * per_wave;
       g0 < n; g0 += ngroups) {
    const uint32_t gi = g0 + sub;
    const bool have = gi < n;
    uint32_t v = 0, tbl = 0, c0 = 0, c1 = 0, j = 0, chunk = 0;
    bool act = false, multi = false;
    double sv = 0.0;
    if (have) {
      const uint2 it = a.items[ib + gi];
      j = it.x;
      chunk = it.y & ~BCM_MULTI;
      multi = (it.y & BCM_MULTI) != 0u;
      v = a.queue[j];
      act = (a.qmask[j] >> s) & 1u;
      if (act) sv = a.sigma[(uint64_t)v * RS + s];
      const uint32_t us = a.out.used[v];
      tbl = a.out.tbl_off[v];
      c0 = chunk * BCM_CH;
      c1 = c0 + BCM_CH < us ? c0 + BCM_CH : us;
    }
    const uint32_t span = c1 - c0;
    uint32_t span_max = span;
#pragma unroll
    for (int x = RS; x < WAVE; x <<= 1) span_max = max(span_max, __shfl_xor(span_max, x, WAVE));
    float ds = 0.f, dc = 0.f;
    for (uint32_t base = 0; base < span_max; base += RS) {
      uint32_t w = TOMB, lm = 0;
      if (base + s < span) {
        w = a.out.slots[slot_index(a.out.arena, tbl, c0 + base + s, a.out.log2B, a.out.B)];
        if (w != TOMB) {
          const uint2 b = a.sn[w];
          lm = b.y == d + 1 ? b.x : 0u;
        }
      }
      for (uint32_t t0 = 0; t0 < (uint32_t)RS; t0 += BCM_ILP) {
        uint32_t ww[BCM_ILP];
        bool hit[BCM_ILP];
        float cw[BCM_ILP];
        // every lane takes part in the shuffles (a lane left out of a
        // cross-lane read gives 0 to the lanes reading it)
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) {
          ww[k] = __shfl(w, (int)(t0 + k), RS);
          hit[k] = (__shfl(lm, (int)(t0 + k), RS) >> s) & 1u;
        }
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) hit[k] = hit[k] && act;
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) cw[k] = hit[k] ? a.coef[(uint64_t)ww[k] * RS + s] : 0.f;
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k)
          if (hit[k]) kahan_add(ds, dc, cw[k]);
      }
    }
    if (have && multi) part[(uint64_t)(a.qpart[j] + chunk) * RS + s] = make_float2(ds, dc);
    // groups of a wavefront finish together; multi-item groups pass have = false
    bcm_finish<RS>(a, have && !multi, v, act, sv, ds - dc, s);
  }
}

// Multi-item vertices of level d: merge their items' partials in chunk
// order, then finish.
template <int RS>
__global__ __launch_bounds__(BCM_THREADS) void k_bcm_fin(BcmArgs a, uint32_t d,
                                                         const float2 *part) {
  const uint32_t lane = lane_id();
  const uint32_t s = lane % RS, sub = lane / RS;
  const uint32_t lb = a.lv[d].lb, n = a.lv[d + 1].lb - lb;
  const uint32_t per_wave = WAVE / RS;
  const uint32_t ngroups = gridDim.x * (BCM_THREADS / WAVE) * per_wave;
  for (uint32_t g0 = (blockIdx.x * (BCM_THREADS / WAVE) + threadIdx.x / WAVE) * per_wave;
       g0 < n; g0 += ngroups) {
    const uint32_t gi = g0 + sub;
    const bool have = gi < n;
    uint32_t v = 0;
    bool act = false;
    double sv = 0.0;
    float ds = 0.f, dc = 0.f;
    if (have) {
      const uint32_t j = a.longl[lb + gi];
      v = a.queue[j];
      act = (a.qmask[j] >> s) & 1u;
      if (act) sv = a.sigma[(uint64_t)v * RS + s];
      const uint32_t k = bcm_chunks(a.out.used[v]), pb = a.qpart[j];
      for (uint32_t c = 0; c < k; ++c) {
        const float2 p = part[(uint64_t)(pb + c) * RS + s];
        bcm_merge(ds, dc, p.x, p.y);
      }
    }
    bcm_finish<RS>(a, have, v, act, sv, ds - dc, s);
  }
}

__global__ void k_used_sum(Store st, unsigned long long *tot) {
  unsigned long long acc = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < st.V;
       v += (uint64_t)gridDim.x * blockDim.x)
    acc += st.used[v];
  acc = block_sum<unsigned long long, 256>(acc);
  if (threadIdx.x == 0 && acc) atomicAdd(tot, acc);
}

void bcm_release(BcmWork &w) {
  hipFree(w.sigma);
  hipFree(w.coef);
  hipFree(w.bcc);
  hipFree(w.sn);
  hipFree(w.fl);
  hipFree(w.queue);
  hipFree(w.qmask);
  hipFree(w.qpart);
  hipFree(w.items);
  hipFree(w.longl);
  hipFree(w.cand);
  hipFree(w.part);
  hipFree(w.lv);
  hipFree(w.ctl);
  hipFree(w.dsrc);
  hipFree(w.used_tot);
  if (w.h_ctl) hipHostFree(w.h_ctl);
  w = BcmWork{};
}

// (Re)allocates buf for cap elements of sz bytes when it holds fewer.
static hipError_t bcm_grow(void **buf, uint64_t &have, uint64_t cap, size_t sz) {
  if (cap <= have && *buf) return hipSuccess;
  hipFree(*buf);
  *buf = nullptr;
  have = 0;
  hipError_t e = hipMalloc(buf, (size_t)(cap ? cap : 1) * sz);
  if (e == hipSuccess) have = cap;
  return e;
}

// Level counter array for at least n levels, keeping the first keep entries.
static gv_status bcm_levels(BcmWork &w, uint64_t n, uint64_t keep, hipStream_t s) {
  if (n <= w.lvcap) return GV_OK;
  uint64_t cap = w.lvcap ? w.lvcap : 64;
  while (cap < n) cap *= 2;
  void *nb = nullptr;
  hipError_t e = hipMalloc(&nb, cap * sizeof(BcmLevel));
  if (e == hipSuccess && keep)
    e = hipMemcpyAsync(nb, w.lv, keep * sizeof(BcmLevel), hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) {
    // counters of the levels not yet reached start at zero
    static_assert(sizeof(BcmLevel) % 4 == 0, "level counters are cleared as 32-bit words");
    fill_u32((uint32_t *)nb + keep * (sizeof(BcmLevel) / 4), 0u,
             (cap - keep) * (sizeof(BcmLevel) / 4), s);
    e = hipStreamSynchronize(s);
  }
  if (e != hipSuccess) {
    hipFree(nb);
    return hip_fail(e, "bc levels");
  }
  hipFree(w.lv);
  w.lv = nb;
  w.lvcap = cap;
  return GV_OK;
}

template <int RS>
static void bcm_forward_group(const BcmArgs &a, uint32_t d0, hipStream_t s) {
  const unsigned gv = grid_for(a.V, BCM_THREADS, 2048), ge = 2048;
  for (uint32_t d = d0; d < d0 + BCM_GROUP; ++d) {
    k_bcm_open<RS><<<gv, BCM_THREADS, 0, s>>>(a, d);
    k_bcm_td<RS><<<ge, BCM_THREADS, 0, s>>>(a, d);
    k_bcm_bu<RS><<<ge, BCM_THREADS, 0, s>>>(a, d);
  }
}

template <int RS>
static void bcm_backward(const BcmArgs &a, uint32_t D, hipStream_t s) {
  const unsigned gq = 1024, ge = 2048;
  // level 0 holds only the sweep's sources, whose own dependencies are not
  // counted, so the sweep stops at level 1
  for (uint32_t d = D; d-- > 1;) {
    if (d + 1 < D) k_bcm_lmask<RS><<<gq, BCM_THREADS, 0, s>>>(a, d + 1);
    k_bcm_bwd<RS><<<ge, BCM_THREADS, 0, s>>>(a, d, a.part);
    k_bcm_fin<RS><<<gq, BCM_THREADS, 0, s>>>(a, d, a.part);
  }
}

template <int RS>
static gv_status bcm_sweeps(const Store &out, const Store *in, const uint32_t *sources,
                            uint32_t n_sources, float *out_dev, BcmWork &w, float bu_ratio,
                            uint64_t h_out, uint64_t h_in, hipStream_t s) {
  const uint32_t V = out.V;
  const uint64_t ch_out = h_out / BCM_CH + 1, ch_in = h_in / BCM_CH + 1;
  // bounds: a vertex is queued at most once per level and at most RS times
  // (once per source); per level its items are at most chunks of its slots
  const uint64_t qmax = (uint64_t)V * RS, imax = (uint64_t)RS * (V + ch_out),
                 lmax = (uint64_t)RS * ch_out;
  if (imax + V + ch_out > 0xFFFFFFFFull) {
    set_error("bc: graph too large for 32-bit work item positions");
    return GV_E_INVALID;
  }
  hipError_t e = hipSuccess;
  if ((e = bcm_grow((void **)&w.sigma, w.n_sigma, (uint64_t)V * RS, sizeof(double))) ||
      (e = bcm_grow((void **)&w.coef, w.n_coef, (uint64_t)V * RS, sizeof(float))) ||
      (e = bcm_grow((void **)&w.bcc, w.n_bcc, V, sizeof(float))) ||
      (e = bcm_grow((void **)&w.sn, w.n_sn, V, sizeof(uint2))) ||
      (e = bcm_grow((void **)&w.fl, w.n_fl, V, sizeof(uint2))) ||
      (e = bcm_grow((void **)&w.part, w.n_part, (2 * ch_out + 2) * RS, sizeof(float2))) ||
      (e = bcm_grow((void **)&w.cand, w.n_cand, in ? 2 * ch_in + 2 : 1, sizeof(uint2))) ||
      (e = bcm_grow((void **)&w.dsrc, w.n_dsrc, BCM_S, sizeof(uint32_t))) ||
      (e = bcm_grow((void **)&w.ctl, w.n_ctl, 1, sizeof(BcmCtl))))
    return hip_fail(e, "bc alloc");
  if (!w.h_ctl && (e = hipHostMalloc(&w.h_ctl, sizeof(BcmCtl), hipHostMallocDefault)))
    return hip_fail(e, "bc alloc");
  uint64_t qcap = std::min<uint64_t>(qmax, std::max<uint64_t>(w.n_queue, 4ull * V + 64));
  uint64_t icap = std::min<uint64_t>(imax, std::max<uint64_t>(w.n_items, 4ull * (V + ch_out) + 64));
  uint64_t lcap = std::min<uint64_t>(lmax, std::max<uint64_t>(w.n_longl, 2 * ch_out + 64));
  gv_status rc = GV_OK;
  BcmArgs a{};
  a.out = out;
  a.in = in ? *in : out;
  a.bc = out_dev;
  a.ns = 0;
  a.V = V;
  a.bu_ratio = in ? bu_ratio : 0.f;
  fill_f32(out_dev, 0.f, V, s);
  fill_f32(w.bcc, 0.f, V, s);
  for (uint32_t b0 = 0; b0 < n_sources && rc == GV_OK;) {
    const uint32_t ns = std::min<uint32_t>(RS, n_sources - b0);
    if ((e = bcm_grow((void **)&w.queue, w.n_queue, qcap, 4)) ||
        (e = bcm_grow((void **)&w.qmask, w.n_qmask, qcap, 4)) ||
        (e = bcm_grow((void **)&w.qpart, w.n_qpart, qcap, 4)) ||
        (e = bcm_grow((void **)&w.items, w.n_items, icap, sizeof(uint2))) ||
        (e = bcm_grow((void **)&w.longl, w.n_longl, lcap, 4)))
      return hip_fail(e, "bc alloc");
    if ((rc = bcm_levels(w, BCM_GROUP + 2, 0, s))) return rc;
    a.sigma = w.sigma;
    a.coef = w.coef;
    a.bcc = w.bcc;
    a.sn = (uint2 *)w.sn;
    a.fl = (uint2 *)w.fl;
    a.queue = w.queue;
    a.qmask = w.qmask;
    a.qpart = w.qpart;
    a.items = (uint2 *)w.items;
    a.longl = w.longl;
    a.cand = (uint2 *)w.cand;
    a.part = (float2 *)w.part;
    a.ctl = (BcmCtl *)w.ctl;
    a.src = w.dsrc;
    a.ns = ns;
    a.qcap = (uint32_t)std::min<uint64_t>(w.n_queue, 0xFFFFFFFFull);
    a.icap = (uint32_t)std::min<uint64_t>(w.n_items, 0xFFFFFFFFull);
    a.lcap = (uint32_t)std::min<uint64_t>(w.n_longl, 0xFFFFFFFFull);
    if ((e = hipMemcpyAsync(w.dsrc, sources + b0, ns * sizeof(uint32_t), hipMemcpyHostToDevice,
                            s)))
      return hip_fail(e, "bc");
    a.lv = (BcmLevel *)w.lv;
    {
      const uint64_t n = std::max<uint64_t>((uint64_t)V * RS, w.lvcap);
      k_bcm_init<RS><<<grid_for(n, 256, 16384), 256, 0, s>>>(a, (uint32_t)w.lvcap);
    }
    k_bcm_src<RS><<<1, WAVE, 0, s>>>(a);
    uint32_t D = BCM_NOLEV;
    bool overflow = false;
    for (uint32_t d0 = 0; D == BCM_NOLEV; d0 += BCM_GROUP) {
      if ((rc = bcm_levels(w, (uint64_t)d0 + BCM_GROUP + 2, d0 + 1, s))) return rc;
      a.lv = (BcmLevel *)w.lv;
      bcm_forward_group<RS>(a, d0, s);
      if ((e = hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(BcmCtl), hipMemcpyDeviceToHost, s)) ||
          (e = hipStreamSynchronize(s)))
        return hip_fail(e, "bc");
      const BcmCtl *hc = (const BcmCtl *)w.h_ctl;
      if (hc->overflow) {
        overflow = true;
        break;
      }
      D = hc->depth;
    }
    if (overflow) {  // rerun this sweep with larger buffers
      if (qcap >= qmax && icap >= imax && lcap >= lmax) {
        set_error("bc: work buffers at their bound overflowed");
        return GV_E_HIP;
      }
      qcap = std::min(qmax, 2 * qcap);
      icap = std::min(imax, 2 * icap);
      lcap = std::min(lmax, 2 * lcap);
      continue;
    }
    bcm_backward<RS>(a, D, s);
    b0 += ns;
  }
  if (rc == GV_OK && (e = hipStreamSynchronize(s))) rc = hip_fail(e, "bc");
  return rc;
}

gv_status bc_multi_run(const Store &out, const Store *in, const uint32_t *sources,
                       uint32_t n_sources, float *out_dev, BcmWork &w, float bu_ratio,
                       hipStream_t s) {
  const uint32_t V = out.V;
  if (!V) return GV_OK;
  hipError_t e = hipSuccess;
  if ((e = bcm_grow((void **)&w.used_tot, w.n_used_tot, 2, sizeof(unsigned long long))))
    return hip_fail(e, "bc alloc");
  fill_u64((uint64_t *)w.used_tot, 0, 2, s);
  k_used_sum<<<grid_for(V, 256, 1024), 256, 0, s>>>(out, w.used_tot);
  if (in) k_used_sum<<<grid_for(V, 256, 1024), 256, 0, s>>>(*in, w.used_tot + 1);
  unsigned long long h[2] = {0, 0};
  if ((e = hipMemcpyAsync(h, w.used_tot, sizeof(h), hipMemcpyDeviceToHost, s)) ||
      (e = hipStreamSynchronize(s)))
    return hip_fail(e, "bc");
  return n_sources <= 16
             ? bcm_sweeps<16>(out, in, sources, n_sources, out_dev, w, bu_ratio, h[0], h[1], s)
             : bcm_sweeps<32>(out, in, sources, n_sources, out_dev, w, bu_ratio, h[0], h[1], s);
}

}  // namespace gv
