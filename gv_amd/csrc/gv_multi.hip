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
constexpr int BCM_S = 32;   // sources per batch = lanes per frontier vertex
constexpr int BCM_ILP = 8;  // neighbour rows requested before the first is used

// Per-vertex source masks of the forward sweep: seen (x) = sources that
// reached the vertex at an earlier level, next (y) = sources reaching it at
// the level being expanded.
__global__ void k_bcm_init(int *dist, double *sigma, float *delta, uint2 *sn, uint32_t V) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (uint64_t)V * BCM_S;
       i += (uint64_t)gridDim.x * blockDim.x) {
    dist[i] = -1;
    sigma[i] = 0.0;
    delta[i] = 0.f;
    if (i < V) sn[i] = make_uint2(0u, 0u);
  }
}

// Sources of the batch enter level 0; a vertex holding several sources is
// queued once.
__global__ void k_bcm_src(const uint32_t *src, uint32_t ns, int *dist, double *sigma, uint2 *sn,
                          uint32_t *queue, unsigned int *cursor) {
  const uint32_t s = threadIdx.x;
  if (s >= ns) return;
  const uint32_t v = src[s];
  dist[(uint64_t)v * BCM_S + s] = 0;
  sigma[(uint64_t)v * BCM_S + s] = 1.0;
  if (atomicOr(&sn[v].x, 1u << s) == 0u) queue[atomicAdd(cursor, 1u)] = v;
}

// After a level: the new frontier's level bits join its seen mask.
__global__ void k_bcm_next(const uint32_t *queue, uint32_t qe, const unsigned int *cursor,
                           uint2 *sn) {
  const uint32_t n = *cursor;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t w = queue[qe + i];
    const uint2 m = sn[w];
    sn[w] = make_uint2(m.x | m.y, 0u);
  }
}

// Work items: a frontier vertex with more than BCM_CH slots is split into
// chunks of BCM_CH slots, each its own item, so a hub's adjacency is walked
// by many lane groups at once.  cnt/coff hold the per-vertex item counts and
// their exclusive scan (coff[nf] = items of the level).
constexpr uint32_t BCM_CH = 256;

__global__ void k_bcm_count(Store st, const uint32_t *q, uint32_t nf, uint32_t *cnt) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nf;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = st.used[q[i]];
    cnt[i] = u > BCM_CH ? (u + BCM_CH - 1) / BCM_CH : 1u;
  }
}

// owner[i] = frontier index j of item i (coff[j] <= i < coff[j + 1]), so a
// wavefront finds its item's vertex with one load.
__global__ void k_bcm_owner(const uint32_t *coff, uint32_t nf, uint32_t *owner) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nf;
       j += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = coff[j], e = coff[j + 1]; i < e; ++i) owner[i] = (uint32_t)j;
}

// One item per lane group: the group's vertex, its slot range [c0, c1), and
// the lane's (dist, sigma) for source s.
struct BcmItem {
  uint32_t v, tbl, c0, c1;
  int dv;
  double sv;
};

__device__ __forceinline__ BcmItem bcm_item(const Store &st, const int *dist, const double *sigma,
                                            const uint32_t *queue, uint32_t qs,
                                            const uint32_t *coff, const uint32_t *owner,
                                            uint64_t gi, bool have, uint32_t s) {
  BcmItem it{0, 0, 0, 0, -2, 0.0};
  if (have) {
    const uint32_t j = owner[gi];
    it.v = queue[qs + j];
    const uint32_t us = st.used[it.v];
    it.tbl = st.tbl_off[it.v];
    it.c0 = ((uint32_t)gi - coff[j]) * BCM_CH;
    it.c1 = it.c0 + BCM_CH < us ? it.c0 + BCM_CH : us;
    it.dv = dist[(uint64_t)it.v * BCM_S + s];
    it.sv = sigma[(uint64_t)it.v * BCM_S + s];
  }
  return it;
}

// Forward, one wavefront per item with its lanes over the item's slots
// (bit-parallel multi-source BFS, Then et al., VLDB 2014): F = the sources
// for which v is at level d (dist[v][s] == d); for a neighbour w, cand =
// F & ~seen[w] are exactly the sources whose shortest-path DAG holds v -> w,
// so sigma[w][s] += sigma[v][s] for each.  The first claim of a bit at this
// level (atomicOr on next[w]) sets dist[w][s] = d + 1, and the claim that
// makes next[w] non-zero queues w.  One 8-byte mask load per edge instead
// of a 32-source distance row.
__global__ __launch_bounds__(256) void k_bcm_fwd(Store st, int *dist, double *sigma, uint2 *sn,
                                                 uint32_t *queue, uint32_t qs, uint32_t qe, int d,
                                                 unsigned int *cursor, const uint32_t *coff,
                                                 const uint32_t *owner) {
  __shared__ double s_sig[4][BCM_S];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint32_t nf = qe - qs;
  const uint32_t nitems = coff[nf];
  for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; it < nitems;
       it += nwaves) {
    const uint32_t j = owner[it];
    const uint32_t v = queue[qs + j];
    const uint32_t us = st.used[v], tbl = st.tbl_off[v];
    const uint32_t c0 = ((uint32_t)it - coff[j]) * BCM_CH;
    const uint32_t c1 = c0 + BCM_CH < us ? c0 + BCM_CH : us;
    const int dv = lane < (uint32_t)BCM_S ? dist[(uint64_t)v * BCM_S + lane] : -1;
    const uint32_t F = (uint32_t)__ballot(lane < (uint32_t)BCM_S && dv == d);
    if (F == 0u) continue;
    if (lane < (uint32_t)BCM_S) s_sig[wid][lane] = sigma[(uint64_t)v * BCM_S + lane];
    wave_lds_sync();
    for (uint32_t base = c0; base < c1; base += WAVE) {
      const uint32_t pos = base + lane;
      const uint32_t w =
          pos < c1 ? st.slots[slot_index(st.arena, tbl, pos, st.log2B, st.B)] : TOMB;
      bool enq = false;
      if (w != TOMB) {
        const uint2 m = sn[w];
        const uint32_t cand = F & ~m.x;
        if (cand) {
          uint32_t claimed = 0u;
          if (cand & ~m.y) {  // some bit not claimed yet at this level
            const uint32_t old = atomicOr(&sn[w].y, cand);
            claimed = cand & ~old;
            enq = old == 0u;
          }
          for (uint32_t b = claimed; b; b &= b - 1)
            dist[(uint64_t)w * BCM_S + (__ffs(b) - 1)] = d + 1;
          for (uint32_t b = cand; b; b &= b - 1) {
            const uint32_t s = __ffs(b) - 1;
            atomicAdd(&sigma[(uint64_t)w * BCM_S + s], s_sig[wid][s]);
          }
        }
      }
      const unsigned long long qm = __ballot(enq);
      if (qm) {
        uint32_t qb = 0;
        if (lane == 0) qb = atomicAdd(cursor, (unsigned int)__popcll(qm));
        qb = __shfl(qb, 0, WAVE);
        if (enq) queue[qe + qb + mask_lt_count(qm)] = w;
      }
    }
    wave_lds_sync();
  }
}

// Backward level masks: next[w] (free after the forward sweep) = the sources
// for which w is at level lev, for the queue entries [qs, qe) of that level
// (set), or 0 again (clear).  Two entries per wavefront, lane = source.
__global__ __launch_bounds__(256) void k_bcm_lmask(const uint32_t *queue, uint32_t qs,
                                                   uint32_t qe, int lev, const int *dist,
                                                   uint2 *sn, int set) {
  const uint32_t lane = lane_id();
  const uint32_t s = lane % BCM_S, sub = lane / BCM_S;
  const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / BCM_S;
  const uint32_t n = qe - qs;
  for (uint64_t g0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE * (WAVE / BCM_S);
       g0 < n; g0 += ngroups) {
    const uint64_t j = g0 + sub;
    const bool have = j < n;
    const uint32_t w = have ? queue[qs + j] : 0u;
    const bool bit = have && set && dist[(uint64_t)w * BCM_S + s] == lev;
    const uint32_t m = (uint32_t)((__ballot(bit) >> (sub * BCM_S)) & 0xFFFFFFFFull);
    if (have && s == 0) sn[w].y = m;
  }
}

__device__ __forceinline__ void bcm_merge(float &s, float &c, float s2, float c2) {
  float t = s + s2;
  float bp = t - s;
  float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 - e;
}

// Backward, per item: Kahan sum over the item's successors of
// sigma_v/sigma_w (1 + delta_w), written as a (sum, compensation) partial.
// Successors at level d + 1 are found through the level masks (next[w],
// set by k_bcm_lmask): one 4-byte mask per neighbour, shuffled to the 32
// source lanes, instead of a 32-source distance row.
__global__ __launch_bounds__(256) void k_bcm_bwd(Store st, const int *dist, const double *sigma,
                                                 const float *delta, const uint2 *sn,
                                                 const uint32_t *queue, uint32_t qs, uint32_t qe,
                                                 int d, const uint32_t *coff,
                                                 const uint32_t *owner, float2 *part) {
  const uint32_t lane = lane_id();
  const uint32_t s = lane % BCM_S, sub = lane / BCM_S;
  const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / BCM_S;
  const uint32_t nf = qe - qs;
  const uint32_t nitems = coff[nf];
  for (uint64_t g0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE * (WAVE / BCM_S);
       g0 < nitems; g0 += ngroups) {
    const uint64_t gi = g0 + sub;
    const bool have = gi < nitems;
    const BcmItem it = bcm_item(st, dist, sigma, queue, qs, coff, owner, gi, have, s);
    const bool act = it.dv == d;
    float ds = 0.f, dc = 0.f;
    const uint32_t span = it.c1 - it.c0;
    const uint32_t span_max = max(span, __shfl_xor(span, BCM_S, WAVE));
    for (uint32_t base = 0; base < span_max; base += BCM_S) {
      uint32_t slot = TOMB;
      if (base + s < span)
        slot = st.slots[slot_index(st.arena, it.tbl, it.c0 + base + s, st.log2B, st.B)];
      const uint32_t lm = slot != TOMB ? sn[slot].y : 0u;  // this slot's level-(d+1) sources
      for (uint32_t t0 = 0; t0 < (uint32_t)BCM_S; t0 += BCM_ILP) {
        // rows of BCM_ILP neighbours in flight; sums stay in neighbour order
        uint32_t w[BCM_ILP];
        bool hit[BCM_ILP];
        double sw[BCM_ILP];
        float dw[BCM_ILP];
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) {
          w[k] = __shfl(slot, (int)(t0 + k), BCM_S);
          hit[k] = (__shfl(lm, (int)(t0 + k), BCM_S) >> s) & 1u;
        }
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) hit[k] = hit[k] && act && w[k] != TOMB;
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k) {
          const uint64_t wi = (uint64_t)w[k] * BCM_S + s;
          sw[k] = hit[k] ? sigma[wi] : 1.0;
          dw[k] = hit[k] ? delta[wi] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < BCM_ILP; ++k)
          if (hit[k]) kahan_add(ds, dc, (float)(it.sv / sw[k]) * (1.0f + dw[k]));
      }
    }
    if (have) part[gi * BCM_S + s] = make_float2(ds, dc);
  }
}

// Per frontier vertex: merge its items' partials in chunk order, store
// delta, then add the S dependencies to BC[v] in a fixed lane order.
__global__ __launch_bounds__(256) void k_bcm_fin(const uint32_t *queue, uint32_t qs, uint32_t qe,
                                                 int d, const int *dist, float *delta,
                                                 const float2 *part, const uint32_t *coff,
                                                 const uint32_t *src, uint32_t ns, float *bc,
                                                 float *bcc) {
  const uint32_t lane = lane_id();
  const uint32_t s = lane % BCM_S, sub = lane / BCM_S;
  const uint64_t ngroups = (uint64_t)gridDim.x * blockDim.x / BCM_S;
  const uint32_t nf = qe - qs;
  const uint32_t my_src = s < ns ? src[s] : 0xFFFFFFFFu;
  for (uint64_t g0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE * (WAVE / BCM_S);
       g0 < nf; g0 += ngroups) {
    const uint64_t j = g0 + sub;
    const bool have = j < nf;
    uint32_t v = 0;
    bool act = false;
    float ds = 0.f, dc = 0.f;
    if (have) {
      v = queue[qs + j];
      act = dist[(uint64_t)v * BCM_S + s] == d;
      for (uint32_t i = coff[j]; i < coff[j + 1]; ++i) {
        const float2 p = part[(uint64_t)i * BCM_S + s];
        bcm_merge(ds, dc, p.x, p.y);
      }
    }
    const float dvl = act ? ds - dc : 0.f;
    if (act) delta[(uint64_t)v * BCM_S + s] = dvl;
    float bs = (act && v != my_src) ? dvl : 0.f, bcomp = 0.f;
#pragma unroll
    for (int k = BCM_S / 2; k >= 1; k >>= 1) {
      float s2 = __shfl_xor(bs, k, BCM_S), c2 = __shfl_xor(bcomp, k, BCM_S);
      bcm_merge(bs, bcomp, s2, c2);
    }
    if (have && s == 0) {
      float b = bc[v], bcv = bcc[v];
      kahan_add(b, bcv, bs - bcomp);
      bc[v] = b;
      bcc[v] = bcv;
    }
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
  hipFree(w.dist);
  hipFree(w.sigma);
  hipFree(w.delta);
  hipFree(w.bcc);
  hipFree(w.part);
  hipFree(w.owner);
  hipFree(w.sn);
  hipFree(w.queue);
  hipFree(w.cnt);
  hipFree(w.coff);
  hipFree(w.tmp);
  hipFree(w.dsrc);
  hipFree(w.cursor);
  hipFree(w.used_tot);
  if (w.h_cursor) hipHostFree(w.h_cursor);
  w = BcmWork{};
}

// Per-vertex state for V vertices, level items for max_items.
static gv_status bcm_reserve(BcmWork &w, uint32_t V, uint64_t max_items) {
  hipError_t e = hipSuccess;
  if (!w.used_tot) {
    if ((e = hipMalloc(&w.used_tot, sizeof(unsigned long long))) != hipSuccess ||
        (e = hipMalloc(&w.dsrc, BCM_S * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&w.cursor, sizeof(unsigned int))) != hipSuccess ||
        (e = hipHostMalloc(&w.h_cursor, sizeof(unsigned int), hipHostMallocDefault)) != hipSuccess)
      return hip_fail(e, "bc alloc");
  }
  if (V > w.V) {
    hipFree(w.dist);
    hipFree(w.sigma);
    hipFree(w.delta);
    hipFree(w.bcc);
    hipFree(w.sn);
    hipFree(w.queue);
    hipFree(w.cnt);
    hipFree(w.coff);
    hipFree(w.tmp);
    w.dist = nullptr;
    w.sigma = nullptr;
    w.delta = w.bcc = nullptr;
    w.sn = nullptr;
    w.queue = w.cnt = w.coff = nullptr;
    w.tmp = nullptr;
    w.V = 0;
    const uint64_t n2 = (uint64_t)V * BCM_S;
    if ((e = hipMalloc(&w.dist, n2 * sizeof(int))) != hipSuccess ||
        (e = hipMalloc(&w.sigma, n2 * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&w.delta, n2 * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&w.bcc, (size_t)V * sizeof(float))) != hipSuccess ||
        (e = hipMalloc(&w.sn, (size_t)V * sizeof(uint2))) != hipSuccess ||
        (e = hipMalloc(&w.queue, n2 * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&w.cnt, ((size_t)V + 1) * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&w.coff, ((size_t)V + 1) * sizeof(uint32_t))) != hipSuccess ||
        (e = hipMalloc(&w.tmp, scan_tmp_bytes((uint64_t)V + 1))) != hipSuccess)
      return hip_fail(e, "bc alloc");
    w.V = V;
  }
  if (max_items > w.items) {
    hipFree(w.part);
    hipFree(w.owner);
    w.part = nullptr;
    w.owner = nullptr;
    w.items = 0;
    const uint64_t cap = max_items + max_items / 4;  // room for graphs that grow
    if ((e = hipMalloc(&w.part, cap * BCM_S * sizeof(float2))) != hipSuccess ||
        (e = hipMalloc(&w.owner, cap * sizeof(uint32_t))) != hipSuccess)
      return hip_fail(e, "bc alloc");
    w.items = cap;
  }
  return GV_OK;
}

gv_status bc_multi_run(const Store &st, const uint32_t *sources, uint32_t n_sources,
                       float *out_dev, BcmWork &w, hipStream_t s) {
  const uint32_t V = st.V;
  if (!V) return GV_OK;
  unsigned long long h_used = 0;
  gv_status rc = GV_OK;
  const uint64_t n2 = (uint64_t)V * BCM_S;
  do {
    hipError_t e = hipSuccess;
    if ((rc = bcm_reserve(w, V, 0))) break;
    fill_u64((uint64_t *)w.used_tot, 0, 1, s);
    k_used_sum<<<grid_for(V, 256, 1024), 256, 0, s>>>(st, w.used_tot);
    if ((e = hipMemcpyAsync(&h_used, w.used_tot, sizeof(h_used), hipMemcpyDeviceToHost, s)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
      rc = hip_fail(e, "bc");
      break;
    }
    // items of one level: at most one per frontier vertex plus one per BCM_CH slots
    const uint64_t max_items = (uint64_t)V + h_used / BCM_CH + 1;
    if (max_items > 0xFFFFFFFFull) {
      set_error("bc: too many work items for 32-bit offsets");
      rc = GV_E_INVALID;
      break;
    }
    if ((rc = bcm_reserve(w, V, max_items))) break;
    int *dist = w.dist;
    double *sigma = w.sigma;
    float *delta = w.delta, *bcc = w.bcc;
    float2 *part = w.part;
    uint2 *sn = w.sn;
    uint32_t *queue = w.queue, *cnt = w.cnt, *coff = w.coff, *dsrc = w.dsrc, *owner = w.owner;
    unsigned int *cursor = w.cursor, *h_cursor = w.h_cursor;
    void *tmp = w.tmp;
    // item offsets of the frontier segment [qs, qe); a level's frontier holds
    // each vertex at most once, so nf <= V
    auto level_items = [&](uint32_t qs, uint32_t qe) {
      const uint32_t nf = qe - qs;
      k_bcm_count<<<grid_for(nf), 256, 0, s>>>(st, queue + qs, nf, cnt);
      scan_u32(cnt, coff, nf, nullptr, tmp, s);
      k_bcm_owner<<<grid_for(nf), 256, 0, s>>>(coff, nf, owner);
    };
    auto item_grid = [&](uint32_t nf) {
      const uint64_t bound = (uint64_t)nf + h_used / BCM_CH + 1;
      return grid_for(bound * BCM_S, 256, 8192);
    };
    auto fwd_grid = [&](uint32_t nf) {  // one wavefront per item
      const uint64_t bound = (uint64_t)nf + h_used / BCM_CH + 1;
      return grid_for(bound * WAVE, 256, 8192);
    };
    fill_f32(out_dev, 0.f, V, s);
    fill_f32(bcc, 0.f, V, s);
    for (uint32_t b0 = 0; b0 < n_sources && rc == GV_OK; b0 += BCM_S) {
      const uint32_t ns = std::min<uint32_t>(BCM_S, n_sources - b0);
      if ((e = hipMemcpyAsync(dsrc, sources + b0, ns * sizeof(uint32_t), hipMemcpyHostToDevice,
                              s)) != hipSuccess) {
        rc = hip_fail(e, "bc");
        break;
      }
      k_bcm_init<<<grid_for(n2, 256, 16384), 256, 0, s>>>(dist, sigma, delta, sn, V);
      fill_u32(cursor, 0, 1, s);
      k_bcm_src<<<1, BCM_S, 0, s>>>(dsrc, ns, dist, sigma, sn, queue, cursor);
      if ((e = hipMemcpyAsync(h_cursor, cursor, sizeof(unsigned), hipMemcpyDeviceToHost, s)) !=
              hipSuccess ||
          (e = hipStreamSynchronize(s)) != hipSuccess) {
        rc = hip_fail(e, "bc");
        break;
      }
      std::vector<uint32_t> lvl{0u, *h_cursor};
      int d = 0;
      while (true) {
        const uint32_t qs = lvl[d], qe = lvl[d + 1];
        if (qe == qs) break;
        fill_u32(cursor, 0, 1, s);
        level_items(qs, qe);
        k_bcm_fwd<<<fwd_grid(qe - qs), 256, 0, s>>>(st, dist, sigma, sn, queue, qs, qe, d, cursor,
                                                    coff, owner);
        k_bcm_next<<<grid_for(qe - qs + h_used / BCM_CH + 1, 256, 4096), 256, 0, s>>>(queue, qe,
                                                                                   cursor, sn);
        if ((e = hipMemcpyAsync(h_cursor, cursor, sizeof(unsigned), hipMemcpyDeviceToHost, s)) !=
                hipSuccess ||
            (e = hipStreamSynchronize(s)) != hipSuccess) {
          rc = hip_fail(e, "bc");
          break;
        }
        lvl.push_back(qe + *h_cursor);
        ++d;
      }
      if (rc) break;
      // queue[lvl[k], lvl[k+1]) = union frontier of level k: every vertex
      // that some source of the batch reaches at depth k (a vertex appears
      // once per distinct depth).  The deepest level has
      // no successors, so the backward sweep starts one level above it.
      const int D = (int)lvl.size() - 2;
      auto lmask = [&](int lev, int set) {
        const uint32_t a = lvl[lev], b = lvl[lev + 1];
        if (b > a)
          k_bcm_lmask<<<grid_for((uint64_t)(b - a) * BCM_S, 256, 8192), 256, 0, s>>>(
              queue, a, b, lev, dist, sn, set);
      };
      for (int lev = D - 2; lev >= 0; --lev) {
        const uint32_t qs = lvl[lev], qe = lvl[lev + 1];
        level_items(qs, qe);
        lmask(lev + 1, 1);
        k_bcm_bwd<<<item_grid(qe - qs), 256, 0, s>>>(st, dist, sigma, delta, sn, queue, qs, qe,
                                                     lev, coff, owner, part);
        lmask(lev + 1, 0);
        k_bcm_fin<<<grid_for((uint64_t)(qe - qs) * BCM_S, 256, 8192), 256, 0, s>>>(
            queue, qs, qe, lev, dist, delta, part, coff, dsrc, ns, out_dev, bcc);
      }
    }
    if (rc == GV_OK && (hipStreamSynchronize(s) != hipSuccess)) rc = hip_fail(hipGetLastError(), "bc");
  } while (0);
  return rc;
}

}  // namespace gv
