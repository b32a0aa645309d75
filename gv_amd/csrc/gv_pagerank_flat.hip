// gv_pagerank_flat.hip -- PageRank iterations as a flat stream over the
// relabelled in-adjacency snapshot (same semantics and stopping rule as
// static_pagerank, pagerank.cu:3-68, pr_kernel.cu:21-45).
//
// Why a flat stream: with one row per wavefront, every row costs a chain of
// dependent loads (row metadata, then its entries, then their gathers)
// before its sum is known; at R-MAT scale 22 that chain, not HBM bandwidth,
// bounded the pull.  Here the row-major entry stream is cut into tiles of
// PF_T = 1024 entries; wavefront w of W takes tiles w, w + W, ..., lane l
// owns 16 consecutive entries of each tile, and the next tile's entries are
// loaded while the current tile's gathers are consumed.  Every tile holds
// the same number of entries whatever the degree distribution, so the work
// splits evenly with no scheduling state at all.
//
// Row sums: each lane Kahan-sums its entries in stream order, cutting at the
// row-start flags; a segmented TwoSum scan across the 64 lanes joins the
// pieces of rows that cross lanes.  A row that ends inside the tile in which
// it started is written to S[row]; the partial of a tile's first row (begun
// in an earlier tile) goes to P[t] and of its last row (may continue) to
// Q[t].  k_pf_rows then forms each row's sum as S, or Q[t0] + P[t0+1..t1] in
// tile order, applies the update and writes the residual partials.  Every
// sum has a fixed order, so results are bit-for-bit deterministic.
//
// Gathers: contributions of labels below nh (the highest out-degrees) are
// staged in LDS per workgroup and iteration; the rest are raw buffer loads,
// whose offset is pushed out of range for LDS-served and padding entries so
// that only those entries touch L1/L2.
#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {

constexpr int PF_THREADS = 1024;
constexpr uint32_t PF_HOT = 40960;  // contributions staged in LDS (160 KiB, all of it)
constexpr uint32_t PF_OOB = 0x80000000u;
constexpr int PF_ROW_THREADS = 256;
constexpr uint32_t PF_ROW_BLOCKS = 1024;
// rows spanning at least PF_LONG tiles are assembled by a wavefront each
constexpr uint32_t PF_LONG = 8;
constexpr uint32_t PF_LONG_BLOCKS = 512;

// ------------------------------------------------------- layout build --
__global__ void k_pf_pad(uint32_t *cols, uint64_t E, uint64_t Epad) {
  for (uint64_t p = E + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < Epad;
       p += (uint64_t)gridDim.x * blockDim.x)
    cols[pf_phys(p)] = PF_PAD;
}

__global__ void k_pf_nzflag(const uint64_t *off, uint32_t V, uint32_t *flag) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x)
    flag[r] = off[r + 1] > off[r] ? 1u : 0u;
}

// nzidx = exclusive scan of the nonempty flags (nzidx[V] = count).
__global__ void k_pf_nzfill(const uint64_t *off, const uint32_t *nzidx, uint32_t V,
                            uint32_t *nzlabel, uint32_t *nzoff, uint32_t *empty) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t i = nzidx[r];
    if (off[r + 1] > off[r]) {
      nzlabel[i] = (uint32_t)r;
      nzoff[i] = (uint32_t)off[r];
    } else {
      empty[r - i] = (uint32_t)r;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) nzoff[nzidx[V]] = (uint32_t)off[V];
}

// tbase[t] = index of the last nonempty row that starts before entry t*PF_T
// (-1 for none).
__global__ void k_pf_tbase(const uint32_t *nzoff, const uint32_t *nzidx, uint32_t V,
                           uint32_t n_tiles, int32_t *tbase) {
  const uint32_t n = nzidx[V];
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = t * PF_T;
    uint32_t lo = 0, hi = n;  // first i with nzoff[i] >= e
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (nzoff[mid] < e) lo = mid + 1; else hi = mid;
    }
    tbase[t] = (int32_t)lo - 1;
  }
}

// Rows whose entries span at least PF_LONG tiles, listed in row order.
__global__ void k_pf_longflag(const uint32_t *nzoff, const uint32_t *nzidx, uint32_t V,
                              uint32_t *flag) {
  const uint32_t n = nzidx[V];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    flag[i] = (nzoff[i + 1] - 1) / PF_T - nzoff[i] / PF_T >= PF_LONG ? 1u : 0u;
}

__global__ void k_pf_longfill(const uint32_t *nzidx, uint32_t V, const uint32_t *lidx,
                              uint32_t *longrows) {
  const uint32_t n = nzidx[V];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    if (lidx[i + 1] > lidx[i]) longrows[lidx[i]] = (uint32_t)i;
}

// ---------------------------------------------------------- iteration --
struct PfArgs {
  const uint32_t *cols;
  const int32_t *tbase;
  float2 *P, *Q, *S;
  const float *x, *cb;
  float *xn, *cn;
  const uint32_t *deg_out, *nzlabel, *nzoff, *empty, *longrows;
  double *partials;
  PrCtl *ctl;
  uint64_t E;
  uint32_t V, nh, n_tiles, n_nz, n_empty, n_long, compensated, max_iter;
  float alpha, base;
  double eps;
};

// left := left (+) right, left first; compensated or plain
__device__ __forceinline__ void pf_join(float &s, float &c, float s2, float c2, uint32_t comp) {
  if (comp) {
    const float t = s + s2;
    const float bp = t - s;
    const float e = (s - (t - bp)) + (s2 - bp);
    s = t;
    c = c + c2 - e;
  } else {
    s = s + s2;
  }
}

__device__ __forceinline__ void pf_load(const uint32_t *cols, uint32_t t, uint32_t lane,
                                        uint32_t (&e)[PF_K]) {
  const uint32_t *p = cols + (uint64_t)t * PF_T + lane;
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) e[k] = __builtin_nontemporal_load(p + k * 64);
}

// Stages the contributions of labels below nh in LDS (every thread of the
// block takes part; the caller then waits at a block barrier).
__device__ __forceinline__ void pf_stage_hot(const PfArgs &a, float *s_hot) {
  constexpr int PER = PF_HOT / 4 / PF_THREADS;
  const uint32_t n4 = a.nh / 4;
  const float4 *cb4 = reinterpret_cast<const float4 *>(a.cb);
  float4 *s4 = reinterpret_cast<float4 *>(s_hot);
  float4 tv[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = threadIdx.x + k * PF_THREADS;
    tv[k] = i < n4 ? cb4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const uint32_t i = threadIdx.x + k * PF_THREADS;
    if (i < n4) s4[i] = tv[k];
  }
  for (uint32_t i = n4 * 4 + threadIdx.x; i < a.nh; i += PF_THREADS) s_hot[i] = a.cb[i];
}

// Cold gathers of one tile: raw buffer loads of the contributions of labels
// at or above nh; LDS-served and padding entries get an out-of-range offset
// (the load returns 0 without touching L1/L2).
__device__ __forceinline__ void pf_gather(__amdgpu_buffer_rsrc_t rcb, uint32_t nh,
                                          const uint32_t (&e)[PF_K], float (&gv)[PF_K]) {
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) {
    const uint32_t w = e[k] & ~PF_FLAG;
    const uint32_t off = w < nh || w == PF_PAD ? PF_OOB : w * 4u;
    gv[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rcb, (int)off, 0, 0));
  }
}

// Row sums of tile t from its entries (cur), their LDS-staged values (hv,
// used for labels below nh) and their gathered values (gv): rows begun and
// ended in the tile go to S, the tile's first-row and last-row partials to
// P[t] and Q[t].
__device__ __forceinline__ void pf_tile_sums(const PfArgs &a, uint32_t t, int32_t tb_cur,
                                             const uint32_t (&cur)[PF_K],
                                             const float (&hv)[PF_K], const float (&gv)[PF_K],
                                             uint32_t lane, uint32_t comp) {
  uint32_t fm = 0;
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) fm |= (cur[k] >> 31) << k;
  const uint32_t nst = (uint32_t)__popc(fm);
  const uint32_t incl = wave_incl_scan<uint32_t>(nst);
  // nonempty row in progress at the start of this lane's entries
  int32_t row = tb_cur + (int32_t)(incl - nst);
  float s = 0.f, c = 0.f, hs = 0.f, hc = 0.f;
  bool seen = false;
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) {
    if ((fm >> k) & 1u) {
      if (seen) {
        a.S[row] = make_float2(s, c);  // row begun and ended in this lane
      } else {
        hs = s;
        hc = c;
      }
      ++row;
      s = 0.f;
      c = 0.f;
      seen = true;
    }
    const uint32_t w = cur[k] & ~PF_FLAG;
    const float g = w < a.nh ? hv[k] : gv[k];
    if (w != PF_PAD) {
      if (comp) kahan_add(s, c, g);
      else s += g;
    }
  }
  if (!seen) {
    hs = s;
    hc = c;
  }
  // segmented inclusive scan over lanes of the value carried out of each
  // lane: its last row's partial (if a row starts in it) or its whole sum
  float vs = seen ? s : hs, vc = seen ? c : hc;
  uint32_t vf = seen ? 1u : 0u;
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const float os = __shfl_up(vs, d, WAVE), oc = __shfl_up(vc, d, WAVE);
    const uint32_t of = __shfl_up(vf, d, WAVE);
    if (lane >= (uint32_t)d && !vf) {
      float ns = os, nc = oc;
      pf_join(ns, nc, vs, vc, comp);
      vs = ns;
      vc = nc;
      vf = of;
    }
  }
  float cs = __shfl_up(vs, 1, WAVE), cc = __shfl_up(vc, 1, WAVE);
  uint32_t cf = __shfl_up(vf, 1, WAVE);
  if (lane == 0) {
    cs = 0.f;
    cc = 0.f;
    cf = 0u;
  }
  if (seen) {
    // the row in progress at this lane's start ends before its first flag
    pf_join(cs, cc, hs, hc, comp);
    const int32_t r0 = tb_cur + (int32_t)(incl - nst);
    if (cf) a.S[r0] = make_float2(cs, cc);        // begun in an earlier lane
    else a.P[t] = make_float2(cs, cc);            // the tile's first row
  }
  const float ls = __shfl(vs, WAVE - 1, WAVE), lc = __shfl(vc, WAVE - 1, WAVE);
  const uint32_t lf = __shfl(vf, WAVE - 1, WAVE);
  if (lane == 0) {
    a.Q[t] = make_float2(ls, lc);
    if (!lf) a.P[t] = make_float2(ls, lc);  // no row starts in this tile
  }
}

__global__ __launch_bounds__(PF_THREADS) void k_pf_tiles(PfArgs a) {
  __shared__ float s_hot[PF_HOT];
  if (a.ctl->done) return;
  pf_stage_hot(a, s_hot);
  __syncthreads();
  // gfx9 buffer descriptor: raw (stride 0), range V*4 bytes
  const __amdgpu_buffer_rsrc_t rcb =
      __builtin_amdgcn_make_buffer_rsrc((void *)a.cb, (short)0, (int)(a.V * 4u), 0x00020000);
  const uint32_t lane = lane_id();
  // tiles are dealt out strided (t = w, w + W, ...): the cost of a tile
  // depends on where it lies in the stream (rows of the highest-degree
  // vertices have mostly cold sources, rows of low-degree ones mostly
  // LDS-staged sources), so contiguous runs per wavefront would not balance
  const uint32_t W = gridDim.x * (PF_THREADS / WAVE);
  const uint32_t gw =
      __builtin_amdgcn_readfirstlane(blockIdx.x * (PF_THREADS / WAVE) + threadIdx.x / WAVE);
  if (gw >= a.n_tiles) return;
  const uint32_t comp = a.compensated;
  uint32_t cur[PF_K], nxt[PF_K];
  pf_load(a.cols, gw, lane, cur);
  int32_t tb_cur = a.tbase[gw];
  for (uint32_t t = gw; t < a.n_tiles; t += W) {
    float hv[PF_K], gv[PF_K];
    pf_gather(rcb, a.nh, cur, gv);
#pragma unroll
    for (uint32_t k = 0; k < PF_K; ++k) {
      const uint32_t w = cur[k] & ~PF_FLAG;
      hv[k] = s_hot[w < PF_HOT ? w : PF_HOT - 1];
    }
    // the next tile's entries are requested after this tile's gathers:
    // vector loads retire in issue order, so the gathers can be consumed
    // while the prefetch is still in flight
    __builtin_amdgcn_sched_barrier(0);
    // unconditional (the last tile reloads itself), so the wait counts for
    // the gathers below are exact on every path
    const uint32_t tn = t + W < a.n_tiles ? t + W : t;
    pf_load(a.cols, tn, lane, nxt);
    const int32_t tb_nxt = a.tbase[tn];
    __builtin_amdgcn_sched_barrier(0);
    pf_tile_sums(a, t, tb_cur, cur, hv, gv, lane, comp);
#pragma unroll
    for (uint32_t k = 0; k < PF_K; ++k) cur[k] = nxt[k];
    tb_cur = tb_nxt;
  }
}

__device__ double pf_long_rows(const PfArgs &a, uint32_t blk);

// Row sums assembled in tile order, the update, residual partials per block.
// Blocks from PF_ROW_BLOCKS on take the long rows (pf_long_rows), so both
// kinds run in one launch.
__global__ __launch_bounds__(PF_ROW_THREADS) void k_pf_rows(PfArgs a) {
  if (a.ctl->done) return;
  if (blockIdx.x >= PF_ROW_BLOCKS) {
    const double bs = pf_long_rows(a, blockIdx.x - PF_ROW_BLOCKS);
    if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
    return;
  }
  double diff = 0.0;
  const uint32_t comp = a.compensated;
  const uint64_t stride = (uint64_t)PF_ROW_BLOCKS * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n_nz; i += stride) {
    const uint32_t e0 = a.nzoff[i], e1 = a.nzoff[i + 1];
    const uint32_t ta = e0 / PF_T, tb = (e1 - 1) / PF_T;
    if (tb - ta >= PF_LONG) continue;  // pf_long_rows
    float2 p;
    if (ta == tb) {
      const bool last = e1 == (ta + 1) * PF_T || i + 1 == a.n_nz;
      p = last ? a.Q[ta] : a.S[i];
    } else {
      p = a.Q[ta];
      for (uint32_t t = ta + 1; t <= tb; ++t) {
        const float2 q = a.P[t];
        pf_join(p.x, p.y, q.x, q.y, comp);
      }
    }
    const uint32_t v = a.nzlabel[i];
    const float xv = a.base + a.alpha * (p.x - p.y);
    a.xn[v] = xv;
    const uint32_t d = a.deg_out[v];
    a.cn[v] = d ? xv / (float)d : 0.f;
    diff += fabs((double)xv - (double)a.x[v]);
  }
  // rows with no entries take x' = base in every iteration: once both
  // buffers hold it (after iterations 0 and 1) they change nothing and add
  // exactly 0 to the residual, so later iterations skip them
  const uint64_t n_empty = a.ctl->iters < 2 ? a.n_empty : 0;
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n_empty; j += stride) {
    const uint32_t v = a.empty[j];
    const float xv = a.base;
    a.xn[v] = xv;
    const uint32_t d = a.deg_out[v];
    a.cn[v] = d ? xv / (float)d : 0.f;
    diff += fabs((double)xv - (double)a.x[v]);
  }
  const double bs = block_sum<double, PF_ROW_THREADS>(diff);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
}

// Long rows, one wavefront each: lane l joins the tile partials P[ta+1+l],
// P[ta+1+l+64], ... in order, the lanes are joined by a fixed butterfly, and
// the first tile's partial Q[ta] goes in front.
__device__ double pf_long_rows(const PfArgs &a, uint32_t blk) {
  const uint32_t lane = lane_id();
  const uint32_t comp = a.compensated;
  const uint32_t nw = PF_LONG_BLOCKS * (PF_ROW_THREADS / WAVE);
  double diff = 0.0;
  for (uint32_t j = blk * (PF_ROW_THREADS / WAVE) + threadIdx.x / WAVE; j < a.n_long; j += nw) {
    const uint32_t i = a.longrows[j];
    const uint32_t ta = a.nzoff[i] / PF_T, tb = (a.nzoff[i + 1] - 1) / PF_T;
    float s = 0.f, c = 0.f;
    for (uint32_t t = ta + 1 + lane; t <= tb; t += WAVE) {
      const float2 q = a.P[t];
      pf_join(s, c, q.x, q.y, comp);
    }
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
      const float s2 = __shfl_xor(s, d, WAVE), c2 = __shfl_xor(c, d, WAVE);
      if (lane & d) {
        float ns = s2, nc = c2;
        pf_join(ns, nc, s, c, comp);
        s = ns;
        c = nc;
      } else {
        pf_join(s, c, s2, c2, comp);
      }
    }
    if (lane == 0) {
      float2 p = a.Q[ta];
      pf_join(p.x, p.y, s, c, comp);
      const uint32_t v = a.nzlabel[i];
      const float xv = a.base + a.alpha * (p.x - p.y);
      a.xn[v] = xv;
      const uint32_t d = a.deg_out[v];
      a.cn[v] = d ? xv / (float)d : 0.f;
      diff += fabs((double)xv - (double)a.x[v]);
    }
  }
  return block_sum<double, PF_ROW_THREADS>(diff);
}

// Fixed-order FP64 residual, then the stopping rule (pagerank.cu:56): done
// after the first update whose residual is below eps, or after max_iter.
__global__ __launch_bounds__(1024) void k_pf_resid(PfArgs a, uint32_t nparts) {
  if (a.ctl->done) return;
  double r = 0.0;
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) r += a.partials[i];
  r = block_sum<double, 1024>(r);
  if (threadIdx.x == 0) {
    a.ctl->residual = r;
    const unsigned it = a.ctl->iters + 1;
    a.ctl->iters = it;
    if (r < a.eps || it >= a.max_iter) a.ctl->done = 1;
  }
}

static int pf_device_cus() {
  static int cus[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cus[dev]) {
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        n <= 0)
      n = 256;
    cus[dev] = n;
  }
  return cus[dev];
}

static gv_status pf_reserve(PrWork &w, uint32_t V, uint64_t n_tiles) {
  if (w.pf_V < V || !w.pf_S) {
    hipFree(w.pf_S);
    hipFree(w.pf_nzlabel);
    hipFree(w.pf_nzoff);
    hipFree(w.pf_empty);
    hipFree(w.pf_nzidx);
    hipFree(w.pf_lidx);
    hipFree(w.pf_long);
    hipFree(w.pf_partials);
    w.pf_S = nullptr;
    w.pf_nzlabel = w.pf_nzoff = w.pf_empty = w.pf_nzidx = nullptr;
    w.pf_lidx = w.pf_long = nullptr;
    w.pf_partials = nullptr;
    w.pf_V = 0;
    const size_t n = (size_t)V + 1;
    GV_HIP(hipMalloc(&w.pf_S, n * sizeof(float2)));
    GV_HIP(hipMalloc(&w.pf_nzlabel, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_nzoff, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_empty, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_nzidx, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_lidx, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_long, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.pf_partials, (PF_ROW_BLOCKS + PF_LONG_BLOCKS) * sizeof(double)));
    w.pf_V = V;
  }
  if (w.pf_tiles_cap < n_tiles || !w.pf_P) {
    hipFree(w.pf_P);
    hipFree(w.pf_Q);
    hipFree(w.pf_tbase);
    w.pf_P = w.pf_Q = nullptr;
    w.pf_tbase = nullptr;
    w.pf_tiles_cap = 0;
    uint64_t cap = n_tiles + n_tiles / 8 + 16;
    if (cap < cdiv(w.snap_min_cap, PF_T) + 16) cap = cdiv(w.snap_min_cap, PF_T) + 16;
    GV_HIP(hipMalloc(&w.pf_P, cap * sizeof(float2)));
    GV_HIP(hipMalloc(&w.pf_Q, cap * sizeof(float2)));
    GV_HIP(hipMalloc(&w.pf_tbase, cap * sizeof(int32_t)));
    w.pf_tiles_cap = cap;
  }
  return GV_OK;
}

gv_status pagerank_snap_prepare(const Store &in, const uint32_t *deg_out, PrWork &w,
                                hipStream_t s, int layout) {
  if (layout == 5) {
    const gv_status st = pagerank_pb_prepare(in, deg_out, w, s);
    if (st != GV_E_UNSUPPORTED) return st;
  }
  return pagerank_flat_prepare(in, deg_out, w, s);
}

gv_status pagerank_snap_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                                double *residual, hipStream_t s, double *iter_ms) {
  if (w.pb_ready) return pagerank_pb_iterate(o, w, out_dev, iters, residual, s, iter_ms);
  return pagerank_flat_iterate(o, w, out_dev, iters, residual, s, iter_ms);
}

gv_status pagerank_flat_prepare(const Store &in, const uint32_t *deg_out, PrWork &w,
                                hipStream_t s) {
  const uint32_t V = in.V;
  w.pf_ready = 0;
  w.pb_ready = 0;
  // stream positions and buffer offsets are 32-bit; labels keep bit 31 free
  if (!V || V >= PF_PAD || (uint64_t)V * 4u > PF_OOB) return GV_E_UNSUPPORTED;
  gv_status st = pr_reserve(w, V);
  if (st) return st;
  st = pr_snapshot(in, deg_out, true, w, s, 1);
  if (st) return st;
  const uint64_t E = w.snap_E;
  const uint64_t n_tiles = cdiv(E, PF_T);
  if (n_tiles * PF_T >= 0xFFFFFFFFull) return GV_E_UNSUPPORTED;
  st = pf_reserve(w, V, n_tiles);
  if (st) return st;
  if (n_tiles * PF_T > E)
    k_pf_pad<<<1, 1024, 0, s>>>(w.snap_cols, E, n_tiles * PF_T);
  k_pf_nzflag<<<grid_for(V), 256, 0, s>>>(w.snap_off, V, w.pf_nzidx);
  scan_u32(w.pf_nzidx, w.pf_nzidx, V, nullptr, w.snap_tmp, s);
  k_pf_nzfill<<<grid_for(V), 256, 0, s>>>(w.snap_off, w.pf_nzidx, V, w.pf_nzlabel, w.pf_nzoff,
                                          w.pf_empty);
  if (n_tiles)
    k_pf_tbase<<<grid_for(n_tiles), 256, 0, s>>>(w.pf_nzoff, w.pf_nzidx, V, (uint32_t)n_tiles,
                                                 w.pf_tbase);
  uint32_t n_nz = 0, n_long = 0;
  GV_HIP(hipMemcpyAsync(&n_nz, w.pf_nzidx + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  if (n_nz) {
    k_pf_longflag<<<grid_for(n_nz), 256, 0, s>>>(w.pf_nzoff, w.pf_nzidx, V, w.pf_lidx);
    scan_u32(w.pf_lidx, w.pf_lidx, n_nz, nullptr, w.snap_tmp, s);
    k_pf_longfill<<<grid_for(n_nz), 256, 0, s>>>(w.pf_nzidx, V, w.pf_lidx, w.pf_long);
    GV_HIP(hipMemcpyAsync(&n_long, w.pf_lidx + n_nz, sizeof(uint32_t), hipMemcpyDeviceToHost,
                          s));
    GV_HIP(hipStreamSynchronize(s));
  }
  w.pf_E = E;
  w.pf_n_tiles = n_tiles;
  w.pf_n_nz = n_nz;
  w.pf_n_long = n_long;
  w.pf_prepV = V;
  w.pf_ready = 1;
  return GV_OK;
}

gv_status pagerank_flat_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                                double *residual, hipStream_t s, double *iter_ms) {
  if (iters) *iters = 0;
  if (residual) *residual = 0.0;
  if (!w.pf_ready || o.max_iter <= 0) {
    set_error("pagerank: no prepared snapshot, or max_iter <= 0");
    return GV_E_INVALID;
  }
  const uint32_t V = w.pf_prepV;
  const uint64_t n_tiles = w.pf_n_tiles;
  const uint32_t *dout = w.rl_dout;  // out-degrees in label order
  const float x0 = 1.0f / (float)V;
  pr_ctl_reset(w.ctl, s);
  pr_init_vectors(w.x[0], w.contrib[0], dout, V, x0, o.init, w.rl_order, s);

  PfArgs a;
  a.cols = w.snap_cols;
  a.tbase = w.pf_tbase;
  a.P = w.pf_P;
  a.Q = w.pf_Q;
  a.S = w.pf_S;
  a.deg_out = dout;
  a.nzlabel = w.pf_nzlabel;
  a.nzoff = w.pf_nzoff;
  a.empty = w.pf_empty;
  a.longrows = w.pf_long;
  a.partials = w.pf_partials;
  a.ctl = w.ctl;
  a.E = w.pf_E;
  a.V = V;
  a.nh = V < PF_HOT ? V : PF_HOT;
  a.n_tiles = (uint32_t)n_tiles;
  a.n_nz = w.pf_n_nz;
  a.n_empty = V - w.pf_n_nz;
  a.n_long = w.pf_n_long;
  a.compensated = o.compensated ? 1u : 0u;
  a.max_iter = (uint32_t)o.max_iter;
  a.alpha = o.alpha;
  a.base = (1.0f - o.alpha) / (float)V;
  a.eps = o.eps;
  const uint32_t tile_grid = w.pf_grid ? w.pf_grid : (uint32_t)pf_device_cus();
  if (iter_ms) {
    // ev[0] before the first iteration, ev[t + 1] after iteration t
    while (w.ev.size() < (size_t)o.max_iter + 1) {
      hipEvent_t e;
      GV_HIP(hipEventCreate(&e));
      w.ev.push_back(e);
    }
    GV_HIP(hipEventRecord(w.ev[0], s));
  }
  if (w.ev_grp[0] == nullptr) {
    for (int i = 0; i < 2; ++i)
      GV_HIP(hipEventCreateWithFlags(&w.ev_grp[i], hipEventDisableTiming));
  }
  int launched = 0;
  // Without a yield hook, iterations go out in groups of PF_GROUP; after each
  // group the done flag is copied to a host slot, and the host reads the
  // previous group's copy, so the device always has a group queued while the
  // host checks.  With a yield hook each iteration is checked before the
  // next is issued.
  constexpr int PF_GROUP = 8;
  const int group = o.yield ? 1 : PF_GROUP;
  int pending = -1;  // host slot of a group copy not yet read
  while (launched < o.max_iter) {
    const int t = launched;
    if (o.yield) o.yield(o.yield_arg);
    a.x = w.x[t & 1];
    a.cb = w.contrib[t & 1];
    a.xn = w.x[(t + 1) & 1];
    a.cn = w.contrib[(t + 1) & 1];
    if (n_tiles) k_pf_tiles<<<tile_grid, PF_THREADS, 0, s>>>(a);
    k_pf_rows<<<PF_ROW_BLOCKS + PF_LONG_BLOCKS, PF_ROW_THREADS, 0, s>>>(a);
    k_pf_resid<<<1, 1024, 0, s>>>(a, PF_ROW_BLOCKS + PF_LONG_BLOCKS);
    if (iter_ms) GV_HIP(hipEventRecord(w.ev[t + 1], s));
    ++launched;
    if (launched % group == 0 && launched < o.max_iter) {
      const int slot = 1 + (launched / group & 1);
      GV_HIP(hipMemcpyAsync(&w.h_ctl[slot], w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
      GV_HIP(hipEventRecord(w.ev_grp[slot - 1], s));
      if (o.yield) {
        GV_HIP(hipEventSynchronize(w.ev_grp[slot - 1]));
        if (w.h_ctl[slot].done) break;
      } else {
        if (pending > 0) {
          GV_HIP(hipEventSynchronize(w.ev_grp[pending - 1]));
          if (w.h_ctl[pending].done) break;
        }
        pending = slot;
      }
    }
  }
  GV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  const uint32_t it = w.h_ctl->iters;
  pr_unpermute(w.x[it & 1], w.rl_newid, V, out_dev, s);
  if (iters) *iters = (int)it;
  if (residual) *residual = w.h_ctl->residual;
  if (iter_ms) {
    // the counted iterations, launch gaps between them included
    float ms = 0.f;
    const int last = (int)it < launched ? (int)it : launched;
    if (last > 0) GV_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[last]));
    *iter_ms = ms;
  }
  return GV_OK;
}

gv_status pagerank_flat_run(const Store &in, const uint32_t *deg_out, const PrOpts &o,
                            float *out_dev, int *iters, double *residual, PrWork &w,
                            hipStream_t s, double *iter_ms) {
  if (iters) *iters = 0;
  if (residual) *residual = 0.0;
  if (!in.V) return GV_OK;
  if (o.max_iter <= 0) return GV_E_UNSUPPORTED;  // handled by the row form
  gv_status st = pagerank_flat_prepare(in, deg_out, w, s);
  if (st) return st;
  return pagerank_flat_iterate(o, w, out_dev, iters, residual, s, iter_ms);
}

}  // namespace gv
