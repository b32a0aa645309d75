// gv_pagerank_seg.hip -- PageRank on an XCD-segmented snapshot of the
// in-adjacency (same semantics and stopping rule as static_pagerank,
// pagerank.cu:3-68).
//
// Why: the pull gather x[u]/deg(u) is random over a V*4-byte array (16.8 MB at
// scale 22), four times one XCD's 4 MB L2, so most 4-byte gathers pull a whole
// line from the Infinity Fabric.  Here the vertices are relabelled by a
// bijection p(u) (multiplication by an odd constant modulo 2^n, cycle-walked
// into [0, V)), which spreads R-MAT's low-ID hubs evenly, and the sources are
// split into PS_SEGS = 8 contiguous ranges of p.  Workgroup b works on segment
// b mod 8; the hardware hands workgroups to the eight XCDs round-robin, so
// XCD k gathers only from segment k's 2 MB slice of the contribution array,
// which stays L2-resident.  Each (vertex, segment) sum goes to P[j*8 + k];
// a combine pass adds the 8 partials in segment order, applies the update
// and produces the next contributions and the FP64 residual partials.
//
// Snapshot build (once per call): emit (segment(p(u)), p(v), p(u)) keys for
// every live in-edge in slot order, stable radix sort on (segment, p(v)),
// split into per-(segment, vertex) offsets and source lists.  Every sum has a
// fixed order, so results are deterministic.
#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {

constexpr uint32_t PS_SEGS = 8;
constexpr uint32_t PS_MUL = 0x9E3779B1u;
constexpr int PS_THREADS = 256;

__device__ __forceinline__ uint32_t ps_perm(uint32_t u, uint32_t mask, uint32_t V) {
  uint32_t x = u;
  do {
    x = (uint32_t)((uint64_t)x * PS_MUL) & mask;
  } while (x >= V);
  return x;
}

struct PsGeom {
  uint32_t V, mask, segsz, bV;
};

__device__ __forceinline__ uint64_t ps_key(const PsGeom &g, uint32_t v, uint32_t u) {
  uint32_t pu = ps_perm(u, g.mask, g.V), pv = ps_perm(v, g.mask, g.V);
  uint64_t seg = pu / g.segsz;
  return (seg << (2 * g.bV)) | ((uint64_t)pv << g.bV) | pu;
}

// Emit one key per live in-edge of every vertex at off_in[v].  A wavefront
// takes 64 consecutive vertices: lanes handle their own small vertices
// (<= 32 slots) sequentially; larger ones are processed by the whole
// wavefront, one at a time, 64 slots per round with ballot compaction.
__global__ __launch_bounds__(256) void k_ps_emit(Store st, const uint64_t *off_in, PsGeom g,
                                                 uint64_t *keys) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t t = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
       t * WAVE < st.V; t += nwaves) {
    const uint32_t v = (uint32_t)(t * WAVE + lane);
    uint32_t us = 0, tbl = 0;
    uint64_t o = 0;
    if (v < st.V) {
      us = st.used[v];
      tbl = st.tbl_off[v];
      o = off_in[v];
    }
    const bool big = us > 32;
    if (!big) {
      uint64_t k = o;
      for (uint32_t j = 0; j < us; ++j) {
        uint32_t w = st.slots[slot_index(st.arena, tbl, j, st.log2B, st.B)];
        if (w != TOMB) keys[k++] = ps_key(g, v, w);
      }
    }
    unsigned long long bigm = __ballot(big);
    while (bigm) {
      const int src_lane = __ffsll((long long)bigm) - 1;
      bigm &= bigm - 1;
      const uint32_t bv = __shfl(v, src_lane, WAVE);
      const uint32_t bus = __shfl(us, src_lane, WAVE);
      const uint32_t btbl = __shfl(tbl, src_lane, WAVE);
      const uint64_t bo = __shfl(o, src_lane, WAVE);
      uint32_t wpos = 0;
      for (uint32_t base = 0; base < bus; base += WAVE) {
        uint32_t j = base + lane;
        uint32_t w = TOMB;
        if (j < bus) w = st.slots[slot_index(st.arena, btbl, j, st.log2B, st.B)];
        bool keep = w != TOMB;
        unsigned long long m = __ballot(keep);
        if (keep) keys[bo + wpos + mask_lt_count(m)] = ps_key(g, bv, w);
        wpos += (uint32_t)__popcll(m);
      }
    }
  }
}

// Sources and per-(segment, vertex) offsets from the sorted keys:
// r = segment * V + p(v); off[r] = first key index with index >= r.
__global__ void k_ps_split(const uint64_t *keys, uint64_t E, PsGeom g, uint32_t *src,
                           uint32_t *off) {
  const uint64_t maskV = (1ull << g.bV) - 1;
  const uint64_t R = (uint64_t)PS_SEGS * g.V;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < E;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t k = keys[i];
    src[i] = (uint32_t)(k & maskV);
    uint64_t r = (k >> (2 * g.bV)) * g.V + ((k >> g.bV) & maskV);
    uint64_t rp;
    if (i == 0) {
      rp = 0;
      off[0] = 0;
    } else {
      uint64_t kp = keys[i - 1];
      rp = (kp >> (2 * g.bV)) * g.V + ((kp >> g.bV) & maskV) + 1;
    }
    for (uint64_t q = rp; q <= r; ++q) off[q] = (uint32_t)i;
    if (i + 1 == E)
      for (uint64_t q = r + 1; q <= R; ++q) off[q] = (uint32_t)E;
  }
}

__global__ void k_ps_vertex_init(const uint32_t *deg_in, const uint32_t *deg_out, PsGeom g,
                                 uint32_t light_max, uint32_t *din_p, uint32_t *dout_p,
                                 uint32_t *heavy, float *xp, float *cp, float x0,
                                 const float *init) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t j = ps_perm((uint32_t)v, g.mask, g.V);
    uint32_t d = deg_out[v];
    uint32_t di = deg_in[v];
    if (di > light_max) atomicOr(&heavy[j >> 5], 1u << (j & 31));
    float xv = init ? init[v] : x0;
    din_p[j] = di;
    dout_p[j] = d;
    xp[j] = xv;
    cp[j] = d ? xv / (float)d : 0.f;
  }
}

__global__ void k_ps_unpermute(const float *xp, PsGeom g, float *out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = xp[ps_perm((uint32_t)v, g.mask, g.V)];
}

struct PsArgs {
  const uint32_t *src, *off, *list, *heavy;
  float *P;
  PsGeom g;
  uint32_t n2, n3, o2, o3;
  uint32_t UL, U2, U3, U;
};

__device__ __forceinline__ void ps_merge(float &s, float &c, float s2, float c2) {
  float t = s + s2;
  float bp = t - s;
  float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 - e;
}

template <int WIDTH>
__device__ __forceinline__ void ps_group_merge(float &s, float &c) {
#pragma unroll
  for (int d = WIDTH / 2; d >= 1; d >>= 1) {
    float s2 = __shfl_xor(s, d, WAVE), c2 = __shfl_xor(c, d, WAVE);
    ps_merge(s, c, s2, c2);
  }
}

// Kahan sum of cp[src[e]] for e = e0, e0 + STRIDE, ... < e1, with four
// source reads issued before their four gathers.
template <int STRIDE>
__device__ __forceinline__ void ps_gather(const uint32_t *src, const float *cp, uint32_t e0,
                                          uint32_t e1, float &s, float &c) {
  for (uint32_t e = e0; e < e1; e += 4 * STRIDE) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t q = e + k * STRIDE;
      w[k] = q < e1 ? __builtin_nontemporal_load(&src[q]) : 0xFFFFFFFFu;
    }
    float gv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) gv[k] = w[k] != 0xFFFFFFFFu ? cp[w[k]] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (w[k] != 0xFFFFFFFFu) kahan_add(s, c, gv[k]);
  }
}

constexpr uint32_t PS_RUN = 8;  // consecutive light vertices per thread

// Light vertices: thread t walks the PS_RUN consecutive vertices j0..j0+7 of
// segment k as one contiguous edge range (edges are sorted by vertex inside a
// segment), eight source reads and eight gathers in flight per step, and
// flushes each vertex's Kahan sum when the walk crosses its end offset.
// Heavy vertices in the window are skipped (their own units write them).
__device__ __forceinline__ void ps_light(const PsArgs &a, const uint32_t *offk, float *Pk,
                                         const float *cp, uint32_t j0) {
  const uint32_t V = a.g.V;
  const uint32_t jn = j0 + PS_RUN < V ? j0 + PS_RUN : V;
  const uint32_t hbits = (a.heavy[j0 >> 5] >> (j0 & 31)) & 0xFFu;  // j0 is a multiple of 8
  uint32_t jj = j0;
  while (jj < jn) {
    if ((hbits >> (jj - j0)) & 1u) {
      ++jj;
      continue;
    }
    uint32_t je = jj;
    while (je < jn && !((hbits >> (je - j0)) & 1u)) ++je;
    // light run jj..je-1: edges [offk[jj], offk[je])
    uint32_t v = jj;
    uint32_t vend = offk[v + 1];
    const uint32_t eend = offk[je];
    float s = 0.f, c = 0.f;
    for (uint32_t e = offk[jj]; e < eend; e += 8) {
      uint32_t w[8];
#pragma unroll
      for (int k = 0; k < 8; ++k)
        w[k] = e + k < eend ? __builtin_nontemporal_load(&a.src[e + k]) : 0u;
      float gv[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) gv[k] = e + k < eend ? cp[w[k]] : 0.f;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        if (e + k < eend) {
          while (e + k >= vend) {
            Pk[v] = s - c;
            s = c = 0.f;
            ++v;
            vend = offk[v + 1];
          }
          kahan_add(s, c, gv[k]);
        }
      }
    }
    for (; v < je; ++v) {  // the last vertex with edges and any empty ones after it
      Pk[v] = s - c;
      s = c = 0.f;
    }
    jj = je;
  }
}

// Segment pass: workgroup b handles segment b % 8 of unit b / 8; units are
// heaviest first: one hub per workgroup, four vertices of one wavefront each,
// then 256 threads x PS_RUN consecutive light vertices.
__global__ __launch_bounds__(PS_THREADS) void k_ps_seg(PsArgs a, const float *cp, PrCtl *ctl) {
  if (ctl->done) return;
  const uint32_t k = blockIdx.x % PS_SEGS;
  const uint32_t u = blockIdx.x / PS_SEGS;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid / WAVE;
  const uint32_t *offk = a.off + (uint64_t)k * a.g.V;
  float *Pk = a.P + (uint64_t)k * a.g.V;
  if (u < a.U3) {
    __shared__ float red_s[PS_THREADS / WAVE], red_c[PS_THREADS / WAVE];
    const uint32_t j = a.list[a.o3 + u];
    float s = 0.f, c = 0.f;
    ps_gather<PS_THREADS>(a.src, cp, offk[j] + tid, offk[j + 1], s, c);
    ps_group_merge<WAVE>(s, c);
    if (lane == 0) {
      red_s[wid] = s;
      red_c[wid] = c;
    }
    __syncthreads();
    if (tid == 0) {
      float ts = red_s[0], tc = red_c[0];
      for (int w = 1; w < PS_THREADS / WAVE; ++w) ps_merge(ts, tc, red_s[w], red_c[w]);
      Pk[j] = ts - tc;
    }
  } else if (u < a.U3 + a.U2) {
    uint32_t gi = (u - a.U3) * (PS_THREADS / WAVE) + wid;
    if (gi < a.n2) {
      const uint32_t j = a.list[a.o2 + gi];
      float s = 0.f, c = 0.f;
      ps_gather<WAVE>(a.src, cp, offk[j] + lane, offk[j + 1], s, c);
      ps_group_merge<WAVE>(s, c);
      if (lane == 0) Pk[j] = s - c;
    }
  } else {
    uint32_t j0 = ((u - a.U3 - a.U2) * PS_THREADS + tid) * PS_RUN;
    if (j0 < a.g.V) ps_light(a, offk, Pk, cp, j0);
  }
}

// Combine: x'[j] = (1-alpha)/V + alpha * sum_k P[k][j] (k in order, Kahan),
// next contributions and the |x' - x| partial of this workgroup.
__global__ __launch_bounds__(PS_THREADS) void k_ps_combine(const float *P, const float *xp,
                                                           const uint32_t *dout_p, uint32_t V,
                                                           float alpha, float base, float *xn,
                                                           float *cn, double *partials,
                                                           PrCtl *ctl) {
  if (ctl->done) return;
  const uint64_t j = (uint64_t)blockIdx.x * PS_THREADS + threadIdx.x;
  double diff = 0.0;
  if (j < V) {
    float s = 0.f, c = 0.f;
#pragma unroll
    for (uint32_t k = 0; k < PS_SEGS; ++k) kahan_add(s, c, P[(uint64_t)k * V + j]);
    float xv = base + alpha * (s - c);
    uint32_t d = dout_p[j];
    xn[j] = xv;
    cn[j] = d ? xv / (float)d : 0.f;
    diff = fabs((double)xv - (double)xp[j]);
  }
  double bs = block_sum<double, PS_THREADS>(diff);
  if (threadIdx.x == 0) partials[blockIdx.x] = bs;
}

__global__ __launch_bounds__(1024) void k_ps_resid(const double *partials, uint32_t np, double eps,
                                                   uint32_t max_iter, PrCtl *ctl) {
  if (ctl->done) return;
  double r = 0.0;
  for (uint32_t i = threadIdx.x; i < np; i += 1024) r += partials[i];
  r = block_sum<double, 1024>(r);
  if (threadIdx.x == 0) {
    ctl->residual = r;
    unsigned it = ctl->iters + 1;
    ctl->iters = it;
    if (r < eps || it >= max_iter) ctl->done = 1;
  }
}

// ------------------------------------------------------------------ host --
void ps_release(PsWork &w) {
  void *ps[] = {w.src, w.off, w.list, w.P, w.xp[0], w.xp[1], w.cp[0], w.cp[1],
                w.din, w.dout, w.heavy, w.partials};
  for (void *p : ps) hipFree(p);
  for (hipEvent_t e : w.ev) hipEventDestroy(e);
  w = PsWork();
}

static gv_status ps_reserve(PsWork &w, uint32_t V, uint64_t E) {
  if (w.V != V) {
    ps_release(w);
    w.V = V;
    size_t n = (size_t)(V ? V : 1);
    GV_HIP(hipMalloc(&w.off, ((size_t)PS_SEGS * n + 1) * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.list, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.P, (size_t)PS_SEGS * n * sizeof(float)));
    for (int i = 0; i < 2; ++i) {
      GV_HIP(hipMalloc(&w.xp[i], n * sizeof(float)));
      GV_HIP(hipMalloc(&w.cp[i], n * sizeof(float)));
    }
    GV_HIP(hipMalloc(&w.din, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.dout, n * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.heavy, (cdiv(n, 32) + 1) * sizeof(uint32_t)));
    w.part_cap = cdiv(n, PS_THREADS) + 16;
    GV_HIP(hipMalloc(&w.partials, w.part_cap * sizeof(double)));
  }
  if (E > w.E_cap) {
    hipFree(w.src);
    w.src = nullptr;
    w.E_cap = E + (E >> 3) + 1024;
    GV_HIP(hipMalloc(&w.src, w.E_cap * sizeof(uint32_t)));
  }
  return GV_OK;
}

gv_status pagerank_seg_run(const Store &in, const uint32_t *deg_out, const PrBins &bins,
                           const PrOpts &o, float *out_dev, int *iters, double *residual,
                           PrWork &pw, PsWork &w, hipStream_t s, double *iter_ms) {
  const float alpha = o.alpha;
  const double eps = o.eps;
  const int max_iter = o.max_iter;
  const uint32_t V = in.V;
  if (iters) *iters = 0;
  if (residual) *residual = 0.0;
  if (!V) return GV_OK;
  gv_status st = pr_reserve(pw, V);  // control block, host mirror, scratch lists
  if (st) return st;
  PsGeom g;
  g.V = V;
  g.bV = id_bits(V);
  if (2 * g.bV + 3 > 64) {
    set_error("segmented PageRank needs fewer than 2^30 vertices");
    return GV_E_UNSUPPORTED;
  }
  g.mask = g.bV >= 32 ? 0xFFFFFFFFu : ((1u << g.bV) - 1);
  g.segsz = (uint32_t)cdiv(V, PS_SEGS);
  // 1. emission offsets from live in-degrees
  uint64_t *off_in = nullptr;
  void *tmp = nullptr;
  GV_HIP(hipMalloc(&off_in, ((size_t)V + 1) * sizeof(uint64_t)));
  GV_HIP(hipMalloc(&tmp, scan_tmp_bytes((uint64_t)V + 1)));
  degrees_u64(in, off_in, s);
  scan_u64(off_in, off_in, V, nullptr, tmp, s);
  uint64_t E = 0;
  GV_HIP(hipMemcpyAsync(&E, off_in + V, 8, hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  hipFree(tmp);
  tmp = nullptr;
  if (E >= 0xFFFFFFFFull) {
    hipFree(off_in);
    set_error("segmented PageRank needs fewer than 2^32 edges");
    return GV_E_UNSUPPORTED;
  }
  st = ps_reserve(w, V, E);
  if (st) {
    hipFree(off_in);
    return st;
  }
  // 2. keys, stable sort on (segment, p(v)), split
  uint64_t *k0 = nullptr, *k1 = nullptr;
  gv_status rc = GV_OK;
  do {
    hipError_t e;
    if ((e = hipMalloc(&k0, (E ? E : 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&k1, (E ? E : 1) * 8)) != hipSuccess ||
        (e = hipMalloc(&tmp, sort_tmp_bytes(E))) != hipSuccess) {
      rc = hip_fail(e, "segmented PageRank alloc");
      break;
    }
    k_ps_emit<<<grid_for((uint64_t)V, 256, 8192), 256, 0, s>>>(in, off_in, g, k0);
    uint64_t *sorted = sort_keys_range(k0, k1, E, g.bV, 2 * g.bV + 3, tmp, s);
    if ((rc = sort_check(tmp, E, g.bV, 2 * g.bV + 3, s)) != GV_OK) break;
    if (E) {
      k_ps_split<<<grid_for(E), 256, 0, s>>>(sorted, E, g, w.src, w.off);
    } else {
      fill_u32(w.off, 0, (uint64_t)PS_SEGS * V + 1, s);
    }
  } while (0);
  hipFree(k0);
  hipFree(k1);
  hipFree(tmp);
  hipFree(off_in);
  if (rc) return rc;
  // 3. permuted vertex arrays, bins on the in-degree
  const float x0 = 1.0f / (float)V;
  pr_ctl_reset(pw.ctl, s);
  fill_u32(w.heavy, 0, cdiv(V, 32) + 1, s);
  k_ps_vertex_init<<<grid_for(V), 256, 0, s>>>(in.deg, deg_out, g, bins.thread_max, w.din,
                                               w.dout, w.heavy, w.xp[0], w.cp[0], x0, o.init);
  if (max_iter <= 0) {
    k_ps_unpermute<<<grid_for(V), 256, 0, s>>>(w.xp[0], g, out_dev);
    GV_HIP(hipStreamSynchronize(s));
    return GV_OK;
  }
  pr_bins_count(w.din, V, bins, pw.ctl, s);
  GV_HIP(hipMemcpyAsync(pw.h_ctl, pw.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  const PrCtl hc = *pw.h_ctl;
  pr_bins_place(w.din, V, bins, w.list, pw.ctl, s);
  PsArgs a;
  a.src = w.src;
  a.off = w.off;
  a.list = w.list;
  a.heavy = w.heavy;
  a.P = w.P;
  a.g = g;
  a.n2 = hc.n_bin[2];
  a.n3 = hc.n_bin[3];
  a.o2 = hc.n_bin[0] + hc.n_bin[1];
  a.o3 = a.o2 + a.n2;
  a.UL = (uint32_t)cdiv(V, (uint64_t)PS_THREADS * PS_RUN);
  a.U2 = (uint32_t)cdiv(a.n2, PS_THREADS / WAVE);
  a.U3 = a.n3;
  a.U = a.UL + a.U2 + a.U3;
  const uint32_t seg_grid = (a.U ? a.U : 1) * PS_SEGS;
  const uint32_t comb_grid = (uint32_t)cdiv(V, PS_THREADS);
  const float base = (1.0f - alpha) / (float)V;
  if (iter_ms) {
    while (w.ev.size() < 2 * (size_t)max_iter) {
      hipEvent_t e;
      GV_HIP(hipEventCreate(&e));
      w.ev.push_back(e);
    }
  }
  int launched = 0;
  const int check_every = 8;
  while (launched < max_iter) {
    const int t = launched;
    if (iter_ms) GV_HIP(hipEventRecord(w.ev[2 * t], s));
    k_ps_seg<<<seg_grid, PS_THREADS, 0, s>>>(a, w.cp[t & 1], pw.ctl);
    k_ps_combine<<<comb_grid, PS_THREADS, 0, s>>>(w.P, w.xp[t & 1], w.dout, V, alpha, base,
                                                  w.xp[(t + 1) & 1], w.cp[(t + 1) & 1],
                                                  w.partials, pw.ctl);
    k_ps_resid<<<1, 1024, 0, s>>>(w.partials, comb_grid, eps, (uint32_t)max_iter, pw.ctl);
    if (iter_ms) GV_HIP(hipEventRecord(w.ev[2 * t + 1], s));
    ++launched;
    if (launched % check_every == 0 && launched < max_iter) {
      GV_HIP(hipMemcpyAsync(pw.h_ctl, pw.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
      GV_HIP(hipStreamSynchronize(s));
      if (pw.h_ctl->done) break;
    }
  }
  GV_HIP(hipMemcpyAsync(pw.h_ctl, pw.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  const uint32_t it = pw.h_ctl->iters;
  k_ps_unpermute<<<grid_for(V), 256, 0, s>>>(w.xp[it & 1], g, out_dev);
  GV_HIP(hipStreamSynchronize(s));
  if (iters) *iters = (int)it;
  if (residual) *residual = pw.h_ctl->residual;
  if (iter_ms) {
    double tot = 0.0;
    for (uint32_t t = 0; t < it && (int)t < launched; ++t) {
      float ms = 0.f;
      GV_HIP(hipEventElapsedTime(&ms, w.ev[2 * t], w.ev[2 * t + 1]));
      tot += ms;
    }
    *iter_ms = tot;
  }
  return GV_OK;
}

}  // namespace gv
