// gv_analytics.hip -- PageRank (row forms and the shared snapshot builder)
// and betweenness centrality over the device-resident block store.
//
// PageRank (static_pagerank, pagerank.cu:3-68): pull form over the
// in-adjacency so every vertex's sum is computed by one fixed lane schedule
// (deterministic, no FP32 atomics).  The default form iterates a flat tiled
// stream (gv_pagerank_flat.hip) over the snapshot built here: a copy of the
// live in-adjacency, relabelled by descending out-degree, written in one pass
// over the slots.  The row forms in this file serve the other layouts:
// vertices binned by row length (one lane, 8 lanes, one wavefront, or chunks
// on many wavefronts whose partials a follow-up kernel joins in chunk order),
// over the relabelled snapshot with LDS-staged hot contributions (k_pr_hot),
// over CSR rows in vertex order, or straight over the block store.  Per-lane
// Kahan sums are merged with TwoSum; the L1 residual is reduced in FP64 in a
// fixed order by a one-block kernel that also sets the done flag, so
// iterations need no host round trip.
//
// Triangle counting (static_tc, tc.cu:3-27) lives in gv_tc.hip.
//
// BC (bc, bc.cu:3-88): level-synchronous Brandes.  Forward: frontier expansion
// with CAS on dist and FP64 atomic sigma (integer-valued, so order-free and
// exact to 2^53).  Backward: per level, a pull over out-neighbours one level
// deeper with FP32 Kahan sums.
#include "gv_analytics.h"
#include "gv_store.h"

#include <vector>

namespace gv {

// ---------------------------------------------------------------- helpers --
__device__ __forceinline__ void two_sum_merge(float &s, float &c, float s2, float c2) {
  float t = s + s2;
  float bp = t - s;
  float e = (s - (t - bp)) + (s2 - bp);
  s = t;
  c = c + c2 - e;
}

// =============================================================== PageRank ==
constexpr int PR_THREADS = 256;
constexpr int PR_GROUP = 8;

gv_status pr_reserve(PrWork &w, uint32_t V) {
  if (w.V >= V && w.ctl) return GV_OK;
  pr_release(w);
  w.V = V;
  size_t n = (size_t)(V ? V : 1);
  for (int i = 0; i < 2; ++i) {
    GV_HIP(hipMalloc(&w.x[i], n * sizeof(float)));
    GV_HIP(hipMalloc(&w.contrib[i], n * sizeof(float)));
  }
  GV_HIP(hipMalloc(&w.list, n * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&w.hub_base, n * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&w.ctl, sizeof(PrCtl)));
  GV_HIP(hipHostMalloc(&w.h_ctl, 3 * sizeof(PrCtl), hipHostMallocDefault));
  return GV_OK;
}

void pr_release(PrWork &w) {
  for (int i = 0; i < 2; ++i) {
    hipFree(w.x[i]);
    hipFree(w.contrib[i]);
  }
  hipFree(w.list);
  hipFree(w.hub_base);
  hipFree(w.chunk_hub);
  hipFree(w.chunk_part);
  hipFree(w.partials);
  hipFree(w.ctl);
  if (w.h_ctl) hipHostFree(w.h_ctl);
  for (hipEvent_t e : w.ev) hipEventDestroy(e);
  w.ev.clear();
  for (hipEvent_t &e : w.ev_grp) {
    if (e) hipEventDestroy(e);
    e = nullptr;
  }
  hipFree(w.snap_off);
  hipFree(w.snap_cols);
  hipFree(w.snap_istart);
  hipFree(w.snap_ivert);
  hipFree(w.snap_ipos);
  hipFree(w.snap_tmp);
  hipFree(w.rl_order);
  hipFree(w.rl_newid);
  hipFree(w.rl_len);
  hipFree(w.rl_dout);
  hipFree(w.rl_keys[0]);
  hipFree(w.rl_keys[1]);
  hipFree(w.rl_tmp);
  pb_release(w);
  hipFree(w.pf_tbase);
  hipFree(w.pf_P);
  hipFree(w.pf_Q);
  hipFree(w.pf_S);
  hipFree(w.pf_nzlabel);
  hipFree(w.pf_nzoff);
  hipFree(w.pf_empty);
  hipFree(w.pf_nzidx);
  hipFree(w.pf_lidx);
  hipFree(w.pf_long);
  hipFree(w.pf_partials);
  // sizing settings of the caller outlive the buffers
  const uint64_t min_cap = w.snap_min_cap;
  const uint32_t grid = w.pf_grid;
  const uint64_t gen = w.gen;
  w = PrWork();
  w.snap_min_cap = min_cap;
  w.pf_grid = grid;
  w.gen = gen + 1;
}

__device__ __forceinline__ int pr_bin(uint32_t used, PrBins b) {
  return used <= b.thread_max ? 0 : used <= b.group_max ? 1 : used <= b.wave_max ? 2 : 3;
}

__global__ void k_pr_reset(PrCtl *c) {
  PrCtl z = {};
  *c = z;
}

// x0 = 1/V, or a caller's vector (warm start of an incremental refresh);
// order (optional) maps the working vertex order to the caller's.
__global__ void k_pr_init(float *x, float *cb, const uint32_t *deg_out, uint32_t V, float x0,
                          const float *init, const uint32_t *order) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t d = deg_out[v];
    float xv = init ? init[order ? order[v] : v] : x0;
    x[v] = xv;
    cb[v] = d ? xv / (float)d : 0.f;
  }
}

__global__ void k_pr_unpermute(const float *x, const uint32_t *newid, uint32_t V, float *out) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = x[newid[v]];
}

// Bin counts: per-block totals, five atomics per block.
__global__ __launch_bounds__(256) void k_pr_count(const uint32_t *used, uint32_t V, PrBins b,
                                                  PrCtl *ctl) {
  uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, nch = 0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t u = used[v];
    int bin = pr_bin(u, b);
    c0 += bin == 0;
    c1 += bin == 1;
    c2 += bin == 2;
    if (bin == 3) {
      ++c3;
      nch += (u + b.wave_max - 1) / b.wave_max;
    }
  }
  uint32_t t0 = block_sum<uint32_t, 256>(c0), t1 = block_sum<uint32_t, 256>(c1);
  uint32_t t2 = block_sum<uint32_t, 256>(c2), t3 = block_sum<uint32_t, 256>(c3);
  uint32_t tn = block_sum<uint32_t, 256>(nch);
  if (threadIdx.x == 0) {
    if (t0) atomicAdd(&ctl->n_bin[0], t0);
    if (t1) atomicAdd(&ctl->n_bin[1], t1);
    if (t2) atomicAdd(&ctl->n_bin[2], t2);
    if (t3) atomicAdd(&ctl->n_bin[3], t3);
    if (tn) atomicAdd(&ctl->n_chunks, tn);
  }
}

// Placement: each block owns a contiguous vertex range, reserves its share of
// every bin with one atomic per bin, then writes its vertices in order.
__global__ __launch_bounds__(256) void k_pr_place(const uint32_t *used, uint32_t V, PrBins b,
                                                  uint32_t *list, PrCtl *ctl) {
  const uint64_t chunk = cdiv_dev(V, gridDim.x);
  const uint64_t lo = blockIdx.x * chunk;
  const uint64_t hi = lo + chunk < V ? lo + chunk : V;
  uint32_t c[4] = {0, 0, 0, 0};
  for (uint64_t v = lo + threadIdx.x; v < hi; v += blockDim.x) {
    int bin = pr_bin(used[v], b);
    c[0] += bin == 0;
    c[1] += bin == 1;
    c[2] += bin == 2;
    c[3] += bin == 3;
  }
  __shared__ uint32_t base[4];
  uint32_t tot[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) tot[k] = block_sum<uint32_t, 256>(c[k]);
  if (threadIdx.x == 0) {
    uint32_t off = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      base[k] = off + (tot[k] ? atomicAdd(&ctl->cursor[k], tot[k]) : 0u);
      off += ctl->n_bin[k];
    }
  }
  __syncthreads();
  uint32_t bb[4] = {base[0], base[1], base[2], base[3]};
  for (uint64_t t0 = lo; t0 < hi; t0 += blockDim.x) {
    uint64_t v = t0 + threadIdx.x;
    int bin = v < hi ? pr_bin(used[v], b) : -1;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      uint32_t tt;
      uint32_t ex = block_excl_scan<uint32_t, 256>(bin == k ? 1u : 0u, tt);
      if (bin == k) list[bb[k] + ex] = (uint32_t)v;
      bb[k] += tt;
    }
  }
}

void pr_ctl_reset(PrCtl *ctl, hipStream_t s) { k_pr_reset<<<1, 1, 0, s>>>(ctl); }
void pr_init_vectors(float *x, float *cb, const uint32_t *deg_out, uint32_t V, float x0,
                     const float *init, const uint32_t *order, hipStream_t s) {
  k_pr_init<<<grid_for(V), 256, 0, s>>>(x, cb, deg_out, V, x0, init, order);
}
void pr_unpermute(const float *x, const uint32_t *newid, uint32_t V, float *out, hipStream_t s) {
  k_pr_unpermute<<<grid_for(V), 256, 0, s>>>(x, newid, V, out);
}

// Chunk lists of the vertices in the chunked bin.  Chunk c of hub h lives at
// hub_base[h] + c, so partials combine in chunk order whatever the schedule.
__global__ void k_pr_hubs(const uint32_t *used, const uint32_t *list, uint32_t o3, uint32_t n3,
                          uint32_t wave_max, uint32_t *hub_base, uint32_t *chunk_hub,
                          unsigned int *cursor) {
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n3;
       h += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t v = list[o3 + h];
    uint32_t nch = (used[v] + wave_max - 1) / wave_max;
    uint32_t b = atomicAdd(cursor, nch);
    hub_base[h] = b;
    for (uint32_t c = 0; c < nch; ++c) chunk_hub[b + c] = (uint32_t)h;
  }
}

// ------------------------------------------------------------ snapshot --
// CSR copy of the live in-adjacency.  Each vertex's slots are cut into
// chunk items of SNAP_CH slots; an item's live count is its length when the
// vertex has no tombstones, else counted; one scan of the counts in
// (vertex, chunk) order gives every item's output position, and the offset
// of vertex v is the position of its first item.  Entries keep slot order.
constexpr uint32_t SNAP_CH = 1024;

__global__ void k_snap_items(const uint32_t *used, uint32_t V, uint32_t *ncks) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    ncks[v] = (used[v] + SNAP_CH - 1) / SNAP_CH;
}

__global__ void k_snap_expand(const uint32_t *istart, uint32_t V, uint32_t *ivert) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = istart[v]; i < istart[v + 1]; ++i) ivert[i] = (uint32_t)v;
}

// Live slots of each item's chunk.  Items are dealt to lanes: wave w of W,
// step r, lane l takes item (r * 64 + l) * W + w, so the items of one hub
// spread over many wavefronts.  A lane settles its item alone when no slot
// needs reading: the vertex has no tombstones (the chunk's slot count), or it
// has one item (its live degree).  The chunks of hubs with tombstones are
// then counted by the whole wavefront, one after another.
__global__ __launch_bounds__(256) void k_snap_count(Store st, const uint32_t *istart,
                                                    const uint32_t *ivert, uint64_t *ipos) {
  const uint32_t ni = istart[st.V];
  const uint32_t lane = lane_id();
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  for (uint64_t r = 0; r * WAVE * W + w < ni; ++r) {
    const uint64_t i = (r * WAVE + lane) * W + w;
    uint32_t lo = 0, hi = 0, tbl = 0;
    bool scan = false;
    if (i < ni) {
      const uint32_t v = ivert[i];
      const uint32_t us = st.used[v], ndeg = st.deg[v];
      lo = ((uint32_t)i - istart[v]) * SNAP_CH;
      hi = lo + SNAP_CH < us ? lo + SNAP_CH : us;
      if (us == ndeg) ipos[i] = hi - lo;
      else if (us <= SNAP_CH) ipos[i] = ndeg;  // the vertex's only item
      else {
        scan = true;
        tbl = st.tbl_off[v];
      }
    }
    for (unsigned long long m = __ballot(scan); m; m &= m - 1) {
      const int j = __ffsll(m) - 1;
      const uint32_t jlo = __builtin_amdgcn_readlane(lo, j), jhi = __builtin_amdgcn_readlane(hi, j);
      const uint32_t jtbl = __builtin_amdgcn_readlane(tbl, j);
      uint32_t cnt = 0;
      for (uint32_t p = jlo; p < jhi; p += WAVE) {
        const bool live = p + lane < jhi &&
                          st.slots[slot_index(st.arena, jtbl, p + lane, st.log2B, st.B)] != TOMB;
        cnt += (uint32_t)__popcll(__ballot(live));
      }
      if (lane == (uint32_t)j) ipos[i] = cnt;
    }
  }
}

__global__ void k_snap_off(const uint32_t *istart, uint32_t V, const uint64_t *ipos,
                           uint64_t *off) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= V;
       v += (uint64_t)gridDim.x * blockDim.x)
    off[v] = ipos[istart[v]];
}

// Live slots of each item packed, in slot order, at the item's output
// position.  With a relabelling (newid non-null) the row of old vertex v is
// row newid[v] of the snapshot (offsets off) and every entry w is written as
// newid[w]; the entry order within the row is kept.  flat != 0 writes the
// tiled stream layout of the flat PageRank iteration instead: entry p of the
// row-major stream is stored at pf_phys(p) and the first entry of every row
// carries PF_FLAG.  hitem (may be null): per item, the entries it wrote below
// hot_below.
// Items are dealt to lanes (wave w of W, step r, lane l takes item
// (r * 64 + l) * W + w); an item of at most SNAP_LANE slots is written by its
// lane alone, four slots (and their new labels) in flight at a time, the
// longer ones by the whole wavefront, one after another.
constexpr uint32_t SNAP_LANE = 64;

__device__ __forceinline__ void snap_put(uint32_t *cols, uint32_t flat, uint64_t q, uint64_t row0,
                                         uint32_t val) {
  if (flat) cols[pf_phys(q)] = val | (q == row0 ? PF_FLAG : 0u);
  else cols[q] = val;
}

__global__ __launch_bounds__(256) void k_snap_write(Store st, const uint32_t *istart,
                                                    const uint32_t *ivert, const uint64_t *ipos,
                                                    const uint32_t *newid, const uint64_t *off,
                                                    uint32_t *cols, uint32_t flat,
                                                    uint32_t *hitem, uint32_t hot_below,
                                                    const uint32_t *fail) {
  if (fail && *fail) return;
  const uint32_t ni = istart[st.V];
  const uint32_t lane = lane_id();
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  for (uint64_t r = 0; r * WAVE * W + w < ni; ++r) {
    const uint64_t i = (r * WAVE + lane) * W + w;
    uint32_t lo = 0, hi = 0, tbl = 0;
    uint64_t dst = 0, row0 = 0;
    if (i < ni) {
      const uint32_t v = ivert[i];
      const uint32_t us = st.used[v];
      tbl = st.tbl_off[v];
      lo = ((uint32_t)i - istart[v]) * SNAP_CH;
      hi = lo + SNAP_CH < us ? lo + SNAP_CH : us;
      dst = ipos[i];
      const uint64_t rs = ipos[istart[v]];
      row0 = newid ? off[newid[v]] : rs;
      if (newid) dst = row0 + (dst - rs);
    }
    if (hi - lo <= SNAP_LANE) {
      // rounds of 16 slots: lo is a multiple of SNAP_CH, so a round lies in
      // one block -- one block-table load as a round enters a new block,
      // four 16-byte slot loads, then the round's 16 new labels, all issued
      // together
      uint32_t hot = 0, blk = 0;
#pragma unroll 1
      for (uint32_t p = lo; p < hi; p += 16) {
        if (p == lo || (p & (st.B - 1)) == 0) blk = st.arena[tbl + (p >> st.log2B)];
        const uint4 *p4 =
            reinterpret_cast<const uint4 *>(st.slots + (blk * st.B + (p & (st.B - 1))));
        const uint32_t nr = hi - p;
        uint32_t x[16];
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
          const uint4 v = 4 * q < nr ? p4[q] : make_uint4(TOMB, TOMB, TOMB, TOMB);
          x[4 * q] = v.x;
          x[4 * q + 1] = v.y;
          x[4 * q + 2] = v.z;
          x[4 * q + 3] = v.w;
        }
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
          if (k >= nr) x[k] = TOMB;
          if (newid && x[k] != TOMB) x[k] = newid[x[k]];
        }
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
          if (x[k] == TOMB) continue;
          snap_put(cols, flat, dst, row0, x[k]);
          ++dst;
          hot += x[k] < hot_below ? 1u : 0u;
        }
      }
      if (hitem && i < ni) hitem[i] = hot;
    }
    for (unsigned long long m = __ballot(hi - lo > SNAP_LANE); m; m &= m - 1) {
      const int j = __ffsll(m) - 1;
      const uint32_t jlo = __builtin_amdgcn_readlane(lo, j), jhi = __builtin_amdgcn_readlane(hi, j);
      const uint32_t jtbl = __builtin_amdgcn_readlane(tbl, j);
      uint64_t jdst = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(dst >> 32), j) << 32) |
                      __builtin_amdgcn_readlane((uint32_t)dst, j);
      const uint64_t jrow0 =
          ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(row0 >> 32), j) << 32) |
          __builtin_amdgcn_readlane((uint32_t)row0, j);
      uint32_t hot = 0;
      for (uint32_t p = jlo; p < jhi; p += WAVE) {
        uint32_t x = TOMB;
        if (p + lane < jhi) x = st.slots[slot_index(st.arena, jtbl, p + lane, st.log2B, st.B)];
        const bool live = x != TOMB;
        const unsigned long long lm = __ballot(live);
        uint32_t val = 0;
        if (live) {
          val = newid ? newid[x] : x;
          snap_put(cols, flat, jdst + mask_lt_count(lm), jrow0, val);
        }
        if (hitem) hot += (uint32_t)__popcll(__ballot(live && val < hot_below));
        jdst += (uint64_t)__popcll(lm);
      }
      if (hitem && lane == (uint32_t)j) hitem[i] = hot;
    }
  }
}

// Relabelling keys: (V - out-degree) above the vertex ID, so an ascending sort
// lists vertices by descending out-degree, ties by ID.
__global__ void k_rl_keys(const uint32_t *deg_out, uint32_t V, uint32_t bV, uint64_t *keys) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = deg_out[v] < V ? deg_out[v] : V;
    keys[v] = ((uint64_t)(V - d) << bV) | v;
  }
}

__global__ void k_rl_perm(const uint64_t *sorted, uint32_t V, uint32_t bV, const uint32_t *rowlen,
                          const uint32_t *deg_out, uint32_t *order, uint32_t *newid,
                          uint32_t *len_new, uint32_t *dout_new, uint64_t *off_new) {
  const uint64_t mask = (1ull << bV) - 1;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(sorted[r] & mask);
    if (v >= V) continue;  // only after a failed sort, which the caller reports
    order[r] = v;
    newid[v] = (uint32_t)r;
    len_new[r] = rowlen[v];
    dout_new[r] = deg_out[v];
    off_new[r] = rowlen[v];
  }
}

// Capture form: the snapshot's entries must fit the columns buffer.
__global__ void k_snap_check(const uint64_t *off_end, uint64_t cap, uint32_t *fail) {
  if (*off_end > cap) atomicCAS(fail, 0u, 1u);
}

uint64_t pr_snapshot_icap(const Store &in) {
  // items: at most one per vertex plus one per SNAP_CH slots of the pool
  return (uint64_t)in.V + in.P * (uint64_t)in.B / SNAP_CH + 1;
}

gv_status pr_snapshot(const Store &in, const uint32_t *deg_out, bool relabel, PrWork &w,
                      hipStream_t s, int flat, uint32_t *hitem, uint32_t hot_below,
                      uint32_t *fail) {
  const uint32_t V = in.V;
  const uint64_t icap = pr_snapshot_icap(in);
  if (icap > 0xFFFFFFFFull) {
    set_error("pagerank snapshot: too many chunk items");
    return GV_E_INVALID;
  }
  if (fail && (!w.snap_istart || icap > w.snap_icap || !w.snap_cols ||
               w.snap_cap < w.snap_min_cap || (relabel && (w.rl_V < V || !w.rl_order)) || flat)) {
    set_error("pagerank capture: the snapshot buffers are not reserved for this graph");
    return GV_E_INVALID;
  }
  if (!w.snap_istart || icap > w.snap_icap) {
    ++w.gen;
    hipFree(w.snap_istart);
    hipFree(w.snap_ivert);
    hipFree(w.snap_ipos);
    hipFree(w.snap_off);
    hipFree(w.snap_tmp);
    w.snap_istart = w.snap_ivert = nullptr;
    w.snap_ipos = w.snap_off = nullptr;
    w.snap_tmp = nullptr;
    w.snap_icap = 0;
    GV_HIP(hipMalloc(&w.snap_istart, ((size_t)V + 1) * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.snap_ivert, icap * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.snap_ipos, (icap + 1) * sizeof(uint64_t)));
    GV_HIP(hipMalloc(&w.snap_off, ((size_t)V + 1) * sizeof(uint64_t)));
    w.snap_tmp_bytes = std::max(scan_tmp_bytes((uint64_t)V + 1), scan_tmp_bytes(icap + 1));
    GV_HIP(hipMalloc(&w.snap_tmp, w.snap_tmp_bytes));
    w.snap_icap = icap;
  }
  k_snap_items<<<grid_for(V), 256, 0, s>>>(in.used, V, w.snap_istart);
  scan_u32(w.snap_istart, w.snap_istart, V, nullptr, w.snap_tmp, s);
  k_snap_expand<<<grid_for(V), 256, 0, s>>>(w.snap_istart, V, w.snap_ivert);
  k_snap_count<<<grid_for(icap * WAVE, 256, 2048), 256, 0, s>>>(in, w.snap_istart, w.snap_ivert,
                                                                 w.snap_ipos);
  scan_u64(w.snap_ipos, w.snap_ipos, icap, w.snap_istart + V, w.snap_tmp, s);
  if (relabel) {
    if (w.rl_V < V || !w.rl_order) {
      ++w.gen;
      hipFree(w.rl_order);
      hipFree(w.rl_newid);
      hipFree(w.rl_len);
      hipFree(w.rl_dout);
      hipFree(w.rl_keys[0]);
      hipFree(w.rl_keys[1]);
      hipFree(w.rl_tmp);
      w.rl_order = w.rl_newid = w.rl_len = w.rl_dout = nullptr;
      w.rl_keys[0] = w.rl_keys[1] = nullptr;
      w.rl_tmp = nullptr;
      w.rl_V = 0;
      const size_t n = (size_t)V + 1;
      GV_HIP(hipMalloc(&w.rl_order, n * sizeof(uint32_t)));
      GV_HIP(hipMalloc(&w.rl_newid, n * sizeof(uint32_t)));
      GV_HIP(hipMalloc(&w.rl_len, n * sizeof(uint32_t)));
      GV_HIP(hipMalloc(&w.rl_dout, n * sizeof(uint32_t)));
      GV_HIP(hipMalloc(&w.rl_keys[0], n * sizeof(uint64_t)));
      GV_HIP(hipMalloc(&w.rl_keys[1], n * sizeof(uint64_t)));
      GV_HIP(hipMalloc(&w.rl_tmp, sort_tmp_bytes(n)));
      w.rl_V = V;
    }
    const uint32_t bV = id_bits(V);
    k_rl_keys<<<grid_for(V), 256, 0, s>>>(deg_out, V, bV, w.rl_keys[0]);
    // capture form: a failed sort is folded into *fail on the device
    const uint64_t *sorted = sort_keys_range(w.rl_keys[0], w.rl_keys[1], V, 0,
                                             bV + id_bits(V + 1), w.rl_tmp, s, fail);
    k_rl_perm<<<grid_for(V), 256, 0, s>>>(sorted, V, bV, in.deg, deg_out, w.rl_order,
                                          w.rl_newid, w.rl_len, w.rl_dout, w.snap_off);
    scan_u64(w.snap_off, w.snap_off, V, nullptr, w.snap_tmp, s);
  } else {
    k_snap_off<<<grid_for((uint64_t)V + 1), 256, 0, s>>>(w.snap_istart, V, w.snap_ipos,
                                                         w.snap_off);
  }
  if (fail) {
    k_snap_check<<<1, 1, 0, s>>>(w.snap_off + V, w.snap_cap, fail);
    k_snap_write<<<grid_for(icap * WAVE, 256, 2048), 256, 0, s>>>(
        in, w.snap_istart, w.snap_ivert, w.snap_ipos, relabel ? w.rl_newid : nullptr, w.snap_off,
        w.snap_cols, 0u, hitem, hot_below, fail);
    return GV_OK;
  }
  uint64_t E = 0;
  GV_HIP(hipMemcpyAsync(&E, w.snap_off + V, sizeof(E), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  if (relabel) {
    gv_status src = sort_check(w.rl_tmp, V, 0, id_bits(V) + id_bits(V + 1), s);
    if (src) return src;
  }
  w.snap_E = E;
  // the flat layout pads the stream to whole tiles
  const uint64_t need = flat ? cdiv(E, PF_T) * PF_T : E;
  if (!w.snap_cols || need > w.snap_cap) {
    ++w.gen;
    hipFree(w.snap_cols);
    w.snap_cols = nullptr;
    w.snap_cap = 0;
    uint64_t cap = need + need / 8 + PF_T;
    if (cap < w.snap_min_cap + PF_T) cap = w.snap_min_cap + PF_T;
    GV_HIP(hipMalloc(&w.snap_cols, cap * sizeof(uint32_t)));
    w.snap_cap = cap;
  }
  k_snap_write<<<grid_for(icap * WAVE, 256, 2048), 256, 0, s>>>(
      in, w.snap_istart, w.snap_ivert, w.snap_ipos, relabel ? w.rl_newid : nullptr, w.snap_off,
      w.snap_cols, flat ? 1u : 0u, hitem, hot_below, nullptr);
  return GV_OK;
}

// ------------------------------------------------------ iteration kernel --
// Adjacency sources of the pull: the block store itself (slot counts include
// tombstones, skipped), or the CSR snapshot (live entries only, streamed
// with non-temporal loads so they do not displace contributions in L2).
struct StoreAdj {
  Store st;
  __device__ __forceinline__ uint32_t len(uint32_t v) const { return st.used[v]; }
  __device__ __forceinline__ uint64_t base(uint32_t v) const { return st.tbl_off[v]; }
  __device__ __forceinline__ uint32_t at(uint64_t b, uint32_t pos) const {
    return st.slots[slot_index(st.arena, (uint32_t)b, pos, st.log2B, st.B)];
  }
};
struct CsrAdj {
  const uint64_t *off;
  const uint32_t *cols;
  __device__ __forceinline__ uint32_t len(uint32_t v) const {
    return (uint32_t)(off[v + 1] - off[v]);
  }
  __device__ __forceinline__ uint64_t base(uint32_t v) const { return off[v]; }
  __device__ __forceinline__ uint32_t at(uint64_t b, uint32_t pos) const {
    return __builtin_nontemporal_load(cols + b + pos);
  }
};

// Work units of one iteration: 256 single-lane vertices, 32 eight-lane
// vertices, 4 one-wavefront vertices, or 4 hub chunks.
struct PrArgs {
  const uint32_t *deg_out;
  const uint32_t *list;
  const uint32_t *hub_base;
  const uint32_t *chunk_hub;
  float2 *chunk_part;
  double *partials;
  PrCtl *ctl;
  float alpha, base;
  double eps;
  uint32_t n0, n1, n2, n3, o1, o2, o3;
  uint32_t U0, U1, U2, U3, U;
  uint32_t n_chunks, wave_max;
  uint32_t max_iter;
  uint32_t compensated;  // 1: Kahan sums per lane; 0: plain FP32 sums
  uint32_t V;
  // staged form (k_pr_hot): contributions of vertices [0, nh) come from LDS;
  // work items are the hub chunks, then bin-2 rows from ib2, bin-1 groups of
  // 8 rows from ib1, bin-0 groups of 64 rows from ib0, n_items in all
  uint32_t nh, ib2, ib1, ib0, n_items;
};

// Kahan-accumulate contributions of positions start, start+STRIDE, ... < end
// of the adjacency row at b.  ILP entry reads are issued before their ILP
// gathers, so each lane keeps many independent loads in flight.
template <int STRIDE, int ILP, class Adj>
__device__ __forceinline__ void pr_gather(const PrArgs &a, const Adj &adj, const float *cb,
                                          uint64_t b, uint32_t start, uint32_t end, float &s,
                                          float &c) {
  for (uint32_t j = start; j < end; j += ILP * STRIDE) {
    uint32_t w[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint32_t pos = j + k * STRIDE;
      w[k] = pos < end ? adj.at(b, pos) : TOMB;
    }
    float g[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) g[k] = w[k] != TOMB ? cb[w[k]] : 0.f;
    if (a.compensated) {
#pragma unroll
      for (int k = 0; k < ILP; ++k)
        if (w[k] != TOMB) kahan_add(s, c, g[k]);
    } else {
#pragma unroll
      for (int k = 0; k < ILP; ++k) s += g[k];
    }
  }
}

template <int WIDTH>
__device__ __forceinline__ void group_merge(float &s, float &c) {
#pragma unroll
  for (int d = WIDTH / 2; d >= 1; d >>= 1) {
    float s2 = __shfl_xor(s, d, WAVE), c2 = __shfl_xor(c, d, WAVE);
    two_sum_merge(s, c, s2, c2);
  }
}

__device__ __forceinline__ double pr_finish(const PrArgs &a, uint32_t v, float s, float c,
                                            const float *x, float *xn, float *cn) {
  float xv = a.base + a.alpha * (s - c);
  xn[v] = xv;
  uint32_t d = a.deg_out[v];
  cn[v] = d ? xv / (float)d : 0.f;
  return fabs((double)xv - (double)x[v]);
}

// One work unit per workgroup: hub chunks (BIN 3), one-wavefront vertices
// (2), eight-lane vertices (1), single-lane vertices (0).  Each bin is its own
// launch, heaviest first, so each gets the register budget -- and so the
// occupancy -- of its own load depth (WILP for the wavefront-wide rows).
// Unit u of the iteration is block u - (units of the bins launched before);
// each workgroup writes its residual partial to slot u, so no counter is
// shared between workgroups.
template <class Adj, int WILP, int BIN>
__global__ __launch_bounds__(PR_THREADS) void k_pr_iter(PrArgs a, Adj adj, const float *x,
                                                        const float *cb, float *xn, float *cn) {
  if (a.ctl->done) return;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wid = tid / WAVE;
  const uint32_t u = blockIdx.x + (BIN == 3   ? 0u
                                   : BIN == 2 ? a.U3
                                   : BIN == 1 ? a.U3 + a.U2
                                              : a.U3 + a.U2 + a.U1);
  double diff = 0.0;
  if (BIN == 3) {
    uint32_t ci = u * (PR_THREADS / WAVE) + wid;
    if (ci < a.n_chunks) {
      uint32_t h = a.chunk_hub[ci];
      uint32_t v = a.list[a.o3 + h];
      uint32_t us = adj.len(v);
      uint64_t b = adj.base(v);
      uint32_t hb = a.hub_base[h];
      uint32_t c0 = (ci - hb) * a.wave_max;
      uint32_t c1 = c0 + a.wave_max < us ? c0 + a.wave_max : us;
      float s = 0.f, c = 0.f;
      pr_gather<WAVE, WILP>(a, adj, cb, b, c0 + lane, c1, s, c);
      group_merge<WAVE>(s, c);
      // chunks of a hub are merged by k_pr_hubfin after this launch: the
      // kernel boundary makes the partials visible without a device fence
      if (lane == 0) a.chunk_part[ci] = make_float2(s, c);
    }
  } else if (BIN == 2) {
    uint32_t g = (u - a.U3) * (PR_THREADS / WAVE) + wid;
    if (g < a.n2) {
      uint32_t v = a.list[a.o2 + g];
      uint32_t us = adj.len(v);
      uint64_t b = adj.base(v);
      float s = 0.f, c = 0.f;
      pr_gather<WAVE, WILP>(a, adj, cb, b, lane, us, s, c);
      group_merge<WAVE>(s, c);
      if (lane == 0) diff = pr_finish(a, v, s, c, x, xn, cn);
    }
  } else if (BIN == 1) {
    uint32_t g = (u - a.U3 - a.U2) * (PR_THREADS / PR_GROUP) + tid / PR_GROUP;
    uint32_t gl = tid & (PR_GROUP - 1);
    uint32_t v = 0, us = 0;
    uint64_t b = 0;
    if (g < a.n1) {
      v = a.list[a.o1 + g];
      us = adj.len(v);
      b = adj.base(v);
    }
    float s = 0.f, c = 0.f;
    pr_gather<PR_GROUP, WILP>(a, adj, cb, b, gl, us, s, c);
    group_merge<PR_GROUP>(s, c);
    if (g < a.n1 && gl == 0) diff = pr_finish(a, v, s, c, x, xn, cn);
  } else {
    uint32_t i = (u - a.U3 - a.U2 - a.U1) * PR_THREADS + tid;
    if (i < a.n0) {
      uint32_t v = a.list[i];
      uint32_t us = adj.len(v);
      uint64_t b = adj.base(v);
      float s = 0.f, c = 0.f;
      pr_gather<1, WILP>(a, adj, cb, b, 0, us, s, c);
      diff = pr_finish(a, v, s, c, x, xn, cn);
    }
  }
  double bs = block_sum<double, PR_THREADS>(diff);
  if (tid == 0) a.partials[u] = bs;
}

// Hub finish: the chunk partials of each chunked vertex merged in chunk
// order, then the vertex update and its residual partial.
template <class Adj>
__global__ void k_pr_hubfin(PrArgs a, Adj adj, const float *x, float *xn, float *cn) {
  if (a.ctl->done) return;
  for (uint64_t h = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; h < a.n3;
       h += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = a.list[a.o3 + h];
    const uint32_t nch = (adj.len(v) + a.wave_max - 1) / a.wave_max;
    const uint32_t hb = a.hub_base[h];
    float ts = 0.f, tc = 0.f;
    for (uint32_t k = 0; k < nch; ++k) {
      const float2 p = a.chunk_part[hb + k];
      two_sum_merge(ts, tc, p.x, p.y);
    }
    a.partials[a.U + h] = pr_finish(a, (uint32_t)v, ts, tc, x, xn, cn);
  }
}

// Fixed-order FP64 reduction of the workgroup and hub partials, then the
// stopping rule of static_pagerank (pagerank.cu:56): done after the first
// update whose residual is below eps, or after max_iter updates.
__global__ __launch_bounds__(1024) void k_pr_resid(PrArgs a) {
  if (a.ctl->done) return;
  const uint32_t np = a.U + a.n3;
  double r = 0.0;
  for (uint32_t i = threadIdx.x; i < np; i += 1024) r += a.partials[i];
  r = block_sum<double, 1024>(r);
  if (threadIdx.x == 0) {
    a.ctl->residual = r;
    unsigned it = a.ctl->iters + 1;
    a.ctl->iters = it;
    if (r < a.eps || it >= a.max_iter) a.ctl->done = 1;
  }
}

// ------------------------------------------- staged iteration (LDS hot set) --
// With the snapshot relabelled by descending out-degree, the most gathered
// contributions are those of the lowest labels: at R-MAT scale 22 the first
// 32768 vertices are the sources of about half of all edges.  The gather rate
// of random 4-byte loads, not HBM bandwidth, bounds the pull, so every
// workgroup copies cb[0, nh) into LDS once per iteration and takes those
// gathers from LDS; only the rest go through L1/L2.  One workgroup per CU
// (128 KiB of LDS each), persistent: wave w of W takes work items w, 2W-1-w,
// 2W+w, ... (a snake over items ordered heaviest first, so every wave gets a
// near-equal share of edges with no shared counter).
constexpr int PRH_THREADS = 1024;
constexpr uint32_t PRH_HOT = 32768;

// Offset past the contribution buffer's range: a raw buffer load there
// returns 0 without a memory access.
constexpr uint32_t PRH_OOB = 0x80000000u;

// Both loads of every entry are issued unconditionally (no exec-mask
// branches, so all ILP loads are in flight together): the LDS read at
// w mod PRH_HOT, and a buffer load of cb[w] whose offset is pushed out of
// range for hot and empty entries, so only cold entries reach L1/L2.
template <int STRIDE, int ILP>
__device__ __forceinline__ void prh_gather(const PrArgs &a, const CsrAdj &adj,
                                           __amdgpu_buffer_rsrc_t rcb, const float *s_hot,
                                           uint64_t b, uint32_t start, uint32_t end, float &s,
                                           float &c) {
  for (uint32_t j = start; j < end; j += ILP * STRIDE) {
    uint32_t w[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint32_t pos = j + k * STRIDE;
      w[k] = pos < end ? adj.at(b, pos) : TOMB;
    }
    float hv[ILP], gv[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) {
      const uint32_t off = w[k] < a.nh || w[k] == TOMB ? PRH_OOB : w[k] * 4u;
      gv[k] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rcb, (int)off, 0, 0));
      hv[k] = s_hot[w[k] & (PRH_HOT - 1)];
    }
    // keep every load of the batch ahead of the first use
    __builtin_amdgcn_sched_barrier(0);
    float g[ILP];
#pragma unroll
    for (int k = 0; k < ILP; ++k) g[k] = w[k] < a.nh ? hv[k] : gv[k];
    if (a.compensated) {
#pragma unroll
      for (int k = 0; k < ILP; ++k)
        if (w[k] != TOMB) kahan_add(s, c, g[k]);
    } else {
#pragma unroll
      for (int k = 0; k < ILP; ++k) s += g[k];
    }
  }
}

template <int ILP_W, int ILP_G, int ILP_L>
__global__ __launch_bounds__(PRH_THREADS) void k_pr_hot(PrArgs a, CsrAdj adj, const float *x,
                                                        const float *cb, float *xn, float *cn) {
  __shared__ float s_hot[PRH_HOT];
  if (a.ctl->done) return;
  {
    // all of a thread's 16-byte loads in flight at once, then the LDS stores
    constexpr int PER = PRH_HOT / 4 / PRH_THREADS;
    const uint32_t n4 = a.nh / 4;
    const float4 *cb4 = reinterpret_cast<const float4 *>(cb);
    float4 *s4 = reinterpret_cast<float4 *>(s_hot);
    float4 t[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * PRH_THREADS;
      t[k] = i < n4 ? cb4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const uint32_t i = threadIdx.x + k * PRH_THREADS;
      if (i < n4) s4[i] = t[k];
    }
    for (uint32_t i = n4 * 4 + threadIdx.x; i < a.nh; i += PRH_THREADS) s_hot[i] = cb[i];
  }
  __syncthreads();
  // gfx9 buffer descriptor: raw (stride 0), range V*4 bytes
  const __amdgpu_buffer_rsrc_t rcb =
      __builtin_amdgcn_make_buffer_rsrc((void *)cb, (short)0, (int)(a.V * 4u), 0x00020000);
  const uint32_t lane = lane_id();
  const uint32_t W = gridDim.x * (PRH_THREADS / WAVE);
  const uint32_t gw = blockIdx.x * (PRH_THREADS / WAVE) + threadIdx.x / WAVE;
  double diff = 0.0;
  for (uint32_t r = 0;; ++r) {
    const uint32_t it = r * W + ((r & 1) ? W - 1 - gw : gw);
    if (it >= a.n_items) break;  // only the last round is partial
    float s = 0.f, c = 0.f;
    if (it < a.ib2) {
      // hub chunk: its partial is merged by k_pr_hotfin in chunk order
      const uint32_t h = a.chunk_hub[it];
      const uint32_t v = a.list[a.o3 + h];
      const uint32_t us = adj.len(v);
      const uint32_t c0 = (it - a.hub_base[h]) * a.wave_max;
      const uint32_t c1 = c0 + a.wave_max < us ? c0 + a.wave_max : us;
      prh_gather<WAVE, ILP_W>(a, adj, rcb, s_hot, adj.base(v), c0 + lane, c1, s, c);
      group_merge<WAVE>(s, c);
      if (lane == 0) a.chunk_part[it] = make_float2(s, c);
    } else if (it < a.ib1) {
      const uint32_t v = a.list[a.o2 + (it - a.ib2)];
      prh_gather<WAVE, ILP_W>(a, adj, rcb, s_hot, adj.base(v), lane, adj.len(v), s, c);
      group_merge<WAVE>(s, c);
      if (lane == 0) diff += pr_finish(a, v, s, c, x, xn, cn);
    } else if (it < a.ib0) {
      const uint32_t g = (it - a.ib1) * (WAVE / PR_GROUP) + lane / PR_GROUP;
      const uint32_t gl = lane & (PR_GROUP - 1);
      uint32_t v = 0, us = 0;
      uint64_t b = 0;
      if (g < a.n1) {
        v = a.list[a.o1 + g];
        us = adj.len(v);
        b = adj.base(v);
      }
      prh_gather<PR_GROUP, ILP_G>(a, adj, rcb, s_hot, b, gl, us, s, c);
      group_merge<PR_GROUP>(s, c);
      if (g < a.n1 && gl == 0) diff += pr_finish(a, v, s, c, x, xn, cn);
    } else {
      const uint32_t i = (it - a.ib0) * WAVE + lane;
      if (i < a.n0) {
        const uint32_t v = a.list[i];
        prh_gather<1, ILP_L>(a, adj, rcb, s_hot, adj.base(v), 0, adj.len(v), s, c);
        diff += pr_finish(a, v, s, c, x, xn, cn);
      }
    }
  }
  const double bs = block_sum<double, PRH_THREADS>(diff);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
}

// Hub chunks merged in chunk order, the hub updates, then the fixed-order
// FP64 residual over the workgroup and hub partials and the stopping rule of
// static_pagerank (pagerank.cu:56).
__global__ __launch_bounds__(1024) void k_pr_hotfin(PrArgs a, CsrAdj adj, const float *x,
                                                    float *xn, float *cn, uint32_t nparts) {
  if (a.ctl->done) return;
  double r = 0.0;
  for (uint32_t h = threadIdx.x; h < a.n3; h += 1024) {
    const uint32_t v = a.list[a.o3 + h];
    const uint32_t nch = (adj.len(v) + a.wave_max - 1) / a.wave_max;
    const uint32_t hb = a.hub_base[h];
    float ts = 0.f, tc = 0.f;
    for (uint32_t k = 0; k < nch; ++k) {
      const float2 p = a.chunk_part[hb + k];
      two_sum_merge(ts, tc, p.x, p.y);
    }
    r += pr_finish(a, v, ts, tc, x, xn, cn);
  }
  for (uint32_t i = threadIdx.x; i < nparts; i += 1024) r += a.partials[i];
  r = block_sum<double, 1024>(r);
  if (threadIdx.x == 0) {
    a.ctl->residual = r;
    const unsigned it = a.ctl->iters + 1;
    a.ctl->iters = it;
    if (r < a.eps || it >= a.max_iter) a.ctl->done = 1;
  }
}

static int device_cus() {
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

gv_status pagerank_run(const Store &in, const uint32_t *deg_out, const PrBins &bins,
                       const PrOpts &o, float *out_dev, int *iters, double *residual, PrWork &w,
                       hipStream_t s, double *iter_ms, int snapshot) {
  const float alpha = o.alpha;
  const double eps = o.eps;
  const int max_iter = o.max_iter;
  const uint32_t V = in.V;
  if (iters) *iters = 0;
  if (residual) *residual = 0.0;
  if (!V) return GV_OK;
  gv_status st = pr_reserve(w, V);
  if (st) return st;
  if (snapshot == 5 && max_iter > 0) {
    st = pagerank_pb_prepare(in, deg_out, w, s);
    if (st == GV_OK) return pagerank_pb_iterate(o, w, out_dev, iters, residual, s, iter_ms);
    if (st != GV_E_UNSUPPORTED) return st;
    snapshot = 1;
  }
  if (snapshot == 5) snapshot = 1;
  if (snapshot == 1) {
    st = pagerank_flat_run(in, deg_out, o, out_dev, iters, residual, w, s, iter_ms);
    if (st != GV_E_UNSUPPORTED) return st;
    st = GV_OK;
  }
  // the row form (also where the stream layouts do not apply: max_iter <= 0,
  // or graphs past their 32-bit positions)
  const float x0 = 1.0f / (float)V;
  k_pr_reset<<<1, 1, 0, s>>>(w.ctl);
  if (max_iter <= 0) {
    k_pr_init<<<grid_for(V), 256, 0, s>>>(w.x[0], w.contrib[0], deg_out, V, x0, o.init, nullptr);
    GV_HIP(hipMemcpyAsync(out_dev, w.x[0], (size_t)V * sizeof(float), hipMemcpyDeviceToDevice, s));
    GV_HIP(hipStreamSynchronize(s));
    return GV_OK;
  }
  st = pr_snapshot(in, deg_out, true, w, s, 0);
  if (st) return st;
  // every per-vertex array below is in the relabelled vertex order; the
  // result is permuted back at the end
  deg_out = w.rl_dout;
  k_pr_init<<<grid_for(V), 256, 0, s>>>(w.x[0], w.contrib[0], deg_out, V, x0, o.init,
                                        w.rl_order);
  // row lengths the bins are cut by: live entries of the snapshot
  const uint32_t *rowlen = w.rl_len;
  k_pr_count<<<grid_for(V, 256, 512), 256, 0, s>>>(rowlen, V, bins, w.ctl);
  GV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  const PrCtl hc = *w.h_ctl;
  const uint32_t n0 = hc.n_bin[0], n1 = hc.n_bin[1], n2 = hc.n_bin[2], n3 = hc.n_bin[3];
  const uint32_t nchunks = hc.n_chunks;
  if (nchunks > w.chunk_cap) {
    hipFree(w.chunk_hub);
    hipFree(w.chunk_part);
    w.chunk_cap = nchunks + 1024;
    GV_HIP(hipMalloc(&w.chunk_hub, w.chunk_cap * sizeof(uint32_t)));
    GV_HIP(hipMalloc(&w.chunk_part, w.chunk_cap * sizeof(float2)));
  }
  k_pr_place<<<grid_for(V, 256, 512), 256, 0, s>>>(rowlen, V, bins, w.list, w.ctl);
  PrArgs a;
  a.deg_out = deg_out;
  a.list = w.list;
  a.hub_base = w.hub_base;
  a.chunk_hub = w.chunk_hub;
  a.chunk_part = (float2 *)w.chunk_part;
  a.ctl = w.ctl;
  a.alpha = alpha;
  a.base = (1.0f - alpha) / (float)V;
  a.eps = eps;
  a.n0 = n0;
  a.n1 = n1;
  a.n2 = n2;
  a.n3 = n3;
  a.o1 = n0;
  a.o2 = n0 + n1;
  a.o3 = n0 + n1 + n2;
  a.U0 = (uint32_t)cdiv(n0, PR_THREADS);
  a.U1 = (uint32_t)cdiv(n1, PR_THREADS / PR_GROUP);
  a.U2 = (uint32_t)cdiv(n2, PR_THREADS / WAVE);
  a.U3 = (uint32_t)cdiv(nchunks, PR_THREADS / WAVE);
  a.U = a.U0 + a.U1 + a.U2 + a.U3;
  a.n_chunks = nchunks;
  a.wave_max = bins.wave_max;
  a.max_iter = (uint32_t)max_iter;
  a.compensated = o.compensated ? 1u : 0u;
  a.V = V;
  a.nh = V < PRH_HOT ? V : PRH_HOT;
  a.ib2 = nchunks;
  a.ib1 = a.ib2 + n2;
  a.ib0 = a.ib1 + (uint32_t)cdiv(n1, WAVE / PR_GROUP);
  a.n_items = a.ib0 + (uint32_t)cdiv(n0, WAVE);
  // staged form: one persistent workgroup per CU
  const bool staged = V <= PRH_OOB / 4;
  const uint32_t hot_grid = (uint32_t)device_cus();
  const uint32_t grid = staged ? hot_grid : a.U ? a.U : 1;
  if (n3)
    k_pr_hubs<<<grid_for(n3), 256, 0, s>>>(rowlen, w.list, a.o3, n3, bins.wave_max, w.hub_base,
                                           w.chunk_hub, &w.ctl->chunk_cursor);
  const uint64_t nparts = (uint64_t)grid + n3;
  if (nparts > w.part_cap) {
    hipFree(w.partials);
    w.part_cap = nparts + 1024;
    GV_HIP(hipMalloc(&w.partials, w.part_cap * sizeof(double)));
  }
  a.partials = w.partials;
  if (iter_ms) {
    while (w.ev.size() < 2 * (size_t)max_iter) {
      hipEvent_t e;
      GV_HIP(hipEventCreate(&e));
      w.ev.push_back(e);
    }
  }
  const CsrAdj cadj{w.snap_off, w.snap_cols};
  int launched = 0;
  const int check_every = 8;
  while (launched < max_iter) {
    int t = launched;
    if (iter_ms) GV_HIP(hipEventRecord(w.ev[2 * t], s));
    const float *x = w.x[t & 1], *cb = w.contrib[t & 1];
    float *xn = w.x[(t + 1) & 1], *cn = w.contrib[(t + 1) & 1];
    if (staged) {
      k_pr_hot<16, 8, 8><<<hot_grid, PRH_THREADS, 0, s>>>(a, cadj, x, cb, xn, cn);
      k_pr_hotfin<<<1, 1024, 0, s>>>(a, cadj, x, xn, cn, hot_grid);
    } else {
      if (a.U3) {
        k_pr_iter<CsrAdj, 32, 3><<<a.U3, PR_THREADS, 0, s>>>(a, cadj, x, cb, xn, cn);
        k_pr_hubfin<CsrAdj><<<grid_for(a.n3), 256, 0, s>>>(a, cadj, x, xn, cn);
      }
      if (a.U2) k_pr_iter<CsrAdj, 32, 2><<<a.U2, PR_THREADS, 0, s>>>(a, cadj, x, cb, xn, cn);
      if (a.U1) k_pr_iter<CsrAdj, 16, 1><<<a.U1, PR_THREADS, 0, s>>>(a, cadj, x, cb, xn, cn);
      if (a.U0) k_pr_iter<CsrAdj, 8, 0><<<a.U0, PR_THREADS, 0, s>>>(a, cadj, x, cb, xn, cn);
    }
    if (!staged) k_pr_resid<<<1, 1024, 0, s>>>(a);
    if (iter_ms) GV_HIP(hipEventRecord(w.ev[2 * t + 1], s));
    ++launched;
    if (launched % check_every == 0 && launched < max_iter) {
      GV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
      GV_HIP(hipStreamSynchronize(s));
      if (w.h_ctl->done) break;
    }
  }
  GV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  const uint32_t it = w.h_ctl->iters;
  k_pr_unpermute<<<grid_for(V), 256, 0, s>>>(w.x[it & 1], w.rl_newid, V, out_dev);
  if (iters) *iters = (int)it;
  if (residual) *residual = w.h_ctl->residual;
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

// Triangle counting: gv_tc.hip.

// ========================================================= betweenness ==
// Frontier rows longer than BC_HUB slots are taken out of the one-wavefront-
// per-row level kernels and spread over a whole grid (BC_HUB_GRID blocks in
// the backward sweep, one partial each).
constexpr uint32_t BC_HUB = 8192;
constexpr uint32_t BC_HUB_GRID = 128;

__global__ void k_bc_init(int *dist, double *sigma, float *delta, uint32_t V) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x) {
    dist[v] = -1;
    sigma[v] = 0.0;
    delta[v] = 0.f;
  }
}

__global__ void k_bc_src(int *dist, double *sigma, uint32_t *queue, uint32_t src,
                         unsigned int *cursor) {
  dist[src] = 0;
  sigma[src] = 1.0;
  queue[0] = src;
  *cursor = 0;
}

// One wavefront per frontier vertex; Adj is the block store (tombstones
// skipped) or a CSR snapshot.
// Frontier vertices queue[qs, qe); newly reached vertices are appended at
// queue[qapp + ...] (qapp = the end of the whole level when a level is
// issued in several launches).
template <class Adj>
__global__ __launch_bounds__(256) void k_bc_fwd(Adj adj, int *dist, double *sigma,
                                                uint32_t *queue, uint32_t qs, uint32_t qe,
                                                uint32_t qapp, int d, unsigned int *cursor,
                                                uint32_t *hubs, unsigned int *nhub) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint32_t nf = qe - qs;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < nf; i += nwaves) {
    const uint32_t u = queue[qs + i];
    const uint32_t us = adj.len(u);
    if (us > BC_HUB) {  // k_bc_fwd_hub
      if (lane == 0) hubs[atomicAdd(nhub, 1u)] = u;
      continue;
    }
    const double su = sigma[u];
    const uint64_t b = adj.base(u);
    for (uint32_t base = 0; base < us; base += WAVE) {
      uint32_t j = base + lane;
      bool fresh = false;
      uint32_t w = TOMB;
      if (j < us) {
        w = adj.at(b, j);
        if (w != TOMB) {
          // dist moves from -1 to a level once, so a plain read of any other
          // value is final; only a -1 needs the CAS to decide.
          int cur = dist[w];
          if (cur == -1) {
            const int old = atomicCAS(&dist[w], -1, d + 1);
            fresh = old == -1;
            cur = fresh ? d + 1 : old;
          }
          if (cur == d + 1) atomicAdd(&sigma[w], su);
        }
      }
      unsigned long long m = __ballot(fresh);
      uint32_t pos = 0;
      if (lane == 0 && m) pos = atomicAdd(cursor, (unsigned)__popcll(m));
      pos = __shfl(pos, 0, WAVE);
      if (fresh) queue[qapp + pos + mask_lt_count(m)] = w;
    }
  }
}

template <class Adj>
__global__ __launch_bounds__(256) void k_bc_bwd(Adj adj, const int *dist, const double *sigma,
                                                float *delta, const uint32_t *queue, uint32_t qs,
                                                uint32_t qe, int d, uint32_t src, float *bc,
                                                float *bcc, uint32_t *hubs, unsigned int *nhub) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint32_t nf = qe - qs;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < nf; i += nwaves) {
    const uint32_t v = queue[qs + i];
    const uint32_t us = adj.len(v);
    if (us > BC_HUB) {  // k_bc_bwd_hub, k_bc_bwd_hubfin
      if (lane == 0) hubs[atomicAdd(nhub, 1u)] = v;
      continue;
    }
    const double sv = sigma[v];
    const uint64_t b = adj.base(v);
    float s = 0.f, c = 0.f;
    for (uint32_t j = lane; j < us; j += WAVE) {
      uint32_t w = adj.at(b, j);
      if (w != TOMB && dist[w] == d + 1) {
        float coef = (float)(sv / sigma[w]);
        kahan_add(s, c, coef * (1.0f + delta[w]));
      }
    }
#pragma unroll
    for (int k = WAVE / 2; k >= 1; k >>= 1) {
      float s2 = __shfl_xor(s, k, WAVE), c2 = __shfl_xor(c, k, WAVE);
      two_sum_merge(s, c, s2, c2);
    }
    if (lane == 0) {
      float dv = s - c;
      delta[v] = dv;
      if (v != src) {
        float b = bc[v], bcomp = bcc[v];
        kahan_add(b, bcomp, dv);
        bc[v] = b;
        bcc[v] = bcomp;
      }
    }
  }
}

// Hub rows of a forward level piece (listed by k_bc_fwd), one after another,
// each spread over the whole grid in 64-slot rounds: with one wavefront per
// row, a level holding a hub of 10^5-10^6 slots waited for that wavefront.
template <class Adj>
__global__ __launch_bounds__(256) void k_bc_fwd_hub(Adj adj, int *dist, double *sigma,
                                                    uint32_t *queue, uint32_t qapp, int d,
                                                    unsigned int *cursor, const uint32_t *hubs,
                                                    const unsigned int *nhub) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t wv = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  const uint32_t nh = *nhub;
  for (uint32_t h = 0; h < nh; ++h) {
    const uint32_t u = hubs[h];
    const double su = sigma[u];
    const uint32_t us = adj.len(u);
    const uint64_t b = adj.base(u);
    for (uint64_t base = wv * WAVE; base < us; base += nwaves * WAVE) {
      const uint32_t j = (uint32_t)base + lane;
      bool fresh = false;
      uint32_t w = TOMB;
      if (j < us) {
        w = adj.at(b, j);
        if (w != TOMB) {
          int cur = dist[w];
          if (cur == -1) {
            const int old = atomicCAS(&dist[w], -1, d + 1);
            fresh = old == -1;
            cur = fresh ? d + 1 : old;
          }
          if (cur == d + 1) atomicAdd(&sigma[w], su);
        }
      }
      unsigned long long m = __ballot(fresh);
      uint32_t pos = 0;
      if (lane == 0 && m) pos = atomicAdd(cursor, (unsigned)__popcll(m));
      pos = __shfl(pos, 0, WAVE);
      if (fresh) queue[qapp + pos + mask_lt_count(m)] = w;
    }
  }
}

// Backward hub rows: block k of BC_HUB_GRID sums the successors in its slot
// range [k*span, (k+1)*span) of each listed row (Kahan, then a fixed-order
// join over its lanes and wavefronts) into hpart[h * BC_HUB_GRID + k].
template <class Adj>
__global__ __launch_bounds__(256) void k_bc_bwd_hub(Adj adj, const int *dist, const double *sigma,
                                                    const float *delta, int d,
                                                    const uint32_t *hubs, const unsigned int *nhub,
                                                    float2 *hpart) {
  __shared__ float2 s_w[256 / WAVE];
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  const uint32_t nh = *nhub;
  for (uint32_t h = 0; h < nh; ++h) {
    const uint32_t v = hubs[h];
    const double sv = sigma[v];
    const uint32_t us = adj.len(v);
    const uint64_t b = adj.base(v);
    const uint32_t span = (uint32_t)(((us + BC_HUB_GRID - 1) / BC_HUB_GRID + WAVE - 1) / WAVE * WAVE);
    const uint32_t lo = blockIdx.x * span;
    const uint32_t hi = lo + span < us ? lo + span : us;
    float s = 0.f, c = 0.f;
    for (uint32_t j = lo + threadIdx.x; j < hi; j += blockDim.x) {
      const uint32_t w = adj.at(b, j);
      if (w != TOMB && dist[w] == d + 1) {
        const float coef = (float)(sv / sigma[w]);
        kahan_add(s, c, coef * (1.0f + delta[w]));
      }
    }
#pragma unroll
    for (int k = WAVE / 2; k >= 1; k >>= 1) {
      float s2 = __shfl_xor(s, k, WAVE), c2 = __shfl_xor(c, k, WAVE);
      two_sum_merge(s, c, s2, c2);
    }
    if (lane == 0) s_w[wid] = make_float2(s, c);
    __syncthreads();
    if (threadIdx.x == 0) {
      float2 t = s_w[0];
      for (int k = 1; k < 256 / WAVE; ++k) two_sum_merge(t.x, t.y, s_w[k].x, s_w[k].y);
      hpart[(uint64_t)h * BC_HUB_GRID + blockIdx.x] = t;
    }
    __syncthreads();
  }
}

// One wavefront per listed hub row: its BC_HUB_GRID partials joined in block
// order, then delta and BC as in k_bc_bwd.
__global__ __launch_bounds__(256) void k_bc_bwd_hubfin(float *delta, const uint32_t *hubs,
                                                       const unsigned int *nhub,
                                                       const float2 *hpart, uint32_t src,
                                                       float *bc, float *bcc) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint32_t nh = *nhub;
  for (uint64_t h = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; h < nh;
       h += nwaves) {
    if (lane != 0) continue;
    float s = 0.f, c = 0.f;
    for (uint32_t k = 0; k < BC_HUB_GRID; ++k) {
      const float2 q = hpart[h * BC_HUB_GRID + k];
      two_sum_merge(s, c, q.x, q.y);
    }
    const uint32_t v = hubs[h];
    const float dv = s - c;
    delta[v] = dv;
    if (v != src) {
      float bv = bc[v], bcomp = bcc[v];
      kahan_add(bv, bcomp, dv);
      bc[v] = bv;
      bcc[v] = bcomp;
    }
  }
}

// Rows longer than BC_HUB slots in the whole graph (sizes hpart).
template <class Adj>
__global__ void k_bc_hubcount(Adj adj, uint32_t V, unsigned int *n) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    if (adj.len((uint32_t)v) > BC_HUB) atomicAdd(n, 1u);
}

gv_status bc_reserve(BcWork &w, uint32_t V) {
  if (w.V >= V && w.dist) return GV_OK;
  bc_release(w);
  const size_t n = (size_t)(V ? V : 1);
  GV_HIP(hipMalloc(&w.dist, n * sizeof(int)));
  GV_HIP(hipMalloc(&w.sigma, n * sizeof(double)));
  GV_HIP(hipMalloc(&w.delta, n * sizeof(float)));
  GV_HIP(hipMalloc(&w.bcc, n * sizeof(float)));
  GV_HIP(hipMalloc(&w.queue, n * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&w.cursor, sizeof(unsigned int)));
  GV_HIP(hipHostMalloc(&w.h_cursor, sizeof(unsigned int), hipHostMallocDefault));
  GV_HIP(hipMalloc(&w.hubs, n * sizeof(uint32_t)));
  GV_HIP(hipMalloc(&w.nhub, sizeof(unsigned int)));
  // hub partials presized (a serving loop reserves once and must not
  // allocate later); a graph with more hub rows grows them in bc_levels
  const uint64_t hc = n < 16384 ? n : 16384;
  GV_HIP(hipMalloc(&w.hpart, hc * BC_HUB_GRID * sizeof(float2)));
  w.hpart_cap = hc;
  w.V = V;
  return GV_OK;
}

void bc_release(BcWork &w) {
  hipFree(w.dist);
  hipFree(w.sigma);
  hipFree(w.delta);
  hipFree(w.bcc);
  hipFree(w.queue);
  hipFree(w.cursor);
  hipFree(w.hubs);
  hipFree(w.nhub);
  hipFree(w.hpart);
  if (w.h_cursor) hipHostFree(w.h_cursor);
  const uint32_t mb = w.max_blocks;
  w = BcWork();
  w.max_blocks = mb;
}

// Level-synchronous Brandes from each source in turn; BC accumulates into
// out_dev (FP32, Kahan-compensated per vertex).
template <class Adj>
static gv_status bc_levels(const Adj &adj, uint32_t V, const uint32_t *sources,
                           uint32_t n_sources, float *out_dev, hipStream_t s, BcWork &w) {
  if (!V) return GV_OK;
  gv_status rc = bc_reserve(w, V);
  if (rc) return rc;
  const unsigned cap = w.max_blocks ? w.max_blocks : 8192u;
  const uint32_t ch = w.yield ? (1u << 20) : 0xFFFFFFFFu;
  // capacity of the backward hub partials: every row longer than BC_HUB
  fill_u32(w.nhub, 0, 1, s);
  k_bc_hubcount<Adj><<<grid_for(V, 256, 4096), 256, 0, s>>>(adj, V, w.nhub);
  if (hipMemcpyAsync(w.h_cursor, w.nhub, sizeof(unsigned), hipMemcpyDeviceToHost, s) !=
          hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    set_error("bc: HIP error");
    return GV_E_HIP;
  }
  const uint64_t nh_all = *w.h_cursor;
  if (nh_all > w.hpart_cap) {
    hipFree(w.hpart);
    w.hpart = nullptr;
    w.hpart_cap = 0;
    GV_HIP(hipMalloc(&w.hpart, (nh_all + nh_all / 4 + 1) * BC_HUB_GRID * sizeof(float2)));
    w.hpart_cap = nh_all + nh_all / 4 + 1;
  }
  const unsigned hub_grid = cap < 1024u ? cap : 1024u;
  fill_f32(out_dev, 0.f, V, s);
  fill_f32(w.bcc, 0.f, V, s);
  for (uint32_t si = 0; si < n_sources && rc == GV_OK; ++si) {
    const uint32_t src = sources[si];
    k_bc_init<<<grid_for(V), 256, 0, s>>>(w.dist, w.sigma, w.delta, V);
    k_bc_src<<<1, 1, 0, s>>>(w.dist, w.sigma, w.queue, src, w.cursor);
    std::vector<uint32_t> lvl{0u, 1u};
    int d = 0;
    while (true) {
      uint32_t qs = lvl[d], qe = lvl[d + 1];
      if (qe == qs) break;
      fill_u32(w.cursor, 0, 1, s);
      // with a yield hook a level is issued in pieces of at most ch vertices
      // and the caller may hold back the next piece
      for (uint64_t c0 = qs; c0 < qe && rc == GV_OK; c0 += ch) {
        const uint32_t c1 = qe - c0 > ch ? (uint32_t)(c0 + ch) : qe;
        if (w.yield) {
          if (c0 > qs && hipStreamSynchronize(s) != hipSuccess) {
            set_error("bc: HIP error");
            rc = GV_E_HIP;
            break;
          }
          w.yield(w.yield_arg);
        }
        fill_u32(w.nhub, 0, 1, s);
        k_bc_fwd<Adj><<<grid_for((uint64_t)(c1 - c0) * WAVE, 256, cap), 256, 0, s>>>(
            adj, w.dist, w.sigma, w.queue, (uint32_t)c0, c1, qe, d, w.cursor, w.hubs, w.nhub);
        if (nh_all)
          k_bc_fwd_hub<Adj><<<hub_grid, 256, 0, s>>>(adj, w.dist, w.sigma, w.queue, qe, d,
                                                     w.cursor, w.hubs, w.nhub);
      }
      if (rc) break;
      if (hipMemcpyAsync(w.h_cursor, w.cursor, sizeof(unsigned), hipMemcpyDeviceToHost, s) !=
              hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("bc: HIP error");
        rc = GV_E_HIP;
        break;
      }
      lvl.push_back(qe + *w.h_cursor);
      ++d;
    }
    // levels 0..D-1 are non-empty; level D-1 has no successors
    const int D = (int)lvl.size() - 2;
    for (int lev = D - 2; lev >= 0 && rc == GV_OK; --lev) {
      const uint32_t qs = lvl[lev], qe = lvl[lev + 1];
      for (uint64_t c0 = qs; c0 < qe && rc == GV_OK; c0 += ch) {
        const uint32_t c1 = qe - c0 > ch ? (uint32_t)(c0 + ch) : qe;
        if (w.yield) {
          // one piece at a time: finish the previous one before yielding
          if (hipStreamSynchronize(s) != hipSuccess) {
            set_error("bc: HIP error");
            rc = GV_E_HIP;
            break;
          }
          w.yield(w.yield_arg);
        }
        fill_u32(w.nhub, 0, 1, s);
        k_bc_bwd<Adj><<<grid_for((uint64_t)(c1 - c0) * WAVE, 256, cap), 256, 0, s>>>(
            adj, w.dist, w.sigma, w.delta, w.queue, (uint32_t)c0, c1, lev, src, out_dev, w.bcc,
            w.hubs, w.nhub);
        if (nh_all) {
          k_bc_bwd_hub<Adj><<<BC_HUB_GRID, 256, 0, s>>>(adj, w.dist, w.sigma, w.delta, lev,
                                                        w.hubs, w.nhub, w.hpart);
          k_bc_bwd_hubfin<<<64, 256, 0, s>>>(w.delta, w.hubs, w.nhub, w.hpart, src, out_dev,
                                             w.bcc);
        }
      }
    }
  }
  if (rc == GV_OK && hipStreamSynchronize(s) != hipSuccess) {
    set_error("bc: HIP error");
    rc = GV_E_HIP;
  }
  return rc;
}

gv_status bc_run(const Store &st, const uint32_t *sources, uint32_t n_sources, float *out_dev,
                 hipStream_t s) {
  BcWork w;
  gv_status rc = bc_levels(StoreAdj{st}, st.V, sources, n_sources, out_dev, s, w);
  bc_release(w);
  return rc;
}

gv_status bc_run_csr(const uint64_t *off, const uint32_t *cols, uint32_t V,
                     const uint32_t *sources, uint32_t n_sources, float *out_dev, hipStream_t s,
                     BcWork &w) {
  return bc_levels(CsrAdj{off, cols}, V, sources, n_sources, out_dev, s, w);
}

}  // namespace gv
