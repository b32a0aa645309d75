// gv_tc.hip -- exact triangle counting over the device-resident store
// (static_tc, tc.cu:3-27; per-vertex definition of tc_kernel.cu:27-73).
//
// t[u] = #{(v, w) : u < v < w, v in N(u), w in N(v), u in N(w)} on the
// symmetric (undirected) store: each triangle counted once, attributed to its
// minimum vertex ID.  Exact uint64 counts.
//
// Orientation: every undirected edge is kept once, pointing to the endpoint
// of higher rank (degree, ID), so N+(u) holds only higher-ranked neighbours
// and every triangle is found exactly once, as the wedge u -> v -> w closed
// by w in N+(u).  The oriented lists are built over fixed 1024-slot chunks of
// the block store (one wavefront per chunk, so a hub costs many chunks, not
// one long wavefront) and are not sorted: membership is tested in an LDS
// open-addressing hash of N+(u), built in parts of at most TC_PART keys.
//
// Work: item (u, b) is the b-th block of 64 entries v of N+(u), with all the
// wedges (u, v, w in N+(v)) behind them; its weight is the sum of |N+(v)|
// over the block.  Items are sorted by weight, heaviest first, and dealt to
// a persistent grid in snake order (wave w of W takes w, 2W-1-w, 2W+w, ...),
// so every wavefront gets a near-equal share of wedges and the tail is one
// light item, however the heavy roots are numbered.
#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {

constexpr uint32_t TC_CH = 1024;    // store slots per build item
// LDS hash slots per wavefront (4 KiB): small enough for 8 blocks (32
// wavefronts) per CU; 2048 slots held 4 blocks per CU and counted 3% slower
// even though roots above 512 keys now rescan their wedges once more per part
constexpr uint32_t TC_TAB = 1024;
constexpr uint32_t TC_PART = TC_TAB / 2;
constexpr uint32_t TC_EMPTY = 0xFFFFFFFFu;
// wedges per lane and step of the count, their loads issued together
// (R-MAT 22: 8 is 1% faster than 4; 2 was 2% slower than 4)
constexpr uint32_t TC_INFLIGHT = 8;

__device__ __forceinline__ bool tc_higher(uint32_t dw, uint32_t w, uint32_t du, uint32_t u) {
  return dw > du || (dw == du && w > u);
}

// ------------------------------------------------- oriented list build --
__global__ void k_tc_items(const uint32_t *used, uint32_t V, uint32_t *nit) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    nit[v] = (used[v] + TC_CH - 1) / TC_CH;
}

// ivert[i] = vertex of item i, for istart = exclusive scan of the item counts
__global__ void k_tc_expand(const uint32_t *istart, uint32_t V, uint32_t *ivert) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = istart[v]; i < istart[v + 1]; ++i) ivert[i] = (uint32_t)v;
}

// One wavefront per chunk item: oriented entries among the item's slots.
__global__ __launch_bounds__(256) void k_tc_count_out(Store st, const uint32_t *istart,
                                                      const uint32_t *ivert, uint64_t *ipos) {
  const uint32_t ni = istart[st.V];
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < ni;
       i += nwaves) {
    const uint32_t u = ivert[i];
    const uint32_t us = st.used[u], du = st.deg[u], tbl = st.tbl_off[u];
    const uint32_t lo = ((uint32_t)i - istart[u]) * TC_CH;
    const uint32_t hi = lo + TC_CH < us ? lo + TC_CH : us;
    uint32_t cnt = 0;
    for (uint32_t p = lo; p < hi; p += WAVE) {
      bool keep = false;
      if (p + lane < hi) {
        const uint32_t w = st.slots[slot_index(st.arena, tbl, p + lane, st.log2B, st.B)];
        keep = w != TOMB && tc_higher(st.deg[w], w, du, u);
      }
      cnt += (uint32_t)__popcll(__ballot(keep));
    }
    if (lane == 0) ipos[i] = cnt;
  }
}

// off[u] = position of u's first item
__global__ void k_tc_off(const uint32_t *istart, uint32_t V, const uint64_t *ipos, uint64_t *off) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v <= V;
       v += (uint64_t)gridDim.x * blockDim.x)
    off[v] = ipos[istart[v]];
}

// One wavefront per chunk item: oriented entries written in slot order.
__global__ __launch_bounds__(256) void k_tc_put(Store st, const uint32_t *istart,
                                                const uint32_t *ivert, const uint64_t *ipos,
                                                uint32_t *cols) {
  const uint32_t ni = istart[st.V];
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t i = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < ni;
       i += nwaves) {
    const uint32_t u = ivert[i];
    const uint32_t us = st.used[u], du = st.deg[u], tbl = st.tbl_off[u];
    const uint32_t lo = ((uint32_t)i - istart[u]) * TC_CH;
    const uint32_t hi = lo + TC_CH < us ? lo + TC_CH : us;
    uint64_t dst = ipos[i];
    for (uint32_t p = lo; p < hi; p += WAVE) {
      uint32_t w = TOMB;
      bool keep = false;
      if (p + lane < hi) {
        w = st.slots[slot_index(st.arena, tbl, p + lane, st.log2B, st.B)];
        keep = w != TOMB && tc_higher(st.deg[w], w, du, u);
      }
      const unsigned long long m = __ballot(keep);
      if (keep) cols[dst + mask_lt_count(m)] = w;
      dst += (uint64_t)__popcll(m);
    }
  }
}

// ---------------------------------------------------------- work items --
// Blocks of 64 entries of N+(u), for roots with at least two entries.
__global__ void k_tc_nblk(const uint64_t *off, uint32_t V, uint32_t *nb) {
  for (uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; u < V;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t d = off[u + 1] - off[u];
    nb[u] = d >= 2 ? (uint32_t)((d + WAVE - 1) / WAVE) : 0u;
  }
}

// One wavefront per item: its wedge count, as a sort key that orders items
// by descending weight (weight/16, saturated at 2^20-1) above the item index.
__global__ __launch_bounds__(256) void k_tc_wkeys(const uint64_t *off, const uint32_t *cols,
                                                  const uint32_t *wroot, const uint32_t *wstart,
                                                  uint32_t n_work, uint64_t *keys) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t it = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; it < n_work;
       it += nwaves) {
    const uint32_t u = wroot[it];
    const uint64_t ou = off[u];
    const uint32_t du = (uint32_t)(off[u + 1] - ou);
    const uint32_t j = ((uint32_t)it - wstart[u]) * WAVE + lane;
    uint64_t len = 0;
    if (j < du) {
      const uint32_t v = cols[ou + j];
      len = off[v + 1] - off[v];
    }
    const uint64_t wsum = wave_sum(len);
    if (lane == 0) {
      uint64_t q = wsum >> 4;
      if (q > 0xFFFFFull) q = 0xFFFFFull;
      keys[it] = ((0xFFFFFull - q) << 32) | it;
    }
  }
}

// ------------------------------------------------------------ counting --
// Per-block LDS cache of per-vertex count increments.  The minimum vertex of
// a triangle is concentrated on few IDs (low IDs, hubs), and global atomics
// on one address serialise; a vertex's increments go to an LDS counter once
// it holds a slot (first come, at most PVC_PROBE probes, never evicted) and
// to the global count otherwise.  pvc_flush adds the block's counters to the
// global counts.  Increments are taken mod 2^64 (a decrement is ~0).
constexpr uint32_t PVC_N = 256;
constexpr uint32_t PVC_PROBE = 4;
constexpr uint32_t PVC_EMPTY = 0xFFFFFFFFu;

struct PvCache {
  uint32_t key[PVC_N];
  unsigned long long cnt[PVC_N];
};

template <int NT>
__device__ __forceinline__ void pvc_init(PvCache &c) {
  for (uint32_t i = threadIdx.x; i < PVC_N; i += NT) {
    c.key[i] = PVC_EMPTY;
    c.cnt[i] = 0;
  }
  __syncthreads();
}

__device__ __forceinline__ void pvc_add(PvCache &c, unsigned long long *pv, uint32_t id,
                                        unsigned long long inc) {
  uint32_t h = (id * 0x9E3779B1u) >> 24;  // 8 bits = log2(PVC_N)
  for (uint32_t p = 0; p < PVC_PROBE; ++p) {
    uint32_t k = c.key[h];
    if (k == PVC_EMPTY) k = atomicCAS(&c.key[h], PVC_EMPTY, id);
    if (k == PVC_EMPTY || k == id) {
      atomicAdd(&c.cnt[h], inc);
      return;
    }
    h = (h + 1) & (PVC_N - 1);
  }
  atomicAdd(&pv[id], inc);
}

template <int NT>
__device__ __forceinline__ void pvc_flush(PvCache &c, unsigned long long *pv) {
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < PVC_N; i += NT)
    if (c.key[i] != PVC_EMPTY && c.cnt[i]) atomicAdd(&pv[c.key[i]], c.cnt[i]);
}

// Triangle (u, v, w) closed at root u through oriented edge v -> w, stored
// at cols[pos].  Its minimum vertex is u (counted by the wave) or else
// min(v, w), which the edge alone fixes: the edge's counter takes it and
// k_tc_edge_flush moves every edge counter to its minimum endpoint, one add
// per edge instead of one contended atomic per triangle.
template <bool PV>
__device__ __forceinline__ void tc_hit(uint32_t u, uint32_t v, uint32_t w, uint64_t pos,
                                       unsigned long long &cnt, unsigned long long &own,
                                       uint32_t *ecnt) {
  ++cnt;
  if (PV) {
    if (u < v && u < w) ++own;
    else atomicAdd(&ecnt[pos], 1u);
  }
}

__device__ __forceinline__ uint32_t tc_slot(uint32_t x, uint32_t shift) {
  return (x * 0x9E3779B1u) >> shift;  // multiplicative hash, top bits
}

// Persistent: wave w takes slots w, 2W-1-w, 2W+w, ... of the sorted order.
// Per item: N+(u) goes into the wave's LDS hash (one part of <= TC_PART keys
// at a time); the wedges behind the item's 64 entries v are flattened 64 at
// a time, each lane walking a monotone cursor through the block's prefix of
// |N+(v)|, and every w is looked up.
template <bool PV>
__global__ __launch_bounds__(256) void k_tc_count(const uint64_t *off, const uint32_t *cols,
                                                  const uint32_t *order, const uint32_t *wroot,
                                                  const uint32_t *wstart, uint32_t n_work,
                                                  unsigned long long *pv, uint32_t *ecnt,
                                                  unsigned long long *total) {
  __shared__ uint32_t s_tab[4][TC_TAB];
  __shared__ uint32_t s_pre[4][WAVE];
  __shared__ uint32_t s_v[4][WAVE];
  __shared__ uint64_t s_off[4][WAVE];
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / WAVE;
  uint32_t *tab = s_tab[wid];
  const uint32_t W = gridDim.x * 4;
  const uint32_t gw = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wid);
  unsigned long long cnt = 0;
  for (uint32_t r = 0;; ++r) {
    const uint32_t slot = r * W + ((r & 1) ? W - 1 - gw : gw);
    if (slot >= n_work) break;  // only the last round is partial
    const uint32_t item = order[slot];
    const uint32_t u = wroot[item];
    const uint64_t ou = off[u];
    const uint32_t du = (uint32_t)(off[u + 1] - ou);
    const uint32_t *nu = cols + ou;
    const uint32_t j = (item - wstart[u]) * WAVE + lane;
    uint32_t v = 0, len = 0;
    uint64_t ov = 0;
    if (j < du) {
      v = nu[j];
      ov = off[v];
      len = (uint32_t)(off[v + 1] - ov);
    }
    const uint32_t pre = wave_incl_scan<uint32_t>(len);
    const uint32_t nwed = __shfl(pre, WAVE - 1, WAVE);
    if (nwed == 0) continue;
    s_pre[wid][lane] = pre;
    s_v[wid][lane] = v;
    s_off[wid][lane] = ov;
    unsigned long long own = 0;
    for (uint32_t p0 = 0; p0 < du; p0 += TC_PART) {
      const uint32_t pn = du - p0 < TC_PART ? du - p0 : TC_PART;
      uint32_t lg = 32 - __clz(2 * pn - 1);  // table of >= 2 pn slots, >= 64
      if (lg < 6) lg = 6;
      const uint32_t hsz = 1u << lg, shift = 32 - lg, hmask = hsz - 1;
      for (uint32_t i = lane; i < hsz; i += WAVE) tab[i] = TC_EMPTY;
      for (uint32_t i = lane; i < pn; i += WAVE) {
        const uint32_t x = nu[p0 + i];
        uint32_t h = tc_slot(x, shift);
        while (atomicCAS(&tab[h], TC_EMPTY, x) != TC_EMPTY) h = (h + 1) & hmask;
      }
      // TC_INFLIGHT wedges per lane and step (f, f + 64, ...), so each lane
      // has that many independent loads in flight before it probes
      uint32_t k = 0;  // this lane's current block entry (advances monotonically)
      for (uint32_t fb = 0; fb < nwed; fb += TC_INFLIGHT * WAVE) {
        uint32_t wv[TC_INFLIGHT], kv[TC_INFLIGHT];
        uint64_t pv_pos[TC_INFLIGHT];
#pragma unroll
        for (int r = 0; r < (int)TC_INFLIGHT; ++r) {
          wv[r] = TC_EMPTY;
          kv[r] = 0;
          pv_pos[r] = 0;
        }
#pragma unroll
        for (int r = 0; r < (int)TC_INFLIGHT; ++r) {
          const uint32_t f = fb + r * WAVE + lane;
          if (f < nwed) {
            while (s_pre[wid][k] <= f) ++k;
            const uint32_t start = k ? s_pre[wid][k - 1] : 0u;
            kv[r] = k;
            pv_pos[r] = s_off[wid][k] + (f - start);
            wv[r] = cols[pv_pos[r]];
          }
        }
#pragma unroll
        for (int r = 0; r < (int)TC_INFLIGHT; ++r) {
          if (wv[r] == TC_EMPTY) continue;
          uint32_t h = tc_slot(wv[r], shift);
          bool found = false;
          while (true) {
            const uint32_t e = tab[h];
            if (e == wv[r]) {
              found = true;
              break;
            }
            if (e == TC_EMPTY) break;
            h = (h + 1) & hmask;
          }
          if (found) tc_hit<PV>(u, s_v[wid][kv[r]], wv[r], pv_pos[r], cnt, own, ecnt);
        }
      }
    }
    if (PV) {
      own = wave_sum(own);
      if (lane == 0 && own) atomicAdd(&pv[u], own);
    }
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

// Edge counters to per-vertex counts: one wavefront per oriented row v; the
// edges whose minimum endpoint is v are summed in the wave.
__global__ __launch_bounds__(256) void k_tc_edge_flush(const uint64_t *off, const uint32_t *cols,
                                                       const uint32_t *ecnt, uint32_t V,
                                                       unsigned long long *pv) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = (uint64_t)gridDim.x * blockDim.x / WAVE;
  for (uint64_t v = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; v < V; v += nwaves) {
    const uint64_t o0 = off[v], o1 = off[v + 1];
    unsigned long long mine = 0;
    for (uint64_t p = o0 + lane; p < o1; p += WAVE) {
      const uint32_t c = ecnt[p];
      if (!c) continue;
      const uint32_t w = cols[p];
      if (v < w) mine += c;
      else atomicAdd(&pv[w], (unsigned long long)c);
    }
    mine = wave_sum(mine);
    if (lane == 0 && mine) atomicAdd(&pv[v], mine);
  }
}

static int tc_device_cus() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    n = 256;
  return n;
}

gv_status triangle_count_run(const Store &st, uint64_t *per_vertex_dev, uint64_t *total,
                             hipStream_t s) {
  const uint32_t V = st.V;
  *total = 0;
  if (per_vertex_dev) fill_u64(per_vertex_dev, 0, V, s);
  if (!V) return GV_OK;
  // build items: at most one per vertex plus one per TC_CH slots of the pool
  const uint64_t icap = (uint64_t)V + st.P * (uint64_t)st.B / TC_CH + 1;
  if (icap >= 0xFFFFFFFFull) {
    set_error("triangle count: too many chunk items");
    return GV_E_INVALID;
  }
  uint32_t *istart = nullptr, *ivert = nullptr, *nb = nullptr, *wroot = nullptr;
  uint64_t *ipos = nullptr, *off = nullptr, *keys[2] = {nullptr, nullptr}, *tot_dev = nullptr;
  uint32_t *cols = nullptr, *ecnt = nullptr;
  void *tmp = nullptr, *stmp = nullptr;
  gv_status rc = GV_OK;
  auto fail = [&](hipError_t e, const char *what) {
    if (e == hipSuccess) return false;
    set_error(std::string("triangle count: ") + what + ": " + hipGetErrorString(e));
    rc = e == hipErrorOutOfMemory ? GV_E_NOMEM : GV_E_HIP;
    return true;
  };
  do {
    const size_t tmp_bytes = scan_tmp_bytes(icap + 1) > scan_tmp_bytes((uint64_t)V + 1)
                                 ? scan_tmp_bytes(icap + 1)
                                 : scan_tmp_bytes((uint64_t)V + 1);
    if (fail(hipMalloc(&istart, ((size_t)V + 1) * sizeof(uint32_t)), "alloc") ||
        fail(hipMalloc(&ivert, icap * sizeof(uint32_t)), "alloc") ||
        fail(hipMalloc(&ipos, (icap + 1) * sizeof(uint64_t)), "alloc") ||
        fail(hipMalloc(&off, ((size_t)V + 1) * sizeof(uint64_t)), "alloc") ||
        fail(hipMalloc(&nb, ((size_t)V + 1) * sizeof(uint32_t)), "alloc") ||
        fail(hipMalloc(&tot_dev, sizeof(uint64_t)), "alloc") ||
        fail(hipMalloc(&tmp, tmp_bytes), "alloc"))
      break;
    k_tc_items<<<grid_for(V), 256, 0, s>>>(st.used, V, istart);
    scan_u32(istart, istart, V, nullptr, tmp, s);
    k_tc_expand<<<grid_for(V), 256, 0, s>>>(istart, V, ivert);
    k_tc_count_out<<<grid_for(icap * WAVE, 256, 4096), 256, 0, s>>>(st, istart, ivert, ipos);
    scan_u64(ipos, ipos, icap, istart + V, tmp, s);
    k_tc_off<<<grid_for((uint64_t)V + 1), 256, 0, s>>>(istart, V, ipos, off);
    uint64_t E = 0;
    if (fail(hipMemcpyAsync(&E, off + V, sizeof(uint64_t), hipMemcpyDeviceToHost, s), "copy") ||
        fail(hipStreamSynchronize(s), "sync"))
      break;
    if (fail(hipMalloc(&cols, (E ? E : 1) * sizeof(uint32_t)), "alloc")) break;
    if (per_vertex_dev) {
      if (fail(hipMalloc(&ecnt, (E ? E : 1) * sizeof(uint32_t)), "alloc")) break;
      fill_u32(ecnt, 0, E, s);
    }
    k_tc_put<<<grid_for(icap * WAVE, 256, 4096), 256, 0, s>>>(st, istart, ivert, ipos, cols);
    // work items
    k_tc_nblk<<<grid_for(V), 256, 0, s>>>(off, V, nb);
    scan_u32(nb, nb, V, nullptr, tmp, s);
    uint32_t n_work = 0;
    if (fail(hipMemcpyAsync(&n_work, nb + V, sizeof(uint32_t), hipMemcpyDeviceToHost, s),
             "copy") ||
        fail(hipStreamSynchronize(s), "sync"))
      break;
    fill_u64(tot_dev, 0, 1, s);
    if (n_work) {
      if (fail(hipMalloc(&wroot, (size_t)n_work * sizeof(uint32_t)), "alloc") ||
          fail(hipMalloc(&keys[0], (size_t)n_work * sizeof(uint64_t)), "alloc") ||
          fail(hipMalloc(&keys[1], (size_t)n_work * sizeof(uint64_t)), "alloc") ||
          fail(hipMalloc(&stmp, sort_tmp_bytes(n_work)), "alloc"))
        break;
      k_tc_expand<<<grid_for(V), 256, 0, s>>>(nb, V, wroot);
      k_tc_wkeys<<<grid_for((uint64_t)n_work * WAVE, 256, 4096), 256, 0, s>>>(off, cols, wroot,
                                                                              nb, n_work, keys[0]);
      const uint64_t *sorted = sort_keys_range(keys[0], keys[1], n_work, 32, 52, stmp, s);
      {
        const gv_status src = sort_check(stmp, n_work, 32, 52, s);
        if (src) {
          rc = src;
          break;
        }
      }
      // the item order is the low word of the sorted keys
      uint32_t *order = reinterpret_cast<uint32_t *>(sorted == keys[0] ? keys[1] : keys[0]);
      keys_to_cols(sorted, n_work, 32, order, s);
      const unsigned grid = (unsigned)tc_device_cus() * 8;  // all resident
      if (per_vertex_dev) {
        k_tc_count<true><<<grid, 256, 0, s>>>(off, cols, order, wroot, nb, n_work,
                                              (unsigned long long *)per_vertex_dev, ecnt,
                                              (unsigned long long *)tot_dev);
        k_tc_edge_flush<<<grid_for((uint64_t)V * WAVE, 256, 4096), 256, 0, s>>>(
            off, cols, ecnt, V, (unsigned long long *)per_vertex_dev);
      } else {
        k_tc_count<false><<<grid, 256, 0, s>>>(off, cols, order, wroot, nb, n_work, nullptr,
                                               nullptr, (unsigned long long *)tot_dev);
      }
    }
    if (fail(hipMemcpyAsync(total, tot_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, s), "copy") ||
        fail(hipStreamSynchronize(s), "sync"))
      break;
  } while (0);
  if (rc) hipStreamSynchronize(s);
  hipFree(istart);
  hipFree(ivert);
  hipFree(ipos);
  hipFree(off);
  hipFree(nb);
  hipFree(wroot);
  hipFree(keys[0]);
  hipFree(keys[1]);
  hipFree(cols);
  hipFree(ecnt);
  hipFree(tmp);
  hipFree(stmp);
  hipFree(tot_dev);
  return rc;
}

// ------------------------------------------------------------ TC delta --
// Exact per-batch change of the counts above (the reference's dynamic_tc,
// tc.cu:29-64, recounts only the endpoints of updated edges, which misses
// triangles whose minimum vertex is the third one).
//
// A batch changes exactly the triangles that contain at least one changed
// edge: a new edge for an insert (counted in the graph after it), a removed
// edge for a delete (counted in the graph before it).  Each changed edge
// {x, y} is taken at its owner y, the endpoint with more slots (ties: higher
// ID), and every w in N(x) is looked up in a hash of N(y).  A triangle with
// several changed edges is counted only at the smallest of them (canonical
// key (min << bV) | max), so each one moves t[min(x, y, w)] by exactly one.
//
// Owners are the source segments of the staged batch (every changed edge is
// staged in both directions).  Owners of at most TCD_SMALL slots: one
// wavefront each, N(y) in a wave-private LDS hash.  Larger owners: items
// (owner, part of N(y) of TCD_PART slots, chunk of TCD_KCH staged keys), one
// 1024-thread block each, the part in a 128 KiB LDS hash and the scans of the
// chunk's N(x) flattened over the block -- a hub's adjacency is hashed once
// per batch (per part), not once per changed edge.
constexpr uint32_t TCD_SMALL = 1024;
constexpr uint32_t TCD_TAB = 2048;     // wave-private hash slots (8 KiB)
constexpr uint32_t TCD_PART = 16384;   // owner slots per block-level part
constexpr uint32_t TCD_BTAB = 32768;   // block hash slots (128 KiB)
constexpr uint32_t TCD_KCH = WAVE;     // staged keys per block-level item
constexpr int TCD_BLOCK = 1024;
constexpr int TCD_UNROLL = 4;          // block-level slot loads per thread in flight

struct TcdArgs {
  const uint64_t *keys;        // staged sorted keys (u << bV) | v, both directions
  const uint32_t *flag;        // per key; the first copy of a key with flag == sflag changed
  const uint32_t *seg_start, *seg_src;
  const uint64_t *hset;        // open-addressing set of the changed canonical keys
  uint64_t hmask;              // its size - 1 (a power of two)
  uint32_t hshift;             // 64 - log2(size)
  uint32_t sflag;
  uint32_t neg;                // 1: removed triangles (delete)
  unsigned long long *pv;      // per-vertex counts, updated in place (may be null)
  unsigned long long *delta;   // change of the total (two's complement)
};

__device__ __forceinline__ uint64_t tcd_hash(const TcdArgs &a, uint64_t key) {
  return (key * 0x9E3779B97F4A7C15ull) >> a.hshift;  // multiplicative, top bits
}

// Is canonical key `key` a changed edge?  e is the set's entry at
// tcd_hash(key), already loaded.
__device__ __forceinline__ bool tcd_changed(const TcdArgs &a, uint64_t key, uint64_t h, uint64_t e) {
  while (true) {
    if (e == key) return true;
    if (e == KEY_SENT) return false;
    h = (h + 1) & a.hmask;
    e = a.hset[h];
  }
}

__device__ __forceinline__ uint64_t tcd_canon(uint32_t p, uint32_t q, uint32_t bV) {
  return p < q ? ((uint64_t)p << bV) | q : ((uint64_t)q << bV) | p;
}

// Triangle {x, y, w} through changed edge {x, y} owned by y: count it here
// unless a smaller edge of it also changed (a probe of the changed-edge set
// for each smaller edge).  Triangles whose minimum is the
// owner are counted in own, added once per owner by the caller.
__device__ __forceinline__ void tcd_hit(const TcdArgs &a, uint32_t bV, uint32_t x,
                                        uint32_t y, uint32_t w, unsigned long long &cnt,
                                        unsigned long long &own, PvCache &pc) {
  const uint64_t e0 = tcd_canon(x, y, bV);
  const uint64_t e1 = tcd_canon(x, w, bV), e2 = tcd_canon(y, w, bV);
  const bool c1 = e1 < e0, c2 = e2 < e0;
  if (c1 || c2) {
    // both probes issued before either is used
    const uint64_t h1 = tcd_hash(a, e1), h2 = tcd_hash(a, e2);
    const uint64_t s1 = c1 ? a.hset[h1] : KEY_SENT, s2 = c2 ? a.hset[h2] : KEY_SENT;
    if (c1 && tcd_changed(a, e1, h1, s1)) return;
    if (c2 && tcd_changed(a, e2, h2, s2)) return;
  }
  ++cnt;
  if (a.pv) {
    if (y < x && y < w) ++own;
    else pvc_add(pc, a.pv, x < w ? x : w, a.neg ? ~0ull : 1ull);
  }
}

// Add a wavefront's owner-minimum triangles of owner y.
__device__ __forceinline__ void tcd_own_flush(const TcdArgs &a, PvCache &pc, uint32_t y,
                                              unsigned long long &own) {
  if (!a.pv) return;
  const unsigned long long t = wave_sum(own);
  own = 0;
  if (lane_id() == 0 && t) pvc_add(pc, a.pv, y, a.neg ? 0ull - t : t);
}

__device__ __forceinline__ bool tcd_lookup(const uint32_t *tab, uint32_t w, uint32_t shift,
                                           uint32_t hmask) {
  uint32_t h = tc_slot(w, shift);
  while (true) {
    const uint32_t e = tab[h];
    if (e == w) return true;
    if (e == TC_EMPTY) return false;
    h = (h + 1) & hmask;
  }
}

// Lane view of staged key kk of owner y (of uy slots): x and its slot count
// when the key is a changed edge owned by y, else 0.
__device__ __forceinline__ uint32_t tcd_key(const Store &st, const TcdArgs &a, uint64_t mask,
                                            uint32_t kk, uint32_t k1, uint32_t y, uint32_t uy,
                                            uint32_t &x, uint32_t &tx) {
  x = 0;
  tx = 0;
  if (kk >= k1) return 0;
  const uint64_t key = a.keys[kk];
  x = (uint32_t)(key & mask);
  if (a.flag[kk] != a.sflag || (kk > 0 && a.keys[kk - 1] == key)) return 0;
  const uint32_t ux = st.used[x];
  if (!(uy > ux || (uy == ux && y > x))) return 0;
  tx = st.tbl_off[x];
  return ux;
}

__global__ __launch_bounds__(256) void k_tcd_small(Store st, TcdArgs a, uint32_t bV, Ctl *ctl) {
  __shared__ uint32_t s_tab[4][TCD_TAB];
  __shared__ uint32_t s_pre[4][WAVE], s_x[4][WAVE], s_tx[4][WAVE];
  __shared__ PvCache s_pc;
  if (ctl->abort) return;
  pvc_init<256>(s_pc);
  const uint32_t T = ctl->n_segs;
  const uint64_t mask = (1ull << bV) - 1;
  const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
  uint32_t *tab = s_tab[wid];
  const uint32_t W = gridDim.x * 4;
  unsigned long long cnt = 0, own = 0;
  for (uint32_t sidx = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + wid); sidx < T;
       sidx += W) {
    const uint32_t y = a.seg_src[sidx], uy = st.used[y];
    if (uy > TCD_SMALL) continue;
    const uint32_t k0 = a.seg_start[sidx], k1 = a.seg_start[sidx + 1];
    const uint32_t ty = st.tbl_off[y];
    bool built = false;
    uint32_t shift = 0, hmask = 0;
    for (uint32_t kb = k0; kb < k1; kb += WAVE) {
      uint32_t x, tx;
      const uint32_t ux = tcd_key(st, a, mask, kb + lane, k1, y, uy, x, tx);
      const uint32_t pre = wave_incl_scan<uint32_t>(ux);
      const uint32_t tot = __shfl(pre, WAVE - 1, WAVE);
      if (tot == 0) continue;
      if (!built) {
        // >= 2 uy slots, >= 64 (uy >= 1 here: y has the changed edge)
        uint32_t lg = uy > 32 ? 32 - __clz(2 * uy - 1) : 6;
        const uint32_t hsz = 1u << lg;
        shift = 32 - lg;
        hmask = hsz - 1;
        for (uint32_t i = lane; i < hsz; i += WAVE) tab[i] = TC_EMPTY;
        wave_lds_sync();
        for (uint32_t i = lane; i < uy; i += WAVE) {
          const uint32_t v = st.slots[slot_index(st.arena, ty, i, st.log2B, st.B)];
          if (v == TOMB) continue;
          uint32_t h = tc_slot(v, shift);
          while (atomicCAS(&tab[h], TC_EMPTY, v) != TC_EMPTY) h = (h + 1) & hmask;
        }
        built = true;
      }
      s_pre[wid][lane] = pre;
      s_x[wid][lane] = x;
      s_tx[wid][lane] = tx;
      wave_lds_sync();
      uint32_t k = 0;  // this lane's key (advances monotonically with f)
      for (uint32_t fb = 0; fb < tot; fb += 2 * WAVE) {
        uint32_t w[2] = {TOMB, TOMB}, kx[2] = {0, 0};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const uint32_t f = fb + r * WAVE + lane;
          if (f < tot) {
            while (s_pre[wid][k] <= f) ++k;
            const uint32_t pos = f - (k ? s_pre[wid][k - 1] : 0u);
            kx[r] = k;
            w[r] = st.slots[slot_index(st.arena, s_tx[wid][k], pos, st.log2B, st.B)];
          }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
          if (w[r] != TOMB && tcd_lookup(tab, w[r], shift, hmask))
            tcd_hit(a, bV, s_x[wid][kx[r]], y, w[r], cnt, own, s_pc);
      }
      wave_lds_sync();
    }
    tcd_own_flush(a, s_pc, y, own);
  }
  if (a.pv) pvc_flush<256>(s_pc, a.pv);
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(a.delta, a.neg ? 0ull - cnt : cnt);
}

// Block-level items per owner: parts(uy) x chunks(keys), 0 for small owners.
__global__ void k_tcd_items(Store st, TcdArgs a, uint32_t *cnt, Ctl *ctl) {
  const uint32_t T = ctl->n_segs;
  for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx < T;
       sidx += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t uy = st.used[a.seg_src[sidx]];
    const uint32_t nk = a.seg_start[sidx + 1] - a.seg_start[sidx];
    cnt[sidx] = (ctl->abort || uy <= TCD_SMALL)
                    ? 0u
                    : ((uy + TCD_PART - 1) / TCD_PART) * ((nk + TCD_KCH - 1) / TCD_KCH);
  }
}

__global__ void k_tcd_expand(const uint32_t *start, uint32_t *iseg, Ctl *ctl) {
  const uint32_t T = ctl->n_segs;
  for (uint64_t sidx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; sidx < T;
       sidx += (uint64_t)gridDim.x * blockDim.x)
    for (uint32_t i = start[sidx]; i < start[sidx + 1]; ++i) iseg[i] = (uint32_t)sidx;
}

// Insert the changed edges (the first copy of each u < v key) into the set.
__global__ void k_tcd_hset(TcdArgs a, uint64_t *hset, uint32_t bV, Ctl *ctl) {
  if (ctl->abort) return;
  const uint32_t nv = ctl->n_valid;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t key = a.keys[i];
    if ((key >> bV) >= (key & ((1ull << bV) - 1))) continue;
    if (a.flag[i] != a.sflag || (i > 0 && a.keys[i - 1] == key)) continue;
    uint64_t h = tcd_hash(a, key);
    while (true) {
      const uint64_t prev = atomicCAS((unsigned long long *)&hset[h], (unsigned long long)KEY_SENT,
                                      (unsigned long long)key);
      if (prev == KEY_SENT || prev == key) break;
      h = (h + 1) & a.hmask;
    }
  }
}

__global__ void k_tcd_total(const uint32_t *start, uint32_t *n_items, Ctl *ctl) {
  *n_items = start[ctl->n_segs];
}

__global__ __launch_bounds__(TCD_BLOCK) void k_tcd_big(Store st, TcdArgs a, const uint32_t *start,
                                                      const uint32_t *iseg, uint32_t n_items,
                                                      uint32_t bV, Ctl *ctl) {
  __shared__ uint32_t s_tab[TCD_BTAB];
  __shared__ uint32_t s_pre[WAVE], s_x[WAVE], s_tx[WAVE];
  __shared__ PvCache s_pc;
  if (ctl->abort) return;
  pvc_init<TCD_BLOCK>(s_pc);
  const uint64_t mask = (1ull << bV) - 1;
  const uint32_t lane = lane_id();
  unsigned long long cnt = 0, own = 0;
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const uint32_t sidx = iseg[it], li = it - start[sidx];
    const uint32_t y = a.seg_src[sidx], uy = st.used[y];
    const uint32_t k0 = a.seg_start[sidx], k1 = a.seg_start[sidx + 1];
    const uint32_t nch = (k1 - k0 + TCD_KCH - 1) / TCD_KCH;
    const uint32_t p = li / nch, c = li - p * nch;
    if (threadIdx.x < WAVE) {
      const uint32_t kb = k0 + c * TCD_KCH;
      const uint32_t ke = kb + TCD_KCH < k1 ? kb + TCD_KCH : k1;
      uint32_t x, tx;
      const uint32_t ux = tcd_key(st, a, mask, kb + lane, ke, y, uy, x, tx);
      s_pre[lane] = wave_incl_scan<uint32_t>(ux);
      s_x[lane] = x;
      s_tx[lane] = tx;
    }
    __syncthreads();
    const uint32_t tot = s_pre[WAVE - 1];
    if (tot) {
      const uint32_t p0 = p * TCD_PART;
      const uint32_t pn = uy - p0 < TCD_PART ? uy - p0 : TCD_PART;
      const uint32_t lg = pn > 32 ? 32 - __clz(2 * pn - 1) : 6;
      const uint32_t hsz = 1u << lg, shift = 32 - lg, hmask = hsz - 1;
      for (uint32_t i = threadIdx.x; i < hsz; i += TCD_BLOCK) s_tab[i] = TC_EMPTY;
      __syncthreads();
      const uint32_t ty = st.tbl_off[y];
      // TCD_UNROLL independent slot loads per thread in flight, then inserts
      for (uint32_t ib = 0; ib < pn; ib += TCD_UNROLL * TCD_BLOCK) {
        uint32_t v[TCD_UNROLL];
#pragma unroll
        for (int r = 0; r < TCD_UNROLL; ++r) {
          const uint32_t i = ib + r * TCD_BLOCK + threadIdx.x;
          v[r] = i < pn ? st.slots[slot_index(st.arena, ty, p0 + i, st.log2B, st.B)] : TOMB;
        }
#pragma unroll
        for (int r = 0; r < TCD_UNROLL; ++r) {
          if (v[r] == TOMB) continue;
          uint32_t h = tc_slot(v[r], shift);
          while (atomicCAS(&s_tab[h], TC_EMPTY, v[r]) != TC_EMPTY) h = (h + 1) & hmask;
        }
      }
      __syncthreads();
      for (uint32_t fb = 0; fb < tot; fb += TCD_UNROLL * TCD_BLOCK) {
        uint32_t w[TCD_UNROLL], kx[TCD_UNROLL];
#pragma unroll
        for (int r = 0; r < TCD_UNROLL; ++r) {
          const uint32_t f = fb + r * TCD_BLOCK + threadIdx.x;
          w[r] = TOMB;
          kx[r] = 0;
          if (f < tot) {
            uint32_t lo = 0, hi = WAVE - 1;  // first k with s_pre[k] > f
            while (lo < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              if (s_pre[mid] <= f) lo = mid + 1;
              else hi = mid;
            }
            const uint32_t pos = f - (lo ? s_pre[lo - 1] : 0u);
            kx[r] = lo;
            w[r] = st.slots[slot_index(st.arena, s_tx[lo], pos, st.log2B, st.B)];
          }
        }
#pragma unroll
        for (int r = 0; r < TCD_UNROLL; ++r)
          if (w[r] != TOMB && tcd_lookup(s_tab, w[r], shift, hmask))
            tcd_hit(a, bV, s_x[kx[r]], y, w[r], cnt, own, s_pc);
      }
      tcd_own_flush(a, s_pc, y, own);
    }
    __syncthreads();
  }
  if (a.pv) pvc_flush<TCD_BLOCK>(s_pc, a.pv);
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(a.delta, a.neg ? 0ull - cnt : cnt);
}

void tcd_release(TcdWork &w) {
  hipFree(w.cnt);
  hipFree(w.start);
  hipFree(w.iseg);
  hipFree(w.hset);
  hipFree(w.tmp);
  hipFree(w.dev);
  w = TcdWork{};
}

static gv_status tcd_grow(void **p, size_t &cap, size_t need, const char *what) {
  if (need <= cap) return GV_OK;
  hipFree(*p);
  *p = nullptr;
  cap = 0;
  hipError_t e = hipMalloc(p, need);
  if (e != hipSuccess) return hip_fail(e, what);
  cap = need;
  return GV_OK;
}

gv_status tc_delta(const Store &st, const Scratch &sc, const uint64_t *keys, uint64_t m,
                   uint32_t sflag, int neg, uint64_t *per_vertex_dev, TcdWork &w,
                   long long *delta, hipStream_t s) {
  *delta = 0;
  if (!m) return GV_OK;
  gv_status rc = GV_OK;
  if ((rc = tcd_grow((void **)&w.cnt, w.cnt_bytes, (m + 1) * sizeof(uint32_t), "tc delta alloc")) ||
      (rc = tcd_grow((void **)&w.start, w.start_bytes, (m + 1) * sizeof(uint32_t),
                     "tc delta alloc")) ||
      (rc = tcd_grow(&w.tmp, w.tmp_bytes, scan_tmp_bytes(m + 1), "tc delta alloc")) ||
      (rc = tcd_grow((void **)&w.dev, w.dev_bytes, 2 * sizeof(unsigned long long),
                     "tc delta alloc")) ||
      (rc = tcd_grow((void **)&w.hset, w.hset_bytes,
                     ((size_t)1 << ceil_log2(m < 64 ? 64 : m)) * sizeof(uint64_t),
                     "tc delta alloc")))
    return rc;
  const uint32_t bV = id_bits(st.V);
  // at most m/2 changed canonical edges in a set of >= m slots: load <= 1/2
  const uint32_t lgh = ceil_log2(m < 64 ? 64 : m);
  TcdArgs a{keys, sc.flag, sc.seg_start, sc.seg_src, w.hset, ((uint64_t)1 << lgh) - 1,
            64 - lgh, sflag, neg ? 1u : 0u, (unsigned long long *)per_vertex_dev, w.dev};
  fill_u64((uint64_t *)w.dev, 0, 2, s);
  fill_u64(w.hset, KEY_SENT, (uint64_t)1 << lgh, s);
  k_tcd_hset<<<grid_for(m), 256, 0, s>>>(a, w.hset, bV, st.ctl);
  k_tcd_items<<<grid_for(m), 256, 0, s>>>(st, a, w.cnt, st.ctl);
  scan_u32(w.cnt, w.start, m, &st.ctl->n_segs, w.tmp, s);
  k_tcd_total<<<1, 1, 0, s>>>(w.start, (uint32_t *)(w.dev + 1), st.ctl);
  k_tcd_small<<<grid_for(cdiv(m, 4) * 256, 256, 2048), 256, 0, s>>>(st, a, bV, st.ctl);
  uint32_t n_items = 0;
  GV_HIP(hipMemcpyAsync(&n_items, w.dev + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  if (n_items) {
    if ((rc = tcd_grow((void **)&w.iseg, w.iseg_bytes, (size_t)n_items * sizeof(uint32_t),
                       "tc delta alloc")))
      return rc;
    k_tcd_expand<<<grid_for(m), 256, 0, s>>>(w.start, w.iseg, st.ctl);
    unsigned grid = (unsigned)tc_device_cus();
    if (grid > n_items) grid = n_items;
    k_tcd_big<<<grid, TCD_BLOCK, 0, s>>>(st, a, w.start, w.iseg, n_items, bV, st.ctl);
  }
  unsigned long long d = 0;
  GV_HIP(hipGetLastError());
  GV_HIP(hipMemcpyAsync(&d, w.dev, sizeof(d), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  *delta = (long long)d;
  return GV_OK;
}

}  // namespace gv
