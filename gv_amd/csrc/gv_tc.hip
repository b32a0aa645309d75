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
constexpr uint32_t TC_TAB = 2048;   // LDS hash slots per wavefront (8 KiB)
constexpr uint32_t TC_PART = TC_TAB / 2;
constexpr uint32_t TC_EMPTY = 0xFFFFFFFFu;

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
template <bool PV>
__device__ __forceinline__ void tc_hit(uint32_t u, uint32_t v, uint32_t w, unsigned long long &cnt,
                                       unsigned long long &own, unsigned long long *pv) {
  ++cnt;
  if (PV) {
    uint32_t mn = u < v ? u : v;
    mn = mn < w ? mn : w;
    if (mn == u) ++own;
    else atomicAdd(&pv[mn], 1ull);
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
                                                  unsigned long long *pv,
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
      // two wedges per lane and step (f and f + 64), so each lane has two
      // independent loads in flight before it probes
      uint32_t k = 0;  // this lane's current block entry (advances monotonically)
      for (uint32_t fb = 0; fb < nwed; fb += 2 * WAVE) {
        uint32_t wv[2] = {TC_EMPTY, TC_EMPTY}, kv[2] = {0, 0};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
          const uint32_t f = fb + r * WAVE + lane;
          if (f < nwed) {
            while (s_pre[wid][k] <= f) ++k;
            const uint32_t start = k ? s_pre[wid][k - 1] : 0u;
            kv[r] = k;
            wv[r] = cols[s_off[wid][k] + (f - start)];
          }
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) {
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
          if (found) tc_hit<PV>(u, s_v[wid][kv[r]], wv[r], cnt, own, pv);
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
  uint32_t *cols = nullptr;
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
      // the item order is the low word of the sorted keys
      uint32_t *order = reinterpret_cast<uint32_t *>(sorted == keys[0] ? keys[1] : keys[0]);
      keys_to_cols(sorted, n_work, 32, order, s);
      const unsigned grid = (unsigned)tc_device_cus() * 4;
      if (per_vertex_dev)
        k_tc_count<true><<<grid, 256, 0, s>>>(off, cols, order, wroot, nb, n_work,
                                              (unsigned long long *)per_vertex_dev,
                                              (unsigned long long *)tot_dev);
      else
        k_tc_count<false><<<grid, 256, 0, s>>>(off, cols, order, wroot, nb, n_work, nullptr,
                                               (unsigned long long *)tot_dev);
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
  hipFree(tmp);
  hipFree(stmp);
  hipFree(tot_dev);
  return rc;
}

}  // namespace gv
