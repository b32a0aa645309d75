// gv_prims.hip -- device scans and the LSD radix sort used to stage update
// batches (replaces the reference's thrust::exclusive_scan calls at
// graphvine.cu:197,291,329 and its host COO->CSR placement at csr.cu:55-78).
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8).
// Sort: 8-bit LSD passes over 4096-key tiles, stable scatter that ranks keys
// with 64-lane ballot matching.  Up to 2^24 keys the sort is single-sweep
// (Merrill and Adinets, "Onesweep", 2022): one kernel counts every digit of
// every pass (global counts do not depend on key order), then each pass is
// one scatter kernel whose tiles find their offsets by decoupled look-back
// over the tiles before them -- 2 + passes launches instead of 4 per pass.
// Larger sorts take one tile-histogram kernel, one scan of the digit-major
// histogram and one scatter per pass.
#include "gv_internal.h"

namespace gv {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr uint64_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
constexpr int SPINE_THREADS = 1024;

constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr uint64_t SORT_TILE = SORT_THREADS * SORT_ITEMS;

__device__ __forceinline__ uint64_t eff_n(uint64_t n, const uint32_t *n_dev) {
  if (!n_dev) return n;
  uint64_t d = *n_dev;
  return d < n ? d : n;
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const T *in, uint64_t n,
                                                              const uint32_t *n_dev, T *tile_sums) {
  const uint64_t ne = eff_n(n, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  T s = 0;
  if (base < ne) {
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
      uint64_t i = base + (uint64_t)j * SCAN_THREADS + threadIdx.x;
      if (i < ne) s += in[i];
    }
  }
  T tot;
  (void)block_excl_scan<T, SCAN_THREADS>(s, tot);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

// Single block: exclusive scan of tile_sums[0..nt) in place; total at [nt].
template <typename T>
__global__ __launch_bounds__(SPINE_THREADS) void k_scan_spine(T *tile_sums, uint64_t nt) {
  T carry = 0;
  for (uint64_t base = 0; base < nt; base += (uint64_t)SPINE_THREADS * 4) {
    T v[4];
    T s = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint64_t i = base + (uint64_t)threadIdx.x * 4 + j;
      v[j] = i < nt ? tile_sums[i] : 0;
      s += v[j];
    }
    T tot;
    T ex = block_excl_scan<T, SPINE_THREADS>(s, tot) + carry;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      uint64_t i = base + (uint64_t)threadIdx.x * 4 + j;
      if (i < nt) tile_sums[i] = ex;
      ex += v[j];
    }
    carry += tot;
  }
  if (threadIdx.x == 0) tile_sums[nt] = carry;
}

// A thread's SCAN_ITEMS consecutive elements through 16-byte accesses when
// the tile lies wholly inside the array and both arrays are 16-byte aligned
// (vec); element by element otherwise.
template <typename T>
__device__ __forceinline__ void scan_load8(const T *in, uint64_t mine, uint64_t ne, bool vec,
                                           T (&v)[SCAN_ITEMS]) {
  static_assert(SCAN_ITEMS * sizeof(T) % 16 == 0, "whole 16-byte groups");
  if (vec) {
    const uint4 *p = reinterpret_cast<const uint4 *>(in + mine);
#pragma unroll
    for (int r = 0; r < (int)(SCAN_ITEMS * sizeof(T) / 16); ++r) {
      const uint4 q = p[r];
      __builtin_memcpy(reinterpret_cast<uint8_t *>(v) + 16 * r, &q, 16);
    }
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) v[j] = mine + j < ne ? in[mine + j] : 0;
  }
}
template <typename T>
__device__ __forceinline__ void scan_store8(T *out, uint64_t mine, uint64_t ne, bool vec,
                                            const T (&v)[SCAN_ITEMS]) {
  if (vec) {
    uint4 *p = reinterpret_cast<uint4 *>(out + mine);
#pragma unroll
    for (int r = 0; r < (int)(SCAN_ITEMS * sizeof(T) / 16); ++r) {
      uint4 q;
      __builtin_memcpy(&q, reinterpret_cast<const uint8_t *>(v) + 16 * r, 16);
      p[r] = q;
    }
  } else {
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j)
      if (mine + j < ne) out[mine + j] = v[j];
  }
}

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const T *in, T *out, uint64_t n,
                                                             const uint32_t *n_dev,
                                                             const T *tile_sums, uint64_t nt) {
  const uint64_t ne = eff_n(n, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[ne] = tile_sums[nt];
  if (base >= ne) return;
  // blocked: thread t owns [base + t*8, base + t*8 + 8)
  const bool vec = base + SCAN_TILE <= ne && ((((uintptr_t)in) | ((uintptr_t)out)) & 15) == 0;
  T v[SCAN_ITEMS];
  const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
  scan_load8<T>(in, mine, ne, vec, v);
  T s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  T tot;
  T ex = block_excl_scan<T, SCAN_THREADS>(s, tot) + tile_sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const T x = v[j];
    v[j] = ex;
    ex += x;
  }
  scan_store8<T>(out, mine, ne, vec, v);
}

// Up to SCAN_FUSED_TILES tiles: the apply kernel adds up the tile sums
// before its own tile (and all of them for the total) itself, so the spine
// launch goes away.  Tile sums are added in tile order by one fixed
// reduction tree, so the result does not depend on scheduling.
constexpr uint64_t SCAN_FUSED_TILES = 1024;

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply_fused(const T *in, T *out, uint64_t n,
                                                                   const uint32_t *n_dev,
                                                                   const T *tile_sums,
                                                                   uint64_t nt) {
  const uint64_t ne = eff_n(n, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  T before = 0, all = 0;
#pragma unroll
  for (int j = 0; j < (int)(SCAN_FUSED_TILES / SCAN_THREADS); ++j) {
    const uint64_t t = (uint64_t)j * SCAN_THREADS + threadIdx.x;
    const T v = t < nt ? tile_sums[t] : 0;
    all += v;
    if (t < blockIdx.x) before += v;
  }
  before = block_sum<T, SCAN_THREADS>(before);
  all = block_sum<T, SCAN_THREADS>(all);
  if (blockIdx.x == 0 && threadIdx.x == 0) out[ne] = all;
  if (base >= ne) return;
  const bool vec = base + SCAN_TILE <= ne && ((((uintptr_t)in) | ((uintptr_t)out)) & 15) == 0;
  T v[SCAN_ITEMS];
  const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
  scan_load8<T>(in, mine, ne, vec, v);
  T s = 0;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) s += v[j];
  T tot;
  T ex = block_excl_scan<T, SCAN_THREADS>(s, tot) + before;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    const T x = v[j];
    v[j] = ex;
    ex += x;
  }
  scan_store8<T>(out, mine, ne, vec, v);
}

size_t scan_tmp_bytes(uint64_t n) { return (cdiv(n ? n : 1, SCAN_TILE) + 1) * sizeof(uint64_t); }

template <typename T>
static void scan_impl(const T *in, T *out, uint64_t n, const uint32_t *n_dev, void *tmp,
                      hipStream_t s) {
  uint64_t nt = cdiv(n ? n : 1, SCAN_TILE);
  T *ts = (T *)tmp;
  k_scan_reduce<T><<<dim3((unsigned)nt), dim3(SCAN_THREADS), 0, s>>>(in, n, n_dev, ts);
  if (nt <= SCAN_FUSED_TILES) {
    k_scan_apply_fused<T><<<dim3((unsigned)nt), dim3(SCAN_THREADS), 0, s>>>(in, out, n, n_dev, ts,
                                                                          nt);
    return;
  }
  k_scan_spine<T><<<dim3(1), dim3(SPINE_THREADS), 0, s>>>(ts, nt);
  k_scan_apply<T><<<dim3((unsigned)nt), dim3(SCAN_THREADS), 0, s>>>(in, out, n, n_dev, ts, nt);
}

void scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, const uint32_t *n_dev, void *tmp,
              hipStream_t s) {
  scan_impl<uint32_t>(in, out, n, n_dev, tmp, s);
}
void scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, const uint32_t *n_dev, void *tmp,
              hipStream_t s) {
  scan_impl<uint64_t>(in, out, n, n_dev, tmp, s);
}

// ------------------------------------------------------------ radix sort --

__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint64_t *keys, uint64_t n,
                                                             uint32_t shift, uint32_t *hist,
                                                             uint64_t nt) {
  __shared__ uint32_t cnt[256];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE;
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * SORT_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nt + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter.  Wave w of the tile owns keys [w*1024, (w+1)*1024) of the
// tile; round j covers 64 consecutive keys, one per lane, so (w, j, lane) is
// index order and ranks assigned round by round are stable.
// RAW_HIST: offs is the raw digit-major histogram (at most SORT_SMALL_TILES
// tiles) and each block derives its digit offsets from it, so the pass needs
// no separate scan: offset(d, t) = sum of all counts of digits < d + counts
// of digit d in tiles < t.
constexpr uint64_t SORT_SMALL_TILES = 128;

template <bool RAW_HIST>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint64_t *in, uint64_t *out,
                                                                uint64_t n, uint32_t shift,
                                                                const uint32_t *offs, uint64_t nt) {
  constexpr int NW = SORT_THREADS / WAVE;
  __shared__ uint32_t wcnt[NW][256];
  for (int i = threadIdx.x; i < NW * 256; i += SORT_THREADS) (&wcnt[0][0])[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / WAVE;
  const uint32_t lane = lane_id();
  const uint64_t base = (uint64_t)blockIdx.x * SORT_TILE + (uint64_t)w * (WAVE * SORT_ITEMS);
  uint64_t key[SORT_ITEMS];
  uint32_t rnk[SORT_ITEMS];
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    bool valid = i < n;
    uint64_t k = valid ? in[i] : 0;
    uint32_t d = (uint32_t)(k >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bb = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t old = 0;
    if (valid) old = wcnt[w][d];
    uint32_t below = mask_lt_count(peers);
    if (valid && below == 0) wcnt[w][d] = old + (uint32_t)__popcll(peers);
    key[j] = k;
    rnk[j] = old + below;
  }
  __syncthreads();
  // the tile is staged in LDS in digit order (local position = digit start
  // in the tile + the wave's offset within the digit + rank), then written
  // out in LDS order, so consecutive threads store consecutive addresses of
  // a digit's run instead of one scattered word each
  __shared__ uint64_t s_keys[SORT_TILE];
  __shared__ uint32_t s_ls[256], s_gb[256];
  {
    const uint32_t d = threadIdx.x;  // SORT_THREADS == 256 digits
    uint32_t gbase;
    if (RAW_HIST) {
      uint32_t tot = 0, below = 0;
      for (uint64_t t = 0; t < nt; ++t) {
        const uint32_t c = offs[(uint64_t)d * nt + t];
        tot += c;
        if (t < blockIdx.x) below += c;
      }
      uint32_t all;
      gbase = block_excl_scan<uint32_t, SORT_THREADS>(tot, all) + below;
    } else {
      gbase = offs[(uint64_t)d * nt + blockIdx.x];
    }
    uint32_t c = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      const uint32_t x = wcnt[i][d];
      wcnt[i][d] = c;
      c += x;
    }
    uint32_t all;
    const uint32_t ls = block_excl_scan<uint32_t, SORT_THREADS>(c, all);
    s_ls[d] = ls;
    s_gb[d] = gbase - ls;  // modulo 2^32: s_gb[d] + local position is the output index
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      s_keys[s_ls[d] + wcnt[w][d] + rnk[j]] = key[j];
    }
  }
  __syncthreads();
  const uint64_t t0 = (uint64_t)blockIdx.x * SORT_TILE;
  const uint32_t tn = n - t0 < SORT_TILE ? (uint32_t)(n - t0) : (uint32_t)SORT_TILE;
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    const uint32_t p = (uint32_t)j * SORT_THREADS + threadIdx.x;
    if (p < tn) {
      const uint64_t k = s_keys[p];
      out[s_gb[(uint32_t)(k >> shift) & 255u] + p] = k;
    }
  }
}

// ------------------------------------------------- single-sweep sort --
// Tile status word per (pass, tile, digit): bits 31..30 = OS_AGG (the tile's
// own count is in bits 29..0) or OS_PREFIX (the count of the digit in all
// tiles up to and including this one); 0 = not yet published.
constexpr uint32_t OS_AGG = 1u << 30;
constexpr uint32_t OS_PREFIX = 2u << 30;
// A tile whose look-back gave up publishes OS_BAD: its offsets are unknown,
// so every tile after it gives up too instead of trusting a wrong prefix.
constexpr uint32_t OS_BAD = OS_AGG | OS_PREFIX;
constexpr uint32_t OS_VAL = OS_AGG - 1;
// Counts must fit 30 bits; past 2^24 keys launches no longer dominate and
// the classic path's workspace (1/8 of this one's) is used.
constexpr uint64_t OS_MAX_KEYS = 1ull << 24;
constexpr int OS_MAX_PASSES = 8;
// Bounded wait: a look-back that sees no progress for this many polls gives
// up, and the sort reports the failure through the caller's abort word
// instead of hanging.
constexpr uint32_t OS_SPIN_LIMIT = 1u << 26;
constexpr int OS_LOOK = 8;                    // predecessors read per look-back round trip
// keys per thread of a single-sweep tile
constexpr int OS_ITEMS = 16;
constexpr uint64_t OS_TILE = SORT_THREADS * OS_ITEMS;

// Single-sweep workspace (32-bit words): ghist[8][256], tile counters[8],
// error word, then status[passes][nt][256].
struct OsTmp {
  uint32_t *ghist, *ctr, *err, *status;
};
static OsTmp os_tmp(void *tmp) {
  uint32_t *w = (uint32_t *)tmp;
  return OsTmp{w, w + OS_MAX_PASSES * 256, w + OS_MAX_PASSES * 256 + OS_MAX_PASSES,
               w + OS_MAX_PASSES * 256 + OS_MAX_PASSES + 8};
}
static uint64_t os_words(uint64_t nt, uint32_t passes) {
  return OS_MAX_PASSES * 256 + OS_MAX_PASSES + 8 + (uint64_t)passes * nt * 256;
}

__global__ void k_os_clear(uint32_t *w, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    w[i] = 0;
}

// Global digit counts of every pass, from the keys in any order.  A block
// takes chunks of SORT_THREADS * OS_ITEMS keys and loads a chunk's keys
// before counting them, so their latencies overlap.
__global__ __launch_bounds__(SORT_THREADS) void k_os_hist(const uint64_t *keys, uint64_t n,
                                                          uint32_t lo, uint32_t passes,
                                                          uint32_t *ghist) {
  __shared__ uint32_t cnt[OS_MAX_PASSES][256];
  for (uint32_t i = threadIdx.x; i < OS_MAX_PASSES * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
  __syncthreads();
  for (uint64_t base = (uint64_t)blockIdx.x * OS_TILE; base < n;
       base += (uint64_t)gridDim.x * OS_TILE) {
    uint64_t kk[OS_ITEMS];
#pragma unroll
    for (int j = 0; j < OS_ITEMS; ++j) {
      const uint64_t i = base + (uint64_t)j * SORT_THREADS + threadIdx.x;
      kk[j] = i < n ? keys[i] : 0;
    }
#pragma unroll
    for (int j = 0; j < OS_ITEMS; ++j) {
      if (base + (uint64_t)j * SORT_THREADS + threadIdx.x >= n) continue;
      for (uint32_t p = 0; p < passes; ++p)
        atomicAdd(&cnt[p][(uint32_t)(kk[j] >> (lo + 8 * p)) & 255u], 1u);
    }
  }
  __syncthreads();
  for (uint32_t p = 0; p < passes; ++p) {
    const uint32_t c = cnt[p][threadIdx.x];
    if (c) atomicAdd(&ghist[p * 256 + threadIdx.x], c);
  }
}

__device__ __forceinline__ void os_publish(uint32_t *w, uint32_t v) {
  __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t os_peek(uint32_t *w) {
  return __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One pass.  A block takes the next tile in start order (atomic counter), so
// it only ever waits on tiles whose blocks are already running.
// spin_limit bounds the polls of one look-back (OS_SPIN_LIMIT unless a test
// lowers it; 0 gives up at the first predecessor, so every tile after tile 0
// fails and the caller's error path runs while every kernel still exits).
__global__ __launch_bounds__(SORT_THREADS) void k_os_scatter(const uint64_t *in, uint64_t *out,
                                                             uint64_t n, uint32_t shift,
                                                             const uint32_t *ghist, uint32_t *ctr,
                                                             uint32_t *status, uint32_t *err,
                                                             uint32_t spin_limit) {
  constexpr int NW = SORT_THREADS / WAVE;
  __shared__ uint32_t wcnt[NW][256];
  __shared__ uint32_t s_tile, s_bad;
  for (int i = threadIdx.x; i < NW * 256; i += SORT_THREADS) (&wcnt[0][0])[i] = 0;
  if (threadIdx.x == 0) {
    s_tile = atomicAdd(ctr, 1u);
    s_bad = 0;
  }
  __syncthreads();
  const uint32_t tile = s_tile;
  const int w = threadIdx.x / WAVE;
  const uint32_t lane = lane_id();
  const uint64_t base = (uint64_t)tile * OS_TILE + (uint64_t)w * (WAVE * OS_ITEMS);
  // this thread's digit total and the tile's keys, all requested before the
  // ranking so their latencies overlap
  const uint32_t gd = ghist[threadIdx.x];
  uint64_t key[OS_ITEMS];
  uint32_t rnk[OS_ITEMS];
#pragma unroll
  for (int j = 0; j < OS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    key[j] = i < n ? in[i] : 0;
  }
#pragma unroll
  for (int j = 0; j < OS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    bool valid = i < n;
    uint64_t k = key[j];
    uint32_t d = (uint32_t)(k >> shift) & 255u;
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned long long bb = __ballot(valid && ((d >> b) & 1u));
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t old = 0;
    if (valid) old = wcnt[w][d];
    uint32_t below = mask_lt_count(peers);
    if (valid && below == 0) wcnt[w][d] = old + (uint32_t)__popcll(peers);
    rnk[j] = old + below;
  }
  __syncthreads();
  {
    const uint32_t d = threadIdx.x;  // SORT_THREADS == 256 digits
    uint32_t mine = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) mine += wcnt[i][d];
    uint32_t *st = status + (uint64_t)tile * 256 + d;
    uint32_t excl = 0;
    if (tile == 0) {
      os_publish(st, OS_PREFIX | mine);
    } else {
      os_publish(st, OS_AGG | mine);
      // look back OS_LOOK tiles per round trip, newest first; tile 0 always
      // holds a prefix, so the walk ends there at the latest
      uint32_t back = 1;  // next predecessor to read: tile - back
      uint32_t spins = 0;
      bool done = false, bad = spin_limit == 0;
      while (!done && !bad) {
        uint32_t v[OS_LOOK];
#pragma unroll
        for (int r = 0; r < OS_LOOK; ++r)
          v[r] = back + r <= tile ? os_peek(st - (uint64_t)(back + r) * 256) : 0u;
        bool stalled = false;
#pragma unroll
        for (int r = 0; r < OS_LOOK; ++r) {
          if (done || bad || stalled || back > tile) continue;
          if (v[r] == 0) {
            stalled = true;  // not published yet: read it again
            continue;
          }
          if ((v[r] & OS_BAD) == OS_BAD) {
            bad = true;  // a predecessor gave up: its count is unknown
            continue;
          }
          excl += v[r] & OS_VAL;
          ++back;
          if (v[r] & OS_PREFIX) done = true;
        }
        if (stalled && !done && !bad) {
          if (++spins > spin_limit) {
            bad = true;
            continue;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (bad) {
        // offsets unknown: this tile writes nothing, and its status tells
        // the tiles after it to give up as well
        atomicOr(err, 1u);
        s_bad = 1;
        os_publish(st, OS_BAD);
      } else {
        os_publish(st, OS_PREFIX | (excl + mine));
      }
    }
    // digit base: counts of all smaller digits in this pass
    uint32_t all;
    uint32_t acc = block_excl_scan<uint32_t, SORT_THREADS>(gd, all) + excl;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      uint32_t c = wcnt[i][d];
      wcnt[i][d] = acc;
      acc += c;
    }
  }
  __syncthreads();
  if (s_bad) return;
#pragma unroll
  for (int j = 0; j < OS_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      out[wcnt[w][d] + rnk[j]] = key[j];
    }
  }
}

// After the passes of a single-sweep sort: a look-back that gave up leaves
// err set; the caller's sticky abort word then takes ABORT_SORT (unless it
// already holds a code), so the operation fails as a whole.
__global__ void k_os_fold(const uint32_t *err, uint32_t *abort) {
  if (*err) atomicCAS(abort, 0u, ABORT_SORT);
}

static uint32_t g_sort_spin_limit = OS_SPIN_LIMIT;
void sort_set_spin_limit(uint32_t limit) { g_sort_spin_limit = limit; }

gv_status sort_check(const void *tmp, uint64_t n, uint32_t lo, uint32_t hi, hipStream_t s) {
  const uint32_t passes = (hi - lo + 7) / 8;
  if (n <= 1 || n > OS_MAX_KEYS || passes > (uint32_t)OS_MAX_PASSES) return GV_OK;
  uint32_t err = 0;
  GV_HIP(hipMemcpyAsync(&err, os_tmp(const_cast<void *>(tmp)).err, sizeof(err),
                        hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  if (err) {
    set_error("radix sort failed: a tile's look-back gave up; results discarded");
    return GV_E_HIP;
  }
  return GV_OK;
}

// Workspace for any sort of at most n keys: callers size it once and sort
// fewer keys (a sort of <= 2^24 keys takes the single-sweep path even when
// the workspace was sized for more).
size_t sort_tmp_bytes(uint64_t n) {
  uint64_t nt = cdiv(n ? n : 1, SORT_TILE);
  uint64_t h = 256 * nt + 1;
  size_t classic = h * sizeof(uint32_t) + 256 + scan_tmp_bytes(h);
  const uint64_t ns = n < OS_MAX_KEYS ? n : OS_MAX_KEYS;
  size_t sweep = os_words(cdiv(ns ? ns : 1, OS_TILE), OS_MAX_PASSES) * sizeof(uint32_t);
  return classic > sweep ? classic : sweep;
}

uint64_t *sort_keys(uint64_t *a, uint64_t *b, uint64_t n, uint32_t bits, void *tmp,
                    hipStream_t s, uint32_t *abort, bool cleared) {
  return sort_keys_range(a, b, n, 0, bits, tmp, s, abort, cleared);
}

SortClear sort_clear_span(void *tmp, uint64_t n, uint32_t lo, uint32_t hi) {
  const uint32_t passes = (hi - lo + 7) / 8;
  if (n <= 1 || n > OS_MAX_KEYS || passes > (uint32_t)OS_MAX_PASSES) return SortClear{};
  const OsTmp t = os_tmp(tmp);
  return SortClear{t.ghist, os_words(cdiv(n, OS_TILE), passes), t.err};
}

uint64_t *sort_keys_range(uint64_t *a, uint64_t *b, uint64_t n, uint32_t lo, uint32_t hi,
                          void *tmp, hipStream_t s, uint32_t *abort, bool cleared) {
  if (n <= 1) return a;
  const uint32_t passes = (hi - lo + 7) / 8;
  if (n <= OS_MAX_KEYS && passes <= (uint32_t)OS_MAX_PASSES) {
    const uint64_t nt = cdiv(n, OS_TILE);
    const OsTmp t = os_tmp(tmp);
    const uint64_t words = os_words(nt, passes);
    if (!cleared)
      k_os_clear<<<dim3((unsigned)(cdiv(words, 256) < 1024 ? cdiv(words, 256) : 1024)), 256, 0,
                   s>>>(t.ghist, words);
    const uint64_t hb = cdiv(n, OS_TILE);
    k_os_hist<<<dim3((unsigned)(hb < 1024 ? hb : 1024)), dim3(SORT_THREADS), 0, s>>>(a, n, lo,
                                                                                     passes, t.ghist);
    uint64_t *src = a, *dst = b;
    for (uint32_t p = 0; p < passes; ++p) {
      k_os_scatter<<<dim3((unsigned)nt), dim3(SORT_THREADS), 0, s>>>(
          src, dst, n, lo + 8 * p, t.ghist + p * 256, t.ctr + p, t.status + (uint64_t)p * nt * 256,
          t.err, g_sort_spin_limit);
      uint64_t *x = src;
      src = dst;
      dst = x;
    }
    if (abort) k_os_fold<<<1, 1, 0, s>>>(t.err, abort);
    return src;
  }
  const uint64_t nt = cdiv(n, SORT_TILE);
  uint64_t h = 256 * nt;
  uint32_t *hist = (uint32_t *)tmp;
  void *stmp = (void *)(((uintptr_t)(hist + h + 1) + 255) & ~(uintptr_t)255);
  uint64_t *src = a, *dst = b;
  for (uint32_t shift = lo; shift < hi; shift += 8) {
    k_radix_hist<<<dim3((unsigned)nt), dim3(SORT_THREADS), 0, s>>>(src, n, shift, hist, nt);
    if (nt <= SORT_SMALL_TILES) {
      k_radix_scatter<true><<<dim3((unsigned)nt), dim3(SORT_THREADS), 0, s>>>(src, dst, n, shift,
                                                                              hist, nt);
    } else {
      scan_u32(hist, hist, h, nullptr, stmp, s);
      k_radix_scatter<false><<<dim3((unsigned)nt), dim3(SORT_THREADS), 0, s>>>(src, dst, n, shift,
                                                                               hist, nt);
    }
    uint64_t *t = src;
    src = dst;
    dst = t;
  }
  return src;
}

}  // namespace gv
