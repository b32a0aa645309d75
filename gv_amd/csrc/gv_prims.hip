// gv_prims.hip -- device scans and the LSD radix sort used to stage update
// batches (replaces the reference's thrust::exclusive_scan calls at
// graphvine.cu:197,291,329 and its host COO->CSR placement at csr.cu:55-78).
//
// Scan: reduce-then-scan over 2048-element tiles (256 threads x 8).
// Sort: 8-bit LSD passes over 4096-key tiles; per pass one tile-histogram
// kernel, one scan of the digit-major histogram, one stable scatter that
// ranks keys with 64-lane ballot matching.
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

template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const T *in, T *out, uint64_t n,
                                                             const uint32_t *n_dev,
                                                             const T *tile_sums, uint64_t nt) {
  const uint64_t ne = eff_n(n, n_dev);
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE;
  if (blockIdx.x == 0 && threadIdx.x == 0) out[ne] = tile_sums[nt];
  if (base >= ne) return;
  // blocked: thread t owns [base + t*8, base + t*8 + 8)
  T v[SCAN_ITEMS];
  T s = 0;
  const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t i = mine + j;
    v[j] = i < ne ? in[i] : 0;
    s += v[j];
  }
  T tot;
  T ex = block_excl_scan<T, SCAN_THREADS>(s, tot) + tile_sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t i = mine + j;
    if (i < ne) out[i] = ex;
    ex += v[j];
  }
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
  T v[SCAN_ITEMS];
  T s = 0;
  const uint64_t mine = base + (uint64_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t i = mine + j;
    v[j] = i < ne ? in[i] : 0;
    s += v[j];
  }
  T tot;
  T ex = block_excl_scan<T, SCAN_THREADS>(s, tot) + before;
#pragma unroll
  for (int j = 0; j < SCAN_ITEMS; ++j) {
    uint64_t i = mine + j;
    if (i < ne) out[i] = ex;
    ex += v[j];
  }
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
  {
    const uint32_t d = threadIdx.x;  // SORT_THREADS == 256 digits
    uint32_t acc;
    if (RAW_HIST) {
      uint32_t tot = 0, below = 0;
      for (uint64_t t = 0; t < nt; ++t) {
        const uint32_t c = offs[(uint64_t)d * nt + t];
        tot += c;
        if (t < blockIdx.x) below += c;
      }
      uint32_t all;
      acc = block_excl_scan<uint32_t, SORT_THREADS>(tot, all) + below;
    } else {
      acc = offs[(uint64_t)d * nt + blockIdx.x];
    }
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      uint32_t c = wcnt[i][d];
      wcnt[i][d] = acc;
      acc += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SORT_ITEMS; ++j) {
    uint64_t i = base + (uint64_t)j * WAVE + lane;
    if (i < n) {
      uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      out[wcnt[w][d] + rnk[j]] = key[j];
    }
  }
}

size_t sort_tmp_bytes(uint64_t n) {
  uint64_t nt = cdiv(n ? n : 1, SORT_TILE);
  uint64_t h = 256 * nt + 1;
  return h * sizeof(uint32_t) + 256 + scan_tmp_bytes(h);
}

uint64_t *sort_keys(uint64_t *a, uint64_t *b, uint64_t n, uint32_t bits, void *tmp,
                    hipStream_t s) {
  return sort_keys_range(a, b, n, 0, bits, tmp, s);
}

uint64_t *sort_keys_range(uint64_t *a, uint64_t *b, uint64_t n, uint32_t lo, uint32_t hi,
                          void *tmp, hipStream_t s) {
  if (n <= 1) return a;
  uint64_t nt = cdiv(n, SORT_TILE);
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
