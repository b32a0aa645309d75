// gv_pagerank_pb.hip -- PageRank iterations with propagation blocking over
// the relabelled in-adjacency snapshot (same semantics and stopping rule as
// static_pagerank, pagerank.cu:3-68, pr_kernel.cu:21-45).
//
// Why: in the flat stream (gv_pagerank_flat.hip) every entry whose source
// is not among the LDS-staged highest out-degrees gathers one 4-byte
// contribution from a 16 MiB-plus vector, one L2 request per entry; at
// R-MAT scale 22 after the update stream that is 58% of the entries, and the
// per-CU rate of such divergent requests bounds the iteration.  Here those
// cold entries never gather from global memory (propagation blocking,
// Beamer, Asanovic and Patterson, IPDPS 2017, arranged for deterministic
// sums):
//   - expansion: cold sources are cut into blocks of PB_SB consecutive
//     labels; for each block, its contributions are staged in LDS and
//     copied out once per out-entry into a values array, in an order that
//     groups the block's entries by destination bin -- sequential writes;
//   - accumulation: the cold entries, in destination-major order, are cut
//     into bins of whole tiles; a bin's values (one contiguous segment
//     per source block) are loaded into LDS, and each entry holds the 15-bit
//     LDS index of its value.  Row sums are then formed exactly as in the
//     flat stream (tiles of 1024 entries, per-lane Kahan sums in stream
//     order, segmented TwoSum scan), so the order of every sum is fixed.
// Entries with a source label below PB_H (the highest out-degrees) form a
// second, hot stream whose entries hold the label itself: the hot
// contributions are staged in LDS once per workgroup and iteration, as in
// the flat stream.  Both streams store 2-byte entries: a 15-bit index and
// the row-start flag.  Every vertex has a row in both streams (a padding
// entry, index PB_PAD in the hot stream and PB_CPAD in the cold one, whose
// LDS slot holds 0, stands in for an empty one), so a stream's row number is
// the vertex label.  A workgroup's cold bins go through two LDS halves: the
// next bin's image is loaded into registers while the current bin's tiles
// are summed from one half, then written into the other.
//
// Per iteration: k_pb_tiles (hot tiles on some workgroups, cold bins on the
// rest) -> k_pb_rows (joins the hot and cold row sums, the update, residual
// partials) -> k_pb_expand (the FP64 residual and the stopping rule on one
// block, the next iteration's expansion on the others).
#include "gv_analytics.h"
#include "gv_store.h"

namespace gv {

// LDS of a tile workgroup, 80 KiB (two workgroups per CU): PB_LDS words
// holding either the hot contributions (labels [0, PB_H), slot PB_PAD = 0,
// the index of the hot stream's padding entries) or two cold bin images of
// up to PB_CPAD values at PB_HALF * h (slot PB_HALF * h + PB_CPAD = 0, the
// cold stream's padding index); the last three words deal the work units.
constexpr uint32_t PB_LDS = 20480;            // words
constexpr uint32_t PB_PAD = 20476;
constexpr uint32_t PB_HALF = 10240;           // stride of the two cold halves
constexpr uint32_t PB_CPAD = 10236;           // zero slot of a cold half (relative)
constexpr uint32_t PB_W_UNIT = 20477;         // hot unit being worked on
constexpr uint32_t PB_W_NEXT = 20478;         // [2]: the cold unit after the current one
static_assert(PB_HALF + PB_CPAD == PB_PAD && PB_W_NEXT + 2 == PB_LDS, "LDS word map");
constexpr uint32_t PB_F16 = 0x8000u;          // row-start flag of a stream entry
constexpr uint32_t PB_IDX = 0x7FFFu;          // index field of a stream entry
// hot source labels: [0, PB_H); a multiple of 4, so the expansion windows
// (PB_H + k * PB_SB) start on 16-byte boundaries
constexpr uint32_t PB_H = PB_PAD & ~3u;
constexpr uint32_t PB_SB = 32768;             // cold sources per expansion block (128 KiB)
// A bin's image holds its cold values in source-block order, each block's
// segment padded to whole groups of 4 (16 B): at most bin_tiles * PF_T +
// 3 * NS slots, which must stay below PB_CPAD.  A bin is one tile per
// wavefront of the workgroup (8 tiles), so that while its tiles are summed
// the next bin's image loads are in flight; that bounds the cold source
// blocks to PB_MAX_NS (V up to 22.3M; beyond, the flat layout is used).
constexpr uint32_t PB_MIN_BIN_TILES = 8;
constexpr uint32_t PB_MAX_NS = 681;
static_assert(PB_MIN_BIN_TILES * PF_T + 3 * PB_MAX_NS < PB_CPAD, "bin image overflows");
__host__ __device__ inline uint32_t pb_bin_tiles(uint32_t NS) {
  (void)NS;
  return PB_MIN_BIN_TILES;
}
// The image is filled in PB_ROUNDS rounds of PB_PER_ROUND 16-byte loads per
// thread, each round's loads issued before its LDS writes (two rounds keep
// the loaded values within the register budget).
constexpr uint32_t PB_PER_ROUND = 5;
constexpr uint32_t PB_ROUNDS = 2;
static_assert(PB_ROUNDS * PB_PER_ROUND * 512 >= PB_PAD / 4, "the rounds cover the hot image");
// A cold bin image in PB_CROUNDS groups of 16 bytes per thread.
constexpr uint32_t PB_CROUNDS = 5;
static_assert(PB_CROUNDS * 512 * 4 >= PB_CPAD, "the rounds cover a cold image");
constexpr uint32_t PB_HOT_UNIT = 64;          // hot tiles per work unit (8 per wavefront)
constexpr uint32_t PB_CH = 1u << 15;          // entries per expansion work item (multiple of 4)
constexpr int PB_THREADS = 512;               // tile workgroups: 8 wavefronts
constexpr int PB_BUILD_THREADS = 1024;
constexpr uint32_t PB_CSRC_FLAG = 0x80000000u;
constexpr uint32_t PB_CSRC_PAD = 0x7FFFFFFFu;
constexpr int PB_ROW_THREADS = 256;
constexpr uint32_t PB_ROW_BLOCKS = 1024;
constexpr uint32_t PB_LONG_BLOCKS = 256;
constexpr uint32_t PB_PARTS = PB_ROW_BLOCKS + PB_LONG_BLOCKS;
// rows spanning at least PB_LONG tiles of a stream are joined by a wavefront
constexpr uint32_t PB_LONG = 8;

// Tiled layout of a 2-byte stream: lane l of a wavefront owns entries
// [16 l, 16 l + 16) of a tile; entries 8 j .. 8 j + 7 of every lane are
// stored contiguously for the wavefront (lane-major), so each of a lane's
// two 16-byte loads is one contiguous 1 KiB read for the wavefront.
__host__ __device__ inline uint64_t pb_phys(uint64_t p) {
  const uint64_t t = p / PF_T, r = p % PF_T;
  const uint64_t l = r / PF_K, kk = r % PF_K;
  return t * PF_T + (kk >> 3) * 512 + l * 8 + (kk & 7);
}

__host__ __device__ inline uint32_t pb_pad4(uint32_t c) { return (c + 3u) & ~3u; }

// ------------------------------------------------------- layout build --
// Row lengths of both streams in label order (a padding entry for an empty
// row).  hscan: exclusive scan of the items' hot counts.
__global__ void k_pb_rowlen(const uint32_t *istart, const uint32_t *hscan, const uint32_t *order,
                            const uint64_t *off, uint32_t V, uint32_t *hl, uint32_t *cl) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = order[r];
    const uint32_t h = hscan[istart[v + 1]] - hscan[istart[v]];
    const uint32_t c = (uint32_t)(off[r + 1] - off[r]) - h;
    hl[r] = h ? h : 1u;
    cl[r] = c ? c : 1u;
  }
}

// Both streams from the snapshot: each entry of an item (a row chunk) goes to
// the hot stream when its source label is below PB_H, else to the cold
// source list, at the positions its row and the hot counts before it give.
// Items are dealt to lanes (wave w of W, step r, lane l takes item
// (r * 64 + l) * W + w); an item of at most PB_SPLIT_LANE entries is split by
// its lane alone, 16 entries per load round, the longer ones by the whole
// wavefront, one after another.
constexpr uint32_t PB_SPLIT_LANE = 64;

__global__ __launch_bounds__(256) void k_pb_split(const uint32_t *istart, const uint32_t *ivert,
                                                  const uint64_t *ipos, uint32_t V,
                                                  const uint32_t *newid, const uint64_t *off,
                                                  const uint32_t *cols, const uint32_t *hscan,
                                                  const uint32_t *hoff, const uint32_t *coff,
                                                  uint16_t *hot, uint32_t *csrc,
                                                  const PbDims *dims) {
  if (dims->fail) return;
  const uint32_t ni = istart[V];
  const uint32_t lane = lane_id();
  const uint64_t W = (uint64_t)gridDim.x * blockDim.x / WAVE;
  const uint64_t w = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
  for (uint64_t r = 0; r * WAVE * W + w < ni; ++r) {
    const uint64_t i = (r * WAVE + lane) * W + w;
    uint32_t len = 0, hb = 0, cb = 0, h0 = 0, c0 = 0;
    uint64_t base = 0;
    if (i < ni) {
      const uint32_t v = ivert[i];
      const uint32_t rw = newid[v];
      const uint64_t rs = ipos[istart[v]];
      len = (uint32_t)(ipos[i + 1] - ipos[i]);
      const uint64_t inrow = ipos[i] - rs;
      base = off[rw] + inrow;
      const uint32_t hin = hscan[i] - hscan[istart[v]];
      h0 = hoff[rw];
      c0 = coff[rw];
      hb = h0 + hin;
      cb = c0 + (uint32_t)inrow - hin;
    }
    if (len <= PB_SPLIT_LANE) {
#pragma unroll 1
      for (uint32_t p = 0; p < len; p += 16) {
        uint32_t e[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) e[k] = p + k < len ? cols[base + p + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
          if (p + k >= len) break;
          if (e[k] < PB_H) {
            hot[pb_phys(hb)] = (uint16_t)(e[k] | (hb == h0 ? PB_F16 : 0u));
            ++hb;
          } else {
            csrc[cb] = e[k] | (cb == c0 ? PB_CSRC_FLAG : 0u);
            ++cb;
          }
        }
      }
    }
    for (unsigned long long m = __ballot(len > PB_SPLIT_LANE); m; m &= m - 1) {
      const int j = __ffsll(m) - 1;
      const uint32_t jlen = __builtin_amdgcn_readlane(len, j);
      const uint64_t jbase = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(base >> 32), j) << 32) |
                             __builtin_amdgcn_readlane((uint32_t)base, j);
      uint32_t jhb = __builtin_amdgcn_readlane(hb, j), jcb = __builtin_amdgcn_readlane(cb, j);
      const uint32_t jh0 = __builtin_amdgcn_readlane(h0, j), jc0 = __builtin_amdgcn_readlane(c0, j);
      for (uint32_t p = 0; p < jlen; p += WAVE) {
        const bool valid = p + lane < jlen;
        const uint32_t e = valid ? cols[jbase + p + lane] : 0u;
        const bool ishot = valid && e < PB_H;
        const bool iscold = valid && !ishot;
        const unsigned long long mh = __ballot(ishot), mc = __ballot(iscold);
        if (ishot) {
          const uint32_t q = jhb + mask_lt_count(mh);
          hot[pb_phys(q)] = (uint16_t)(e | (q == jh0 ? PB_F16 : 0u));
        }
        if (iscold) {
          const uint32_t q = jcb + mask_lt_count(mc);
          csrc[q] = e | (q == jc0 ? PB_CSRC_FLAG : 0u);
        }
        jhb += (uint32_t)__popcll(mh);
        jcb += (uint32_t)__popcll(mc);
      }
    }
  }
}

// Padding entries for rows empty in a stream, and after the last entry up to
// the end of the last tile.
__global__ void k_pb_pads(const uint32_t *istart, const uint32_t *hscan, const uint32_t *order,
                          const uint64_t *off, uint32_t V, const uint32_t *hoff,
                          const uint32_t *coff, uint16_t *hot, uint32_t *csrc,
                          const PbDims *dims) {
  if (dims->fail) return;
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = order[r];
    const uint32_t h = hscan[istart[v + 1]] - hscan[istart[v]];
    const uint32_t c = (uint32_t)(off[r + 1] - off[r]) - h;
    if (!h) hot[pb_phys(hoff[r])] = (uint16_t)(PB_PAD | PB_F16);
    if (!c) csrc[coff[r]] = PB_CSRC_PAD | PB_CSRC_FLAG;
  }
}

__global__ void k_pb_tail(const uint32_t *hoff, const uint32_t *coff, uint32_t V,
                          const PbDims *dims, uint16_t *hot, uint32_t *csrc) {
  if (dims->fail) return;
  const uint64_t hend = dims->hend, cend = dims->cend;
  const uint64_t h0 = hoff[V], c0 = coff[V];
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t st = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t p = h0 + i0; p < hend; p += st) hot[pb_phys(p)] = (uint16_t)PB_PAD;
  for (uint64_t p = c0 + i0; p < cend; p += st) csrc[p] = PB_CSRC_PAD;
}

// tbase[t] = label of the row in progress at the start of tile t (the last
// row whose first entry lies before t * PF_T), -1 for tile 0.  cold: the
// cold stream's tiles, else the hot stream's.
__global__ void k_pb_tbase(const uint32_t *off, uint32_t V, const PbDims *dims, uint32_t cold,
                           int32_t *tbase) {
  if (dims->fail) return;
  const uint32_t n_tiles = cold ? dims->nct : dims->nht;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tiles;
       t += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = t * PF_T;
    uint32_t lo = 0, hi = V;  // first r with off[r] >= e
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (off[mid] < e) lo = mid + 1; else hi = mid;
    }
    tbase[t] = (int32_t)lo - 1;
  }
}

__device__ __forceinline__ uint32_t pb_span(const uint32_t *off, uint32_t r) {
  return (off[r + 1] - 1) / PF_T - off[r] / PF_T;
}

// Rows spanning at least PB_LONG tiles in either stream, in label order.
__global__ void k_pb_longflag(const uint32_t *hoff, const uint32_t *coff, uint32_t V,
                              uint32_t *flag) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x)
    flag[r] = pb_span(hoff, (uint32_t)r) >= PB_LONG || pb_span(coff, (uint32_t)r) >= PB_LONG
                  ? 1u : 0u;
}

__global__ void k_pb_longfill(const uint32_t *lidx, uint32_t V, uint32_t *longrows) {
  for (uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; r < V;
       r += (uint64_t)gridDim.x * blockDim.x)
    if (lidx[r + 1] > lidx[r]) longrows[lidx[r]] = (uint32_t)r;
}

// Cold entries of bin b per source block k, padded to groups of 4:
// cnt[k * NB + b] (k-major, so an exclusive scan gives the expansion
// order: source block, then bin); the bin's image size in groups goes to
// ngrp[b].
// One workgroup per bin; a grid sized for more bins than the layout has
// (a captured call) leaves the others idle.
__global__ __launch_bounds__(PB_BUILD_THREADS) void k_pb_bincnt(const uint32_t *csrc,
                                                                const PbDims *dims, uint32_t NS,
                                                                uint32_t bin, uint32_t *cnt,
                                                                uint32_t *ngrp) {
  __shared__ uint32_t hist[PB_MAX_NS];
  const uint32_t b = blockIdx.x;
  const uint32_t NB = dims->NB;
  const uint64_t cend = dims->cend;
  if (dims->fail || b >= NB) return;
  for (uint32_t k = threadIdx.x; k < NS; k += PB_BUILD_THREADS) hist[k] = 0;
  __syncthreads();
  const uint64_t q0 = (uint64_t)b * bin;
  const uint64_t q1 = q0 + bin < cend ? q0 + bin : cend;
  for (uint64_t q = q0 + threadIdx.x; q < q1; q += PB_BUILD_THREADS) {
    const uint32_t src = csrc[q] & ~PB_CSRC_FLAG;
    if (src != PB_CSRC_PAD) atomicAdd(&hist[(src - PB_H) / PB_SB], 1u);
  }
  __syncthreads();
  uint32_t ng = 0;
  for (uint32_t k = threadIdx.x; k < NS; k += PB_BUILD_THREADS) {
    const uint32_t c = pb_pad4(hist[k]);
    cnt[(uint64_t)k * NB + b] = c;
    ng += c / 4;
  }
  ng = block_sum<uint32_t, PB_BUILD_THREADS>(ng);
  if (threadIdx.x == 0) ngrp[b] = ng;
}

// Expansion order and image index of every cold entry of bin b.
// seg: exclusive scan of the padded counts (k-major), so segment (k, b) of
// the expansion stream starts at seg[k * NB + b] and spans seg[k * NB + b +
// 1] - seg[k * NB + b] slots (a multiple of 4; the padding slots expand
// harmlessly from window offset 0).  In the bin's image the segments follow
// one another in k order (lpre[k] = slots of blocks before k); gaddr lists,
// per group of 4 image slots, where that group lies in the values array (in
// units of 4 floats).  Within a segment entries take places by LDS atomics,
// so that order is not fixed -- but each entry's index names its own value
// and the sums run in stream order, so the results do not depend on it.
__global__ __launch_bounds__(PB_BUILD_THREADS) void k_pb_scatter(
    const uint32_t *csrc, const PbDims *dims, uint32_t NS, uint32_t bin, const uint32_t *seg,
    const uint32_t *gbase, uint32_t *gaddr, uint16_t *lidx, uint16_t *cold) {
  __shared__ uint32_t lpre[PB_MAX_NS];
  __shared__ uint32_t ctr[PB_MAX_NS];
  __shared__ uint32_t sbase[PB_MAX_NS];  // segment (k, b)'s first slot
  const uint32_t b = blockIdx.x;
  const uint32_t NB = dims->NB;
  const uint64_t cend = dims->cend;
  if (dims->fail || b >= NB) return;
  uint32_t carry = 0;
  const uint32_t g0 = gbase[b];
  for (uint32_t k0 = 0; k0 < NS; k0 += PB_BUILD_THREADS) {
    const uint32_t k = k0 + threadIdx.x;
    const uint64_t at = (uint64_t)k * NB + b;
    const uint32_t c = k < NS ? seg[at + 1] - seg[at] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan<uint32_t, PB_BUILD_THREADS>(c, tot) + carry;
    if (k < NS) {
      lpre[k] = ex;
      ctr[k] = 0;
      const uint32_t s0 = seg[at];
      sbase[k] = s0;
      for (uint32_t g = 0; g < c / 4; ++g) gaddr[g0 + ex / 4 + g] = s0 / 4 + g;
    }
    carry += tot;
  }
  __syncthreads();
  const uint64_t q0 = (uint64_t)b * bin;
  const uint64_t q1 = q0 + bin < cend ? q0 + bin : cend;
  for (uint64_t q = q0 + threadIdx.x; q < q1; q += PB_BUILD_THREADS) {
    const uint32_t raw = csrc[q];
    const uint32_t f = raw & PB_CSRC_FLAG ? PB_F16 : 0u;
    const uint32_t src = raw & ~PB_CSRC_FLAG;
    uint32_t idx = PB_CPAD;
    if (src != PB_CSRC_PAD) {
      const uint32_t k = (src - PB_H) / PB_SB;
      const uint32_t pos = atomicAdd(&ctr[k], 1u);
      lidx[sbase[k] + pos] = (uint16_t)(src - PB_H - k * PB_SB);
      idx = lpre[k] + pos;
    }
    cold[pb_phys(q)] = (uint16_t)(idx | f);
  }
  __syncthreads();
  // padding slots of each segment expand from window offset 0
  for (uint32_t k = threadIdx.x; k < NS; k += PB_BUILD_THREADS) {
    const uint64_t at = (uint64_t)k * NB + b;
    const uint32_t s0 = sbase[k], c = seg[at + 1] - s0;
    for (uint32_t p = ctr[k]; p < c; ++p) lidx[s0 + p] = 0;
  }
}

// Expansion work items: the slots of source block k (seg[k * NB] ..
// seg[(k + 1) * NB]) in pieces of at most PB_CH.
// (no items at all after a failure: the expansion then does nothing)
__global__ void k_pb_itemcnt(const uint32_t *seg, uint32_t NS, const PbDims *dims, uint32_t *nit) {
  const uint32_t NB = dims->NB;
  const bool fail = dims->fail != 0;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < NS;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t n = fail ? 0u : seg[(k + 1) * NB] - seg[k * NB];
    nit[k] = (n + PB_CH - 1) / PB_CH;
  }
}

__global__ void k_pb_itemfill(const uint32_t *seg, uint32_t NS, const PbDims *dims,
                              const uint32_t *nit, uint4 *items) {
  if (dims->fail) return;
  const uint32_t NB = dims->NB;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < NS;
       k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t s0 = seg[k * NB], s1 = seg[(k + 1) * NB];
    uint32_t j = nit[k];
    for (uint32_t a = s0; a < s1; a += PB_CH, ++j)
      items[j] = make_uint4((uint32_t)k, a, a + PB_CH < s1 ? a + PB_CH : s1, 0u);
  }
}

// ---------------------------------------------------------- iteration --
struct PbArgs {
  const uint4 *hot, *cold;        // tiled 2-byte streams, 16-byte units
  const int32_t *htbase, *ctbase;
  float2 *Ph, *Qh, *Sh, *Pc, *Qc, *Sc;
  const uint32_t *hoff, *coff;
  const uint32_t *gaddr, *gbase;
  const uint16_t *lidx;
  float *vals;
  const uint4 *items;
  const uint32_t *n_items;
  const float *x, *cb;
  float *xn, *cn;
  const uint32_t *deg_out, *longrows;
  double *partials;
  PrCtl *ctl;
  const PbDims *dims;  // tile, bin and long-row counts of the prepared layout
  uint32_t V, nh, n_htiles, n_ctiles, n_hot_units, n_bins, bin_tiles, n_long, compensated,
      max_iter;
  float alpha, base;
  double eps;
};

// The counts the layout was prepared with, from the device (the host does
// not know them when the call is captured); false after a failed prepare.
__device__ __forceinline__ bool pb_sizes(PbArgs &a) {
  const PbDims *d = a.dims;
  if (d->fail) return false;
  a.n_htiles = d->nht;
  a.n_ctiles = d->nct;
  a.n_hot_units = (d->nht + PB_HOT_UNIT - 1) / PB_HOT_UNIT;
  a.n_bins = d->NB;
  a.n_long = d->n_long;
  return true;
}

// left := left (+) right, left first; compensated or plain
__device__ __forceinline__ void pb_join(float &s, float &c, float s2, float c2, uint32_t comp) {
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

typedef unsigned int pb_u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void pb_load(const uint4 *s4, uint32_t t, uint32_t lane,
                                        uint4 (&q)[2]) {
  const pb_u32x4 *p = reinterpret_cast<const pb_u32x4 *>(s4 + (uint64_t)t * (PF_T / 8) + lane);
  const pb_u32x4 a = __builtin_nontemporal_load(p), b = __builtin_nontemporal_load(p + 64);
  q[0] = make_uint4(a.x, a.y, a.z, a.w);
  q[1] = make_uint4(b.x, b.y, b.z, b.w);
}

// Entry kk (0..15) of a lane's tile part.
__device__ __forceinline__ uint32_t pb_entry(const uint4 (&q)[2], uint32_t kk) {
  const uint4 v = q[kk >> 3];
  const uint32_t e = kk & 7;
  const uint32_t w = e < 2 ? v.x : e < 4 ? v.y : e < 6 ? v.z : v.w;
  return (e & 1) ? (w >> 16) : (w & 0xFFFFu);
}

// Cross-lane moves for the scans below, on the DPP network (no LDS): a lane
// whose source lies outside the pattern, or in a row the row mask leaves
// out, reads 0 -- the identity of every combine used with them.
template <int CTRL, int RM>
__device__ __forceinline__ float pb_dpp(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, RM, 0xF, true));
}
template <int CTRL, int RM>
__device__ __forceinline__ uint32_t pb_dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, RM, 0xF, true);
}
constexpr int DPP_ROW_SHR = 0x110;    // + n: row_shr:n (within rows of 16 lanes)
constexpr int DPP_BCAST15 = 0x142;    // lane 15 of each row to the next row
constexpr int DPP_BCAST31 = 0x143;    // lane 31 to rows 2 and 3
constexpr int DPP_WAVE_SHR1 = 0x138;  // whole wavefront, one lane up

// Inclusive prefix sum over the wavefront.
__device__ __forceinline__ uint32_t pb_wave_incl(uint32_t v) {
  v += pb_dpp<DPP_ROW_SHR + 1, 0xF>(v);
  v += pb_dpp<DPP_ROW_SHR + 2, 0xF>(v);
  v += pb_dpp<DPP_ROW_SHR + 4, 0xF>(v);
  v += pb_dpp<DPP_ROW_SHR + 8, 0xF>(v);
  v += pb_dpp<DPP_BCAST15, 0xA>(v);
  v += pb_dpp<DPP_BCAST31, 0xC>(v);
  return v;
}

// One step of the segmented scan: a lane whose value carries no row start
// takes the value on its left joined in front (left first) and that value's
// row-start flag.
template <int CTRL, int RM, bool COMP>
__device__ __forceinline__ void pb_seg_step(float &vs, float &vc, uint32_t &vf) {
  constexpr uint32_t comp = COMP ? 1u : 0u;
  const float os = pb_dpp<CTRL, RM>(vs), oc = pb_dpp<CTRL, RM>(vc);
  const uint32_t of = pb_dpp<CTRL, RM>(vf);
  float ns = os, nc = oc;
  pb_join(ns, nc, vs, vc, comp);
  vs = vf ? vs : ns;
  vc = vf ? vc : nc;
  vf = vf ? vf : of;
}

// Row sums of tile t of a stream (rows numbered by label) from the lane's
// entries and their image values: rows begun and ended in the tile go to S,
// the tile's first-row and last-row partials to P[t] and Q[t].  Padding
// entries read the image slot holding 0, and Kahan-adding 0 leaves the
// represented sum unchanged, so they need no test.
template <bool COMP>
__device__ __forceinline__ void pb_tile_sums(float2 *S, float2 *P, float2 *Q, uint32_t t,
                                             int32_t tb, const uint4 (&q)[2],
                                             const float *lds, uint32_t lane, uint32_t nrows) {
  constexpr uint32_t comp = COMP ? 1u : 0u;
  uint32_t fm = 0;
  float g[PF_K];
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) {
    const uint32_t e = pb_entry(q, k);
    fm |= (e >> 15) << k;
    g[k] = lds[e & PB_IDX];
  }
  const uint32_t nst = (uint32_t)__popc(fm);
  const uint32_t incl = pb_wave_incl(nst);
  // label of the row in progress at this lane's start
  const int32_t r0 = tb + (int32_t)(incl - nst);
  int32_t row = r0;
  float s = 0.f, c = 0.f, hs = 0.f, hc = 0.f;
  bool seen = false;
#pragma unroll
  for (uint32_t k = 0; k < PF_K; ++k) {
    const bool f = (fm >> k) & 1u;
    // a row begun in this lane ends here
    if (f && seen && (uint32_t)row < nrows) S[row] = make_float2(s, c);
    hs = f && !seen ? s : hs;
    hc = f && !seen ? c : hc;
    row += f ? 1 : 0;
    s = f ? 0.f : s;
    c = f ? 0.f : c;
    seen = seen || f;
    if (COMP) kahan_add(s, c, g[k]);
    else s += g[k];
  }
  if (!seen) {
    hs = s;
    hc = c;
  }
  // segmented inclusive scan over lanes of the value carried out of each
  // lane: its last row's partial (if a row starts in it) or its whole sum
  float vs = seen ? s : hs, vc = seen ? c : hc;
  uint32_t vf = seen ? 1u : 0u;
  pb_seg_step<DPP_ROW_SHR + 1, 0xF, COMP>(vs, vc, vf);
  pb_seg_step<DPP_ROW_SHR + 2, 0xF, COMP>(vs, vc, vf);
  pb_seg_step<DPP_ROW_SHR + 4, 0xF, COMP>(vs, vc, vf);
  pb_seg_step<DPP_ROW_SHR + 8, 0xF, COMP>(vs, vc, vf);
  pb_seg_step<DPP_BCAST15, 0xA, COMP>(vs, vc, vf);
  pb_seg_step<DPP_BCAST31, 0xC, COMP>(vs, vc, vf);
  // what the lanes before this one carry into it (0 for lane 0)
  float cs = pb_dpp<DPP_WAVE_SHR1, 0xF>(vs), cc = pb_dpp<DPP_WAVE_SHR1, 0xF>(vc);
  const uint32_t cf = pb_dpp<DPP_WAVE_SHR1, 0xF>(vf);
  if (seen) {
    // the row in progress at this lane's start ends before its first flag
    pb_join(cs, cc, hs, hc, comp);
    if (cf) {
      if ((uint32_t)r0 < nrows) S[r0] = make_float2(cs, cc);  // begun in an earlier lane
    } else {
      P[t] = make_float2(cs, cc);  // the tile's first row
    }
  }
  const float ls = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vs), 63));
  const float lc = __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, vc), 63));
  const uint32_t lf = (uint32_t)__builtin_amdgcn_readlane((int)vf, 63);
  if (lane == 0) {
    Q[t] = make_float2(ls, lc);
    if (!lf) P[t] = make_float2(ls, lc);  // no row starts in this tile
  }
}

// One round of an image fill: groups g = threadIdx.x + (R * 5 + i) *
// PB_THREADS, i < 5, of a source of n groups (through the group map gmap
// when given, else directly); all five loads are issued before the LDS
// writes.  Groups past n reload the last one and write image slots no entry
// names; no group reaches the zero slot.
template <uint32_t R>
__device__ __forceinline__ void pb_fill_round(float4 *img4, const float4 *src,
                                              const uint32_t *gmap, uint32_t n) {
  static_assert(PB_PER_ROUND == 5, "the round below is written out for five loads");
  const uint32_t b = threadIdx.x + R * PB_PER_ROUND * PB_THREADS;
  const uint32_t g0 = b, g1 = b + PB_THREADS, g2 = b + 2 * PB_THREADS, g3 = b + 3 * PB_THREADS,
                 g4 = b + 4 * PB_THREADS;
  uint32_t a0 = g0 < n ? g0 : n - 1, a1 = g1 < n ? g1 : n - 1, a2 = g2 < n ? g2 : n - 1,
           a3 = g3 < n ? g3 : n - 1, a4 = g4 < n ? g4 : n - 1;
  if (gmap) {
    a0 = gmap[a0];
    a1 = gmap[a1];
    a2 = gmap[a2];
    a3 = gmap[a3];
    a4 = gmap[a4];
  }
  const float4 v0 = src[a0], v1 = src[a1], v2 = src[a2], v3 = src[a3], v4 = src[a4];
  __builtin_amdgcn_sched_barrier(0);
  img4[g0] = v0;
  img4[g1] = v1;
  img4[g2] = v2;
  img4[g3] = v3;
  if (R + 1 < PB_ROUNDS || g4 < PB_PAD / 4) img4[g4] = v4;
}

// A cold bin's group map entries for this thread: group k * PB_THREADS +
// threadIdx.x of the bin's ng groups in ga[k] (0 past the end).
__device__ __forceinline__ void pb_cold_map(uint32_t (&ga)[PB_CROUNDS], const uint32_t *gmap,
                                            uint32_t ng) {
#pragma unroll
  for (uint32_t k = 0; k < PB_CROUNDS; ++k) {
    const uint32_t g = k * PB_THREADS + threadIdx.x;
    ga[k] = g < ng ? gmap[g] : 0u;
  }
}

// The bin's image groups named by ga into registers (past ng: not loaded).
__device__ __forceinline__ void pb_cold_load(float4 (&r)[PB_CROUNDS], const float4 *vals,
                                             const uint32_t (&ga)[PB_CROUNDS], uint32_t ng) {
#pragma unroll
  for (uint32_t k = 0; k < PB_CROUNDS; ++k) {
    const uint32_t g = k * PB_THREADS + threadIdx.x;
    r[k] = g < ng ? vals[ga[k]] : make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// ... and from registers into one LDS half (never its zero slot: an image
// holds at most PB_CPAD values).
__device__ __forceinline__ void pb_cold_store(float4 *half4, const float4 (&r)[PB_CROUNDS],
                                              uint32_t ng) {
#pragma unroll
  for (uint32_t k = 0; k < PB_CROUNDS; ++k) {
    const uint32_t g = k * PB_THREADS + threadIdx.x;
    if (g < ng) half4[g] = r[k];
  }
}

// Work units from one counter per iteration: first the hot units (PB_HOT_UNIT
// tiles of the hot stream, staged hot contributions), then the cold bins (a
// bin's values loaded into an LDS half, then its tiles, one per wavefront).
// A workgroup stages the hot contributions once, at its first hot unit; once
// it has taken a bin it never takes a hot unit again.  Bins go through the
// two halves in turn: the image of the next bin the workgroup was dealt is
// loaded into registers while the current bin's tiles are summed, then
// written into the other half; one barrier per bin orders both.
template <bool COMP>
__global__ __launch_bounds__(PB_THREADS, 4) void k_pb_tiles(PbArgs a) {
  static_assert(PB_THREADS / WAVE == PB_MIN_BIN_TILES, "one cold tile per wavefront and bin");
  __shared__ float s_lds[PB_LDS];
  uint32_t *s_words = reinterpret_cast<uint32_t *>(s_lds);
  if (a.ctl->done || !pb_sizes(a)) return;
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x / WAVE;
  unsigned int *cursor = &a.ctl->cursor[a.ctl->iters & 1u];
  const uint32_t n_units = a.n_hot_units + a.n_bins;
  bool staged = false;
  // the next unit is requested before the current one is worked on, so the
  // counter's round trip overlaps the work
  uint32_t next = 0;
  if (threadIdx.x == 0) next = atomicAdd(cursor, 1u);
  uint32_t u;
  for (;;) {
    if (threadIdx.x == 0) {
      s_words[PB_W_UNIT] = next;
      next = atomicAdd(cursor, 1u);
    }
    __syncthreads();
    u = s_words[PB_W_UNIT];
    __syncthreads();
    if (u >= n_units) return;
    if (u >= a.n_hot_units) break;
    if (!staged) {
      // every load of the stage is issued before the first LDS write; the
      // vector part covers whole groups of 4 labels below nh (groups past
      // them reload the last one, writing image slots no hot entry names)
      const uint32_t nh4 = a.nh / 4;
      if (nh4) {
        const float4 *c4 = reinterpret_cast<const float4 *>(a.cb);
        float4 *img4 = reinterpret_cast<float4 *>(s_lds);
        pb_fill_round<0>(img4, c4, nullptr, nh4);
        pb_fill_round<1>(img4, c4, nullptr, nh4);
      }
      // the last partial group of hot labels, then the zero slot; the
      // vector rounds above rewrite that group's slots with the reloaded
      // last whole group, so its labels are written after a barrier
      if (a.nh & 3u) __syncthreads();
      for (uint32_t i = nh4 * 4 + threadIdx.x; i < a.nh; i += PB_THREADS) s_lds[i] = a.cb[i];
      if (threadIdx.x == 0) s_lds[PB_PAD] = 0.f;
      __syncthreads();
      staged = true;
    }
    // the wavefront's tiles of the unit, the next one's entries loaded
    // before the current one is summed
    constexpr uint32_t STEP = PB_THREADS / WAVE;
    const uint32_t t0 = u * PB_HOT_UNIT + wv;
    uint4 cur[2], nxt[2];
    pb_load(a.hot, t0 < a.n_htiles ? t0 : 0u, lane, cur);
    int32_t tb_cur = a.htbase[t0 < a.n_htiles ? t0 : 0u];
    for (uint32_t t = t0; t < (u + 1) * PB_HOT_UNIT && t < a.n_htiles; t += STEP) {
      const uint32_t tn = t + STEP < (u + 1) * PB_HOT_UNIT && t + STEP < a.n_htiles ? t + STEP : t;
      pb_load(a.hot, tn, lane, nxt);
      const int32_t tb_nxt = a.htbase[tn];
      pb_tile_sums<COMP>(a.Sh, a.Ph, a.Qh, t, tb_cur, cur, s_lds, lane, a.V);
      cur[0] = nxt[0];
      cur[1] = nxt[1];
      tb_cur = tb_nxt;
    }
  }
  // cold bins, from u on (every unit after a cold one is cold).  Three bins
  // are in flight: u's tiles are summed from LDS half h while the image of
  // the next one, n, is loaded into registers (through its group map, read
  // one bin earlier) and the group map of the one after, m, is loaded.
  // Thread 0 deals n and m ahead through two LDS words used in turn.
  const float4 *v4 = reinterpret_cast<const float4 *>(a.vals);
  float4 *img4 = reinterpret_cast<float4 *>(s_lds);
  uint32_t *slot[2] = {&s_words[PB_W_NEXT], &s_words[PB_W_UNIT]};
  if (threadIdx.x == 0) {
    *slot[1] = next;  // n
    next = atomicAdd(cursor, 1u);
  }
  {
    // the first image straight into half 0, and the zero slots of both
    // halves (the hot image held values there)
    const uint32_t b = u - a.n_hot_units;
    const uint32_t g0 = a.gbase[b], ng = a.gbase[b + 1] - g0;
    uint32_t ga[PB_CROUNDS];
    float4 r[PB_CROUNDS];
    pb_cold_map(ga, a.gaddr + g0, ng);
    pb_cold_load(r, v4, ga, ng);
    pb_cold_store(img4, r, ng);
    if (threadIdx.x == 0) {
      s_lds[PB_CPAD] = 0.f;
      s_lds[PB_HALF + PB_CPAD] = 0.f;
    }
  }
  __syncthreads();
  uint32_t n = *slot[1], ng_n = 0;
  uint32_t ga_n[PB_CROUNDS];
  if (n < n_units) {
    const uint32_t g0 = a.gbase[n - a.n_hot_units];
    ng_n = a.gbase[n - a.n_hot_units + 1] - g0;
    pb_cold_map(ga_n, a.gaddr + g0, ng_n);
  }
  for (uint32_t i = 0, h = 0;; ++i, h ^= 1u) {
    // this wavefront's tile of bin u, loaded first
    const uint32_t t = (u - a.n_hot_units) * PB_MIN_BIN_TILES + wv;
    const bool has_t = t < a.n_ctiles;
    uint4 ent[2];
    pb_load(a.cold, has_t ? t : 0u, lane, ent);
    const int32_t tb = a.ctbase[has_t ? t : 0u];
    if (threadIdx.x == 0) {
      *slot[i & 1u] = next;  // m
      next = atomicAdd(cursor, 1u);
    }
    // u's image is complete in half h, and the bin before it is no longer
    // read from half h ^ 1
    __syncthreads();
    const uint32_t m = *slot[i & 1u];
    float4 r[PB_CROUNDS];
    if (n < n_units) pb_cold_load(r, v4, ga_n, ng_n);
    uint32_t ga_m[PB_CROUNDS], ng_m = 0;
    if (m < n_units) {
      const uint32_t g0 = a.gbase[m - a.n_hot_units];
      ng_m = a.gbase[m - a.n_hot_units + 1] - g0;
      pb_cold_map(ga_m, a.gaddr + g0, ng_m);
    }
    if (has_t) pb_tile_sums<COMP>(a.Sc, a.Pc, a.Qc, t, tb, ent, s_lds + h * PB_HALF, lane, a.V);
    if (n >= n_units) break;
    pb_cold_store(img4 + (h ^ 1u) * (PB_HALF / 4), r, ng_n);
    u = n;
    n = m;
    ng_n = ng_m;
#pragma unroll
    for (uint32_t k = 0; k < PB_CROUNDS; ++k) ga_n[k] = ga_m[k];
  }
}

// Sum of row r in one stream (not a long row): S[r] when it begins and ends
// inside a tile without being the tile's last row, else Q of its first tile
// joined with P of the following tiles in order.
__device__ __forceinline__ float2 pb_part(const uint32_t *off, const float2 *S, const float2 *P,
                                          const float2 *Q, uint32_t r, uint32_t V,
                                          uint32_t comp) {
  const uint32_t e0 = off[r], e1 = off[r + 1];
  const uint32_t ta = e0 / PF_T, tb = (e1 - 1) / PF_T;
  if (ta == tb) {
    const bool last = e1 == (ta + 1) * PF_T || r + 1 == V;
    return last ? Q[ta] : S[r];
  }
  float2 p = Q[ta];
  for (uint32_t t = ta + 1; t <= tb; ++t) {
    const float2 x = P[t];
    pb_join(p.x, p.y, x.x, x.y, comp);
  }
  return p;
}

__device__ __forceinline__ double pb_finish(const PbArgs &a, uint32_t r, float2 h, float2 c) {
  pb_join(h.x, h.y, c.x, c.y, a.compensated);
  const float xv = a.base + a.alpha * (h.x - h.y);
  a.xn[r] = xv;
  const uint32_t d = a.deg_out[r];
  a.cn[r] = d ? xv / (float)d : 0.f;
  return fabs((double)xv - (double)a.x[r]);
}

// A long row's part in one stream, by one wavefront: lane l joins P[ta+1+l],
// P[ta+1+l+64], ... in order, the lanes are joined by a fixed butterfly, and
// Q[ta] goes in front.  Rows not long in this stream take the short path.
__device__ __forceinline__ float2 pb_long_part(const uint32_t *off, const float2 *S,
                                               const float2 *P, const float2 *Q, uint32_t r,
                                               uint32_t V, uint32_t comp, uint32_t lane) {
  const uint32_t ta = off[r] / PF_T, tb = (off[r + 1] - 1) / PF_T;
  if (tb - ta < PB_LONG) return pb_part(off, S, P, Q, r, V, comp);
  float s = 0.f, c = 0.f;
  for (uint32_t t = ta + 1 + lane; t <= tb; t += WAVE) {
    const float2 q = P[t];
    pb_join(s, c, q.x, q.y, comp);
  }
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    const float s2 = __shfl_xor(s, d, WAVE), c2 = __shfl_xor(c, d, WAVE);
    if (lane & d) {
      float ns = s2, nc = c2;
      pb_join(ns, nc, s, c, comp);
      s = ns;
      c = nc;
    } else {
      pb_join(s, c, s2, c2, comp);
    }
  }
  float2 p = Q[ta];
  pb_join(p.x, p.y, s, c, comp);
  return p;
}

// Row offsets of both streams for row r: whether either spans PB_LONG
// tiles or more (a long row, joined by a wavefront), else the row's S/P/Q
// reads are those of pb_part.
__global__ __launch_bounds__(PB_ROW_THREADS) void k_pb_rows(PbArgs a) {
  if (a.ctl->done || !pb_sizes(a)) return;
  // the next iteration's tile kernel deals its units from the other counter
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl->cursor[(a.ctl->iters + 1) & 1u] = 0;
  const uint32_t comp = a.compensated;
  double diff = 0.0;
  if (blockIdx.x >= PB_ROW_BLOCKS) {
    const uint32_t lane = lane_id();
    const uint32_t nw = PB_LONG_BLOCKS * (PB_ROW_THREADS / WAVE);
    for (uint32_t j = (blockIdx.x - PB_ROW_BLOCKS) * (PB_ROW_THREADS / WAVE) + threadIdx.x / WAVE;
         j < a.n_long; j += nw) {
      const uint32_t r = a.longrows[j];
      const float2 h = pb_long_part(a.hoff, a.Sh, a.Ph, a.Qh, r, a.V, comp, lane);
      const float2 c = pb_long_part(a.coff, a.Sc, a.Pc, a.Qc, r, a.V, comp, lane);
      if (lane == 0) diff += pb_finish(a, r, h, c);
    }
  } else {
    // R rows per thread and step, each step's loads issued together: the
    // offsets, then the row partials they select, then x and degrees
    constexpr uint32_t R = 4;
    const uint64_t stride = (uint64_t)PB_ROW_BLOCKS * blockDim.x;
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < a.V;
         i0 += R * stride) {
      uint32_t h0[R], h1[R], c0[R], c1[R];
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) {
        const uint64_t i = i0 + k * stride;
        const uint32_t r = i < a.V ? (uint32_t)i : 0u;
        h0[k] = a.hoff[r];
        h1[k] = a.hoff[r + 1];
        c0[k] = a.coff[r];
        c1[k] = a.coff[r + 1];
      }
      float2 hp[R], cp[R];
      float xo[R];
      uint32_t dg[R];
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) {
        const uint64_t i = i0 + k * stride;
        const uint32_t r = i < a.V ? (uint32_t)i : 0u;
        const uint32_t hta = h0[k] / PF_T, htb = (h1[k] - 1) / PF_T;
        const uint32_t cta = c0[k] / PF_T, ctb = (c1[k] - 1) / PF_T;
        // the row's first read in each stream: S[r] when it is not its tile's
        // last row, else Q of its first tile (further tiles are joined below)
        const bool hs = hta == htb && !(h1[k] == (hta + 1) * PF_T || r + 1 == a.V);
        const bool cs = cta == ctb && !(c1[k] == (cta + 1) * PF_T || r + 1 == a.V);
        hp[k] = hs ? a.Sh[r] : a.Qh[hta];
        cp[k] = cs ? a.Sc[r] : a.Qc[cta];
        xo[k] = a.x[r];
        dg[k] = a.deg_out[r];
      }
#pragma unroll
      for (uint32_t k = 0; k < R; ++k) {
        const uint64_t i = i0 + k * stride;
        if (i >= a.V) continue;
        const uint32_t r = (uint32_t)i;
        const uint32_t hta = h0[k] / PF_T, htb = (h1[k] - 1) / PF_T;
        const uint32_t cta = c0[k] / PF_T, ctb = (c1[k] - 1) / PF_T;
        if (htb - hta >= PB_LONG || ctb - cta >= PB_LONG) continue;  // a long row
        float2 h = hp[k], c = cp[k];
        for (uint32_t t = hta + 1; t <= htb; ++t) {
          const float2 x = a.Ph[t];
          pb_join(h.x, h.y, x.x, x.y, comp);
        }
        for (uint32_t t = cta + 1; t <= ctb; ++t) {
          const float2 x = a.Pc[t];
          pb_join(c.x, c.y, x.x, x.y, comp);
        }
        pb_join(h.x, h.y, c.x, c.y, comp);
        const float xv = a.base + a.alpha * (h.x - h.y);
        a.xn[r] = xv;
        a.cn[r] = dg[k] ? xv / (float)dg[k] : 0.f;
        diff += fabs((double)xv - (double)xo[k]);
      }
    }
  }
  const double bs = block_sum<double, PB_ROW_THREADS>(diff);
  if (threadIdx.x == 0) a.partials[blockIdx.x] = bs;
}

// Block 0 (with resid): fixed-order FP64 residual of the iteration just
// finished and the stopping rule (pagerank.cu:56).  Other blocks: the
// expansion of the contributions cb for the next iteration, 4 slots (one
// 8-byte load of window offsets, one 16-byte store of values) per thread
// and step.
__global__ __launch_bounds__(PB_BUILD_THREADS) void k_pb_expand(PbArgs a, uint32_t resid) {
  __shared__ float s_win[PB_SB];
  if (!pb_sizes(a)) return;
  if (resid) {
    if (blockIdx.x == 0) {
      if (a.ctl->done) return;
      double r = 0.0;
      for (uint32_t i = threadIdx.x; i < PB_PARTS; i += PB_BUILD_THREADS) r += a.partials[i];
      r = block_sum<double, PB_BUILD_THREADS>(r);
      if (threadIdx.x == 0) {
        a.ctl->residual = r;
        const unsigned it = a.ctl->iters + 1;
        a.ctl->iters = it;
        if (r < a.eps || it >= a.max_iter) a.ctl->done = 1;
      }
      return;
    }
  }
  if (a.ctl->done) return;
  const uint32_t nb = gridDim.x - resid, bid = blockIdx.x - resid;
  const uint32_t n_items = *a.n_items;
  const uint2 *l4 = reinterpret_cast<const uint2 *>(a.lidx);
  float4 *v4 = reinterpret_cast<float4 *>(a.vals);
  // a contiguous run of items per workgroup: consecutive items mostly share
  // their source block, whose window is then staged once
  const uint32_t i0 = (uint32_t)((uint64_t)n_items * bid / nb);
  const uint32_t i1 = (uint32_t)((uint64_t)n_items * (bid + 1) / nb);
  uint32_t staged_k = 0xFFFFFFFFu;
  for (uint32_t it = i0; it < i1; ++it) {
    const uint4 item = a.items[it];
    const uint32_t s0 = PB_H + item.x * PB_SB;
    const uint32_t ns = a.V - s0 < PB_SB ? a.V - s0 : PB_SB;
    if (item.x != staged_k) {
      if (staged_k != 0xFFFFFFFFu) __syncthreads();  // the previous window's last reads
      staged_k = item.x;
      // the window's loads all issued before its LDS writes (s0 is a multiple
      // of 4, so whole groups are 16-byte loads)
      constexpr uint32_t NW = PB_SB / 4 / PB_BUILD_THREADS;
      const float4 *c4 = reinterpret_cast<const float4 *>(a.cb + s0);
      float4 wv[NW];
#pragma unroll
      for (uint32_t j = 0; j < NW; ++j) {
        const uint32_t i = threadIdx.x + j * PB_BUILD_THREADS;
        wv[j] = 4 * i + 3 < ns ? c4[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      }
      __builtin_amdgcn_sched_barrier(0);
      float4 *w4 = reinterpret_cast<float4 *>(s_win);
#pragma unroll
      for (uint32_t j = 0; j < NW; ++j) w4[threadIdx.x + j * PB_BUILD_THREADS] = wv[j];
      // a partial last group was written as zeros above: its labels go in
      // after a barrier
      if (ns & 3u) __syncthreads();
      for (uint32_t i = (ns & ~3u) + threadIdx.x; i < ns; i += PB_BUILD_THREADS)
        s_win[i] = a.cb[s0 + i];
      __syncthreads();
    }
    constexpr uint32_t U = 4;
    const uint32_t ga = item.y / 4, gb = item.z / 4;
    for (uint32_t g = ga + threadIdx.x; g < gb; g += U * PB_BUILD_THREADS) {
      uint2 l[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t q = g + u * PB_BUILD_THREADS;
        l[u] = q < gb ? l4[q] : make_uint2(0u, 0u);
      }
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t q = g + u * PB_BUILD_THREADS;
        if (q < gb)
          v4[q] = make_float4(s_win[l[u].x & 0xFFFFu], s_win[l[u].x >> 16],
                              s_win[l[u].y & 0xFFFFu], s_win[l[u].y >> 16]);
      }
    }
  }
}

// ------------------------------------------------------------ workspace --
static int pb_device_cus() {
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

// Grow-only device buffer (headroom of an eighth, at least `floor` bytes).
// A move bumps the workspace's generation (captured calls hold addresses).
static gv_status pb_grow(PrWork &w, void **p, size_t &cap, size_t need, size_t floor = 0) {
  if (*p && need <= cap) return GV_OK;
  ++w.gen;
  hipFree(*p);
  *p = nullptr;
  cap = 0;
  size_t c = need + need / 8 + 256;
  if (c < floor) c = floor;
  GV_HIP(hipMalloc(p, c));
  cap = c;
  return GV_OK;
}

void pb_release(PrWork &w) {
  void **bufs[] = {(void **)&w.pb_hot,   (void **)&w.pb_cold,  (void **)&w.pb_csrc,
                   (void **)&w.pb_lidx,  (void **)&w.pb_vals,  (void **)&w.pb_hoff,
                   (void **)&w.pb_coff,  (void **)&w.pb_htb,   (void **)&w.pb_ctb,
                   (void **)&w.pb_Sh,    (void **)&w.pb_Sc,    (void **)&w.pb_PQh,
                   (void **)&w.pb_PQc,   (void **)&w.pb_hitem, (void **)&w.pb_cnt,
                   (void **)&w.pb_ngrp,  (void **)&w.pb_gaddr, (void **)&w.pb_items,
                   (void **)&w.pb_nit,   (void **)&w.pb_lidxr, (void **)&w.pb_long,
                   (void **)&w.pb_tmp,   (void **)&w.pb_partials, (void **)&w.pb_dims};
  for (void **b : bufs) {
    hipFree(*b);
    *b = nullptr;
  }
  if (w.h_pb_dims) hipHostFree(w.h_pb_dims);
  w.h_pb_dims = nullptr;
  w.pb_caps = PbCaps{};
  w.pb_ready = 0;
  w.pb_res_V = 0;
  w.pb_res_E = 0;
  ++w.gen;
}

// Capacities (entries) the device-side size checks of a captured call use;
// a plain call grows its buffers to the sizes it reads back instead.
struct PbLimits {
  unsigned long long hot, cold, htb, ctb, pqh, pqc, cnt, ngrp, lidx, vals, gaddr, items;
  uint32_t check;
};

static PbLimits pb_limits(const PrWork &w, bool capture) {
  PbLimits l{};
  l.check = capture ? 1u : 0u;
  if (!capture) return l;
  const PbCaps &c = w.pb_caps;
  l.hot = c.hot / 2;
  l.cold = std::min(c.cold / 2, c.csrc / 4);
  l.htb = c.htb / 4;
  l.ctb = c.ctb / 4;
  l.pqh = c.PQh / 16;
  l.pqc = c.PQc / 16;
  l.cnt = c.cnt / 4;
  l.ngrp = c.ngrp / 4;
  l.lidx = c.lidx / 2;
  l.vals = c.vals / 4;
  l.gaddr = c.gaddr / 4;
  l.items = c.items / 16;
  return l;
}

__global__ void k_pb_dims_reset(PbDims *d) {
  PbDims z = {};
  *d = z;
}

// Stream lengths, tile and bin counts and the table size, once both streams'
// row offsets are scanned; with l.check, a layout that exceeds the reserved
// buffers or the layout's 32-bit positions sets fail.
__global__ void k_pb_dims1(const uint32_t *hoff, const uint32_t *coff, const uint64_t *snap_off,
                           uint32_t V, uint32_t NS, uint32_t BT, PbLimits l, PbDims *d) {
  const unsigned long long E = snap_off[V];
  const unsigned long long nht = ((unsigned long long)hoff[V] + PF_T - 1) / PF_T;
  const unsigned long long nct = ((unsigned long long)coff[V] + PF_T - 1) / PF_T;
  const unsigned long long NB = (nct + BT - 1) / BT;
  d->E = E;
  d->hend = nht * PF_T;
  d->cend = nct * PF_T;
  d->nht = (uint32_t)nht;
  d->nct = (uint32_t)nct;
  d->NB = (uint32_t)NB;
  d->table = (uint32_t)(NS * NB);
  if (!l.check) return;
  const bool bad = E + 2ull * V + 2ull * BT * PF_T >= 0x7FFFFFFFull ||
                   (unsigned long long)NS * NB >= (1ull << 29) || d->hend > l.hot ||
                   d->cend > l.cold || nht + 1 > l.htb || nct + 1 > l.ctb || nht + 1 > l.pqh ||
                   nct + 1 > l.pqc || NS * NB + 1 > l.cnt || NB + 1 > l.ngrp;
  if (bad && !d->fail) d->fail = 2;
}

// Long rows, expansion slots and image groups, once the table is scanned.
__global__ void k_pb_dims2(const uint32_t *lidxr, const uint32_t *cnt, const uint32_t *ngrp,
                           uint32_t V, uint32_t NS, PbLimits l, PbDims *d) {
  if (d->fail) return;
  d->n_long = lidxr[V];
  d->P1 = cnt[d->table];
  d->G = ngrp[d->NB];
  if (!l.check) return;
  const unsigned long long items = ((unsigned long long)d->P1 + PB_CH - 1) / PB_CH + NS + 1;
  // (no expansion items without cold source blocks)
  if (d->P1 + 8ull > l.lidx || d->P1 + 8ull > l.vals || d->G + 8ull > l.gaddr ||
      (NS && items > l.items))
    d->fail = 3;
}

// The result in vertex order from whichever buffer the last update wrote
// (its parity is known only on the device in a captured call).
__global__ void k_pb_unpermute(const float *x0, const float *x1, const PrCtl *ctl,
                               const PbDims *dims, const uint32_t *newid, uint32_t V,
                               float *out) {
  if (dims->fail) return;
  const float *x = ctl->iters & 1u ? x1 : x0;
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V;
       v += (uint64_t)gridDim.x * blockDim.x)
    out[v] = x[newid[v]];
}

gv_status pagerank_pb_prepare(const Store &in, const uint32_t *deg_out, PrWork &w, hipStream_t s,
                              bool capture) {
  const uint32_t V = in.V;
  w.pb_ready = 0;
  w.pf_ready = 0;
  if (!V) return GV_E_UNSUPPORTED;
  const uint32_t NS = V > PB_H ? (uint32_t)cdiv(V - PB_H, PB_SB) : 0u;
  if (NS > PB_MAX_NS) return GV_E_UNSUPPORTED;
  if (capture && (w.pb_res_V != V || !w.pb_dims || w.V < V || !w.ctl)) {
    set_error("pagerank capture: call gv_pagerank_reserve for this graph first");
    return GV_E_INVALID;
  }
  gv_status st = pr_reserve(w, V);
  if (st) return st;
  PbCaps &cp = w.pb_caps;
  if (!w.pb_dims) {
    ++w.gen;
    GV_HIP(hipMalloc(&w.pb_dims, sizeof(PbDims)));
    GV_HIP(hipHostMalloc(&w.h_pb_dims, sizeof(PbDims), hipHostMallocDefault));
  }
  PbDims *dims = w.pb_dims;
  k_pb_dims_reset<<<1, 1, 0, s>>>(dims);
  // hot entries per snapshot item, counted while the snapshot is written
  if (!capture && (st = pb_grow(w, (void **)&w.pb_hitem, cp.hitem, (pr_snapshot_icap(in) + 1) * 4)))
    return st;
  st = pr_snapshot(in, deg_out, true, w, s, 0, w.pb_hitem, PB_H, capture ? &dims->fail : nullptr);
  if (st) return st;
  // stream positions, expansion positions and the segment table are 32-bit
  const uint32_t BT = pb_bin_tiles(NS);
  const uint32_t bin = BT * PF_T;
  if (!capture && w.snap_E + 2ull * V + 2ull * bin >= 0x7FFFFFFFull) return GV_E_UNSUPPORTED;
  const uint32_t ni_cap = (uint32_t)w.snap_icap;
  const size_t vfloor = (size_t)(V + 1) * 4;
  if (!capture &&
      ((st = pb_grow(w, (void **)&w.pb_hoff, cp.hoff, vfloor)) ||
       (st = pb_grow(w, (void **)&w.pb_coff, cp.coff, vfloor)) ||
       (st = pb_grow(w, (void **)&w.pb_lidxr, cp.lidxr, vfloor)) ||
       (st = pb_grow(w, (void **)&w.pb_long, cp.lng, vfloor)) ||
       (st = pb_grow(w, (void **)&w.pb_Sh, cp.Sh, (size_t)V * 8)) ||
       (st = pb_grow(w, (void **)&w.pb_Sc, cp.Sc, (size_t)V * 8)) ||
       (st = pb_grow(w, (void **)&w.pb_partials, cp.partials, PB_PARTS * sizeof(double)))))
    return st;
  // hot counts per item scanned; row lengths of both streams, scanned
  scan_u32(w.pb_hitem, w.pb_hitem, ni_cap, w.snap_istart + V, w.snap_tmp, s);
  k_pb_rowlen<<<grid_for(V), 256, 0, s>>>(w.snap_istart, w.pb_hitem, w.rl_order, w.snap_off, V,
                                          w.pb_hoff, w.pb_coff);
  scan_u32(w.pb_hoff, w.pb_hoff, V, nullptr, w.snap_tmp, s);
  scan_u32(w.pb_coff, w.pb_coff, V, nullptr, w.snap_tmp, s);
  const PbLimits lim = pb_limits(w, capture);
  k_pb_dims1<<<1, 1, 0, s>>>(w.pb_hoff, w.pb_coff, w.snap_off, V, NS, BT, lim, dims);
  // launch sizes: the layout's own (read back) or, captured, the reserved ones
  uint64_t nht, nct, NB, table;
  if (capture) {
    nht = cp.htb / 4 - 1;
    nct = cp.ctb / 4 - 1;
    NB = cp.ngrp / 4 - 1;
    table = cp.cnt / 4 - 1;
  } else {
    GV_HIP(hipMemcpyAsync(w.h_pb_dims, dims, sizeof(PbDims), hipMemcpyDeviceToHost, s));
    GV_HIP(hipStreamSynchronize(s));
    const PbDims &h = *w.h_pb_dims;
    nht = h.nht;
    nct = h.nct;
    NB = h.NB;
    table = h.table;
    if (table >= (1ull << 29)) return GV_E_UNSUPPORTED;
    // the streams are sized for the snapshot buffer's capacity, so a caller
    // that presized it (the request loop) does not reallocate here either
    const uint64_t E = w.snap_E;
    const uint64_t ecap = std::max<uint64_t>(E, w.snap_min_cap) + V;
    const uint64_t tcap = cdiv(ecap, PF_T) + 2;
    const uint64_t bcap = cdiv(tcap, PB_MIN_BIN_TILES) + 1;
    const uint64_t hend = h.hend, cend = h.cend;
    if ((st = pb_grow(w, (void **)&w.pb_hot, cp.hot, hend * 2, tcap * PF_T * 2)) ||
        (st = pb_grow(w, (void **)&w.pb_cold, cp.cold, cend * 2, tcap * PF_T * 2)) ||
        (st = pb_grow(w, (void **)&w.pb_csrc, cp.csrc, cend * 4, tcap * PF_T * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_htb, cp.htb, (nht + 1) * 4, tcap * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_ctb, cp.ctb, (nct + 1) * 4, tcap * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_PQh, cp.PQh, (nht + 1) * 16, tcap * 16)) ||
        (st = pb_grow(w, (void **)&w.pb_PQc, cp.PQc, (nct + 1) * 16, tcap * 16)) ||
        (st = pb_grow(w, (void **)&w.pb_cnt, cp.cnt, (table + 1) * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_ngrp, cp.ngrp, ((size_t)NB + 1) * 4, (bcap + 1) * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_nit, cp.nit, ((size_t)NS + 1) * 4 + 8)) ||
        (st = pb_grow(w, (void **)&w.pb_tmp, cp.tmp,
                      std::max(scan_tmp_bytes(table + 1), scan_tmp_bytes((uint64_t)bcap + 1)))))
      return st;
  }
  // both streams
  k_pb_split<<<grid_for((uint64_t)ni_cap * WAVE, 256, 2048), 256, 0, s>>>(
      w.snap_istart, w.snap_ivert, w.snap_ipos, V, w.rl_newid, w.snap_off, w.snap_cols,
      w.pb_hitem, w.pb_hoff, w.pb_coff, w.pb_hot, w.pb_csrc, dims);
  k_pb_pads<<<grid_for(V), 256, 0, s>>>(w.snap_istart, w.pb_hitem, w.rl_order, w.snap_off, V,
                                        w.pb_hoff, w.pb_coff, w.pb_hot, w.pb_csrc, dims);
  k_pb_tail<<<4, 256, 0, s>>>(w.pb_hoff, w.pb_coff, V, dims, w.pb_hot, w.pb_csrc);
  if (nht) k_pb_tbase<<<grid_for(nht), 256, 0, s>>>(w.pb_hoff, V, dims, 0u, w.pb_htb);
  if (nct) k_pb_tbase<<<grid_for(nct), 256, 0, s>>>(w.pb_coff, V, dims, 1u, w.pb_ctb);
  k_pb_longflag<<<grid_for(V), 256, 0, s>>>(w.pb_hoff, w.pb_coff, V, w.pb_lidxr);
  scan_u32(w.pb_lidxr, w.pb_lidxr, V, nullptr, w.snap_tmp, s);
  k_pb_longfill<<<grid_for(V), 256, 0, s>>>(w.pb_lidxr, V, w.pb_long);
  // cold bins: padded (source block, bin) counts, expansion positions, image
  // group offsets
  if (NS) {
    k_pb_bincnt<<<(unsigned)NB, PB_BUILD_THREADS, 0, s>>>(w.pb_csrc, dims, NS, bin, w.pb_cnt,
                                                          w.pb_ngrp);
  } else {
    fill_u32(w.pb_ngrp, 0, NB, s);
  }
  scan_u32(w.pb_cnt, w.pb_cnt, table, capture ? &dims->table : nullptr, w.pb_tmp, s);
  scan_u32(w.pb_ngrp, w.pb_ngrp, NB, capture ? &dims->NB : nullptr, w.pb_tmp, s);
  k_pb_dims2<<<1, 1, 0, s>>>(w.pb_lidxr, w.pb_cnt, w.pb_ngrp, V, NS, lim, dims);
  uint64_t pcap = 0;
  if (!capture) {
    GV_HIP(hipMemcpyAsync(w.h_pb_dims, dims, sizeof(PbDims), hipMemcpyDeviceToHost, s));
    GV_HIP(hipStreamSynchronize(s));
    const uint64_t P1 = w.h_pb_dims->P1, G = w.h_pb_dims->G;
    const uint64_t ecap = std::max<uint64_t>(w.snap_E, w.snap_min_cap) + V;
    pcap = std::max<uint64_t>(P1, ecap + ecap / 8) + 8;
    if ((st = pb_grow(w, (void **)&w.pb_lidx, cp.lidx, (P1 + 8) * 2, pcap * 2)) ||
        (st = pb_grow(w, (void **)&w.pb_vals, cp.vals, (P1 + 8) * 4, pcap * 4)) ||
        (st = pb_grow(w, (void **)&w.pb_gaddr, cp.gaddr, (G + 8) * 4, pcap)))
      return st;
  }
  if (NS) {
    k_pb_scatter<<<(unsigned)NB, PB_BUILD_THREADS, 0, s>>>(w.pb_csrc, dims, NS, bin, w.pb_cnt,
                                                           w.pb_ngrp, w.pb_gaddr, w.pb_lidx,
                                                           w.pb_cold);
    k_pb_itemcnt<<<grid_for(NS), 256, 0, s>>>(w.pb_cnt, NS, dims, w.pb_nit);
    scan_u32(w.pb_nit, w.pb_nit, NS, nullptr, w.pb_tmp, s);
    if (!capture) {
      const uint64_t icap = cdiv(std::max<uint64_t>(pcap, 1), PB_CH) + NS + 1;
      if ((st = pb_grow(w, (void **)&w.pb_items, cp.items, icap * 16))) return st;
    }
    k_pb_itemfill<<<grid_for(NS), 256, 0, s>>>(w.pb_cnt, NS, dims, w.pb_nit, w.pb_items);
  } else {
    // every cold entry is padding: the stream's entries all take index PB_CPAD
    k_pb_scatter<<<(unsigned)NB, PB_BUILD_THREADS, 0, s>>>(w.pb_csrc, dims, 0, bin, w.pb_cnt,
                                                           w.pb_ngrp, w.pb_gaddr, w.pb_lidx,
                                                           w.pb_cold);
    fill_u32(w.pb_nit, 0, 1, s);
  }
  GV_HIP(hipGetLastError());
  w.pb_NS = NS;
  w.pb_NB = (uint32_t)NB;
  w.pb_bin_tiles = BT;
  w.pb_nht = (uint32_t)nht;
  w.pb_nct = (uint32_t)nct;
  w.pb_n_long = capture ? 0u : w.h_pb_dims->n_long;
  w.pb_prepV = V;
  w.pb_E = capture ? 0 : w.snap_E;
  w.pb_ready = 1;
  return GV_OK;
}

gv_status pagerank_pb_reserve(const Store &in, const uint32_t *deg_out, PrWork &w,
                              uint64_t max_edges, hipStream_t s) {
  const uint32_t V = in.V;
  if (!V) return GV_E_UNSUPPORTED;
  const uint32_t NS = V > PB_H ? (uint32_t)cdiv(V - PB_H, PB_SB) : 0u;
  const uint32_t BT = pb_bin_tiles(NS);
  if (NS > PB_MAX_NS || max_edges + 2ull * V + 2ull * BT * PF_T >= 0x7FFFFFFFull) {
    set_error("gv_pagerank_reserve: the default layout's 32-bit positions do not cover this size");
    return GV_E_UNSUPPORTED;
  }
  w.pb_res_V = 0;
  if (w.snap_min_cap < max_edges) w.snap_min_cap = max_edges;
  // a plain preparation of the current graph sizes every buffer with the
  // headroom of snap_min_cap
  gv_status st = pagerank_pb_prepare(in, deg_out, w, s, false);
  if (st) return st;
  if (w.snap_cap < w.snap_min_cap + PF_T) {
    ++w.gen;
    hipFree(w.snap_cols);
    w.snap_cols = nullptr;
    w.snap_cap = 0;
    const uint64_t cap = w.snap_min_cap + w.snap_min_cap / 8 + PF_T;
    GV_HIP(hipMalloc(&w.snap_cols, cap * sizeof(uint32_t)));
    w.snap_cap = cap;
  }
  const uint64_t ecap = std::max<uint64_t>(w.snap_E, w.snap_min_cap) + V;
  const uint64_t tcap = cdiv(ecap, PF_T) + 2;
  const uint64_t bcap = cdiv(tcap, PB_MIN_BIN_TILES) + 1;
  const uint64_t table = (uint64_t)NS * bcap;
  if (table >= (1ull << 29)) {
    set_error("gv_pagerank_reserve: too many (source block, bin) pairs for the default layout");
    return GV_E_UNSUPPORTED;
  }
  PbCaps &cp = w.pb_caps;
  if ((st = pb_grow(w, (void **)&w.pb_cnt, cp.cnt, (table + 1) * 4)) ||
      (st = pb_grow(w, (void **)&w.pb_tmp, cp.tmp,
                    std::max(scan_tmp_bytes(table + 1), scan_tmp_bytes(bcap + 1)))))
    return st;
  GV_HIP(hipStreamSynchronize(s));
  w.pb_res_V = V;
  w.pb_res_E = max_edges;
  return GV_OK;
}

gv_status pagerank_pb_last(PrWork &w, int *iters, double *residual, hipStream_t s) {
  if (!w.pb_dims || !w.ctl) {
    set_error("no PageRank call has run on this handle");
    return GV_E_INVALID;
  }
  GV_HIP(hipMemcpyAsync(w.h_pb_dims, w.pb_dims, sizeof(PbDims), hipMemcpyDeviceToHost, s));
  GV_HIP(hipMemcpyAsync(w.h_ctl, w.ctl, sizeof(PrCtl), hipMemcpyDeviceToHost, s));
  GV_HIP(hipStreamSynchronize(s));
  if (w.h_pb_dims->fail) {
    set_error("the last PageRank call's snapshot exceeded the buffers reserved for it "
              "(gv_pagerank_reserve); nothing was computed");
    return GV_E_INVALID;
  }
  if (iters) *iters = (int)w.h_ctl->iters;
  if (residual) *residual = w.h_ctl->residual;
  return GV_OK;
}

gv_status pagerank_pb_iterate(const PrOpts &o, PrWork &w, float *out_dev, int *iters,
                              double *residual, hipStream_t s, double *iter_ms, bool capture) {
  if (iters) *iters = 0;
  if (residual) *residual = 0.0;
  if (!w.pb_ready || o.max_iter <= 0) {
    set_error("pagerank: no prepared snapshot, or max_iter <= 0");
    return GV_E_INVALID;
  }
  if (capture && (iters || residual || iter_ms || o.yield)) {
    set_error("pagerank capture: no statistics or yield hook inside a capture");
    return GV_E_INVALID;
  }
  const uint32_t V = w.pb_prepV;
  const uint32_t *dout = w.rl_dout;  // out-degrees in label order
  const float x0 = 1.0f / (float)V;
  pr_ctl_reset(w.ctl, s);
  pr_init_vectors(w.x[0], w.contrib[0], dout, V, x0, o.init, w.rl_order, s);

  PbArgs a;
  a.hot = reinterpret_cast<const uint4 *>(w.pb_hot);
  a.cold = reinterpret_cast<const uint4 *>(w.pb_cold);
  a.htbase = w.pb_htb;
  a.ctbase = w.pb_ctb;
  const uint64_t tcap_h = w.pb_caps.PQh / 16, tcap_c = w.pb_caps.PQc / 16;
  a.Ph = w.pb_PQh;
  a.Qh = w.pb_PQh + tcap_h;
  a.Pc = w.pb_PQc;
  a.Qc = w.pb_PQc + tcap_c;
  a.Sh = w.pb_Sh;
  a.Sc = w.pb_Sc;
  a.hoff = w.pb_hoff;
  a.coff = w.pb_coff;
  a.gaddr = w.pb_gaddr;
  a.gbase = w.pb_ngrp;
  a.lidx = w.pb_lidx;
  a.vals = w.pb_vals;
  a.items = w.pb_items;
  a.n_items = w.pb_nit + w.pb_NS;
  a.deg_out = dout;
  a.longrows = w.pb_long;
  a.partials = w.pb_partials;
  a.ctl = w.ctl;
  a.dims = w.pb_dims;
  a.V = V;
  a.nh = V < PB_H ? V : PB_H;
  // tile, bin and long-row counts are read from a.dims by the kernels
  a.n_htiles = a.n_ctiles = a.n_hot_units = a.n_bins = a.n_long = 0;
  a.bin_tiles = w.pb_bin_tiles;
  a.compensated = o.compensated ? 1u : 0u;
  a.max_iter = (uint32_t)o.max_iter;
  a.alpha = o.alpha;
  a.base = (1.0f - o.alpha) / (float)V;
  a.eps = o.eps;
  // two tile workgroups per CU (80 KiB of LDS each); the request loop's
  // CU share scales the grid through pf_grid (workgroups per CU pair)
  const uint32_t cus = w.pf_grid ? w.pf_grid : (uint32_t)pb_device_cus();
  const uint32_t grid = 2 * cus;
  const uint32_t egrid = 2 * cus + 1;
  a.x = w.x[0];
  a.cb = w.contrib[0];
  if (w.pb_NS) k_pb_expand<<<egrid - 1, PB_BUILD_THREADS, 0, s>>>(a, 0u);
  auto launch_iter = [&](int t) {
    a.x = w.x[t & 1];
    a.cb = w.contrib[t & 1];
    a.xn = w.x[(t + 1) & 1];
    a.cn = w.contrib[(t + 1) & 1];
    if (a.compensated) k_pb_tiles<true><<<grid, PB_THREADS, 0, s>>>(a);
    else k_pb_tiles<false><<<grid, PB_THREADS, 0, s>>>(a);
    k_pb_rows<<<PB_PARTS, PB_ROW_THREADS, 0, s>>>(a);
    // the residual, and the expansion of the contributions just written
    PbArgs ae = a;
    ae.cb = a.cn;
    k_pb_expand<<<w.pb_NS ? egrid : 1, PB_BUILD_THREADS, 0, s>>>(ae, 1u);
  };
  if (capture) {
    // every launch is recorded; those after convergence exit at once
    for (int t = 0; t < o.max_iter; ++t) launch_iter(t);
    k_pb_unpermute<<<grid_for(V), 256, 0, s>>>(w.x[0], w.x[1], w.ctl, w.pb_dims, w.rl_newid, V,
                                               out_dev);
    GV_HIP(hipGetLastError());
    return GV_OK;
  }
  if (iter_ms) {
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
  // Iterations go out in groups (one at a time with a yield hook); after each
  // group the done flag is copied to a host slot and the host reads the
  // previous group's copy, so the device always has a group queued.
  constexpr int PB_GROUP = 8;
  const int group = o.yield ? 1 : PB_GROUP;
  int pending = -1;
  while (launched < o.max_iter) {
    const int t = launched;
    if (o.yield) o.yield(o.yield_arg);
    launch_iter(t);
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
    float ms = 0.f;
    const int last = (int)it < launched ? (int)it : launched;
    if (last > 0) GV_HIP(hipEventElapsedTime(&ms, w.ev[0], w.ev[last]));
    *iter_ms = ms;
  }
  return GV_OK;
}

}  // namespace gv
