// Exact quadtree build on gfx950: the same cells, bounds, leaf chains and
// preorder numbering as the reference's O(n) build (src/core/tree.zig), which
// is itself d3-quadtree's addAll tree (tree.zig:1-22).
//
// Pipeline (all on one stream, one host sync to size the cell buffers):
//  1. k_box: bounding box of the non-NaN points with addAll's comparisons
//     (ties between -0 and +0 go to the lower index, as a serial scan keeps
//     the first), then d3's cover in the last block (tree.zig:148-187,
//     316-340).
//  2. k_keys: 32 levels of d3's midpoint test per point -> 64-bit Morton key
//     (tree.zig:252-284); NaN points compacted out by a scan (rest, in index
//     order).
//  3. Stable LSD radix sort of (key, node), 8 passes of 8 bits (sort.zig):
//     per pass a tile histogram, one scan block per digit, and a stable
//     ballot-ranked scatter.
//  4. Key groups (runs of equal keys). A group whose points all coincide is
//     one leaf at depth L = 1 + max(divergence depth with the previous and the
//     next group); the cells on its path that start at the group are the
//     depths (dleft, L]. A group with distinct positions below key resolution
//     ("mixed") is internal at depth 32 and continues with d3's midpoint
//     split, serially (tree.zig:449-479). Preorder index of a cell =
//     (cells of earlier groups) + (depth - dleft - 1): preorder is the order
//     of (first point, depth).
//  5. k_assign names each cell's (group, depth) at its preorder index; then
//     k_emit, one thread per cell, writes the cell whole (bounds by the
//     midpoint walk along the key, children by binary search of the digit
//     boundaries inside the cell's key range), so no two threads write one
//     cell and no thread walks a long chain of searches.
#include <cstring>
#include <utility>

#include "common.h"

namespace d3f {
namespace {

constexpr int BOX_THREADS = 256;
constexpr int BOX_BLOCKS = 1024;
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;
constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per tile
static_assert(RS_THREADS == 256, "one thread per radix digit");
constexpr int EW_THREADS = 256;                 // element-wise kernels

struct BoxPartial {
  double x0, y0, x1, y1;
  uint32_t ix0, iy0, ix1, iy1;
  uint32_t m, pad;
};

// First-index-wins min / max: the serial scan of tree.zig:237-244 keeps the
// first of equal values (it only replaces on a strict < or >).
__device__ inline void take_min(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void take_max(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}

__device__ inline void combine(BoxPartial& a, const BoxPartial& b) {
  take_min(a.x0, a.ix0, b.x0, b.ix0);
  take_min(a.y0, a.iy0, b.y0, b.iy0);
  take_max(a.x1, a.ix1, b.x1, b.ix1);
  take_max(a.y1, a.iy1, b.y1, b.iy1);
  a.m += b.m;
}

// d3-quadtree cover (tree.zig:148-187), one thread.
__device__ void cover(double e[4], double x, double y) {
  if (x != x || y != y) return;
  double x0 = e[0], y0 = e[1], x1 = e[2], y1 = e[3];
  if (x0 != x0) {
    x0 = floor(x);
    x1 = x0 + 1;
    y0 = floor(y);
    y1 = y0 + 1;
  } else {
    double z = x1 - x0;
    if (z == 0 || z != z) z = 1;
    uint32_t guard = 0;
    while ((x0 > x || x >= x1 || y0 > y || y >= y1) && guard < MAX_DEPTH) {
      unsigned i = (unsigned)(y < y0) << 1 | (unsigned)(x < x0);
      z *= 2;
      switch (i) {
        case 0: x1 = x0 + z; y1 = y0 + z; break;
        case 1: x0 = x1 - z; y1 = y0 + z; break;
        case 2: x1 = x0 + z; y0 = y1 - z; break;
        default: x0 = x1 - z; y0 = y1 - z; break;
      }
      ++guard;
    }
  }
  e[0] = x0; e[1] = y0; e[2] = x1; e[3] = y1;
}

__global__ __launch_bounds__(BOX_THREADS) void k_box(const double* __restrict__ xs, const double* __restrict__ ys,
                                                     uint32_t n, BoxPartial* __restrict__ partials) {
  const double inf = __builtin_huge_val();
  BoxPartial p{inf, inf, -inf, -inf, NONE, NONE, NONE, NONE, 0, 0};
  for (uint32_t i = blockIdx.x * BOX_THREADS + threadIdx.x; i < n; i += gridDim.x * BOX_THREADS) {
    double x = xs[i], y = ys[i];
    if (x != x || y != y) continue;
    take_min(p.x0, p.ix0, x, i);
    take_min(p.y0, p.iy0, y, i);
    take_max(p.x1, p.ix1, x, i);
    take_max(p.y1, p.iy1, y, i);
    p.m++;
  }
  __shared__ BoxPartial sh[BOX_THREADS];
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int s = BOX_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) combine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// The blocks' partials in block order, then d3's cover of the two corners;
// a kernel of its own, so the partials are visible without a fence per block.
__global__ __launch_bounds__(BOX_THREADS) void k_box_final(uint32_t n, uint32_t blocks,
                                                           const BoxPartial* __restrict__ partials,
                                                           TreeMeta* __restrict__ meta) {
  const double inf = __builtin_huge_val();
  __shared__ BoxPartial sh[BOX_THREADS];
  BoxPartial q{inf, inf, -inf, -inf, NONE, NONE, NONE, NONE, 0, 0};
  for (uint32_t b = threadIdx.x; b < blocks; b += BOX_THREADS) combine(q, partials[b]);
  sh[threadIdx.x] = q;
  __syncthreads();
  for (int s = BOX_THREADS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) combine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    BoxPartial r = sh[0];
    const double nan = __builtin_nan("");
    double e[4] = {nan, nan, nan, nan};
    meta->n = n;
    meta->m = r.m;
    meta->len = 0;
    meta->depth = 0;
    meta->mixed = 0;
    meta->overflow = 0;
    meta->ncut = 0;
    meta->ntop = 0;
    meta->topdepth = 0;
    meta->long_group = 0;
    meta->stop = 0;
    if (r.x0 > r.x1 || r.y0 > r.y1) {  // no finite point (tree.zig:338)
      meta->m = 0;
      meta->valid = 0;
    } else {
      cover(e, r.x0, r.y0);
      cover(e, r.x1, r.y1);
      meta->valid = 1;
    }
    for (int k = 0; k < 4; ++k) meta->extent[k] = e[k];
    meta->ticket = 0;
  }
}

// d3's 32-level midpoint descent (tree.zig:265-279).
__device__ inline uint64_t exact_key(double x, double y, const double* e) {
  double x0 = e[0], y0 = e[1], x1 = e[2], y1 = e[3];
  uint64_t key = 0;
#pragma unroll 4
  for (int lev = 0; lev < KEY_LEVELS; ++lev) {
    double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
    bool right = x >= xm;
    if (right) x0 = xm; else x1 = xm;
    bool bottom = y >= ym;
    if (bottom) y0 = ym; else y1 = ym;
    key = key << 2 | (uint64_t)bottom << 1 | (uint64_t)right;
  }
  return key;
}

__global__ __launch_bounds__(EW_THREADS) void k_keys(const double* __restrict__ xs, const double* __restrict__ ys,
                                                     uint32_t n, const TreeMeta* __restrict__ meta,
                                                     uint64_t* __restrict__ keys, uint32_t* __restrict__ perm,
                                                     uint64_t* __restrict__ keys_all, uint32_t* __restrict__ flags) {
  uint32_t i = blockIdx.x * EW_THREADS + threadIdx.x;
  if (i >= n || !meta->valid) return;
  double x = xs[i], y = ys[i];
  bool valid = !(x != x || y != y);
  uint64_t key = 0;
  if (valid) {
    double e[4] = {meta->extent[0], meta->extent[1], meta->extent[2], meta->extent[3]};
    key = exact_key(x, y, e);
  }
  if (meta->m == n) {
    keys[i] = key;
    perm[i] = i;
  } else {
    keys_all[i] = key;
    flags[i] = valid ? 1u : 0u;
  }
}

__global__ __launch_bounds__(EW_THREADS) void k_compact(uint32_t n, const TreeMeta* __restrict__ meta,
                                                        const uint64_t* __restrict__ keys_all,
                                                        const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ offs,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ perm,
                                                        uint32_t* __restrict__ rest) {
  uint32_t i = blockIdx.x * EW_THREADS + threadIdx.x;
  if (i >= n || !meta->valid || meta->m == n) return;
  uint32_t o = offs[i];
  if (flags[i]) {
    perm[o] = i;
    keys[o] = keys_all[i];
  } else {
    rest[i - o] = i;
  }
}

// All nodes NaN: rest = 0..n-1 (the tree is empty, tree.zig:262-266).
__global__ __launch_bounds__(EW_THREADS) void k_rest_all(uint32_t n, const TreeMeta* __restrict__ meta,
                                                         uint32_t* __restrict__ rest) {
  uint32_t i = blockIdx.x * EW_THREADS + threadIdx.x;
  if (i >= n || meta->valid) return;
  rest[i] = i;
}

// ---------------------------------------------------------------- scan
// Exclusive scan of u32 over [0, *count) in two kernels: the sum of every
// tile of SCAN_TILE items, then each tile adds the sums of the tiles before it
// and scans itself. `skip` (when not null) makes every kernel return at once
// when skip[0] == skip[1].

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
  constexpr int NW = SCAN_THREADS / 64;
  __shared__ uint32_t wsum[NW];
  __shared__ uint32_t tot;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < NW ? wsum[lane] : 0;
    uint32_t winc = w;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      uint32_t t = __shfl_up(winc, d, 64);
      if (lane >= d) winc += t;
    }
    if (lane < NW) wsum[lane] = winc - w;
    if (lane == NW - 1) tot = winc;
  }
  __syncthreads();
  const uint32_t res = wsum[wave] + inc - v;
  *total = tot;
  __syncthreads();
  return res;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                              const uint32_t* __restrict__ count,
                                                              uint32_t* __restrict__ partials,
                                                              const uint32_t* __restrict__ skip) {
  if (skip && skip[0] == skip[1]) return;
  const uint32_t m = *count;
  const uint32_t base = blockIdx.x * SCAN_TILE;
  if (base >= m) return;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint32_t i = base + k * SCAN_THREADS + threadIdx.x;
    if (i < m) s += in[i];
  }
  uint32_t total;
  block_exclusive_scan(s, &total);
  if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out,
                                                             const uint32_t* __restrict__ count,
                                                             const uint32_t* __restrict__ partials,
                                                             uint32_t* __restrict__ total_out,
                                                             const uint32_t* __restrict__ skip) {
  if (skip && skip[0] == skip[1]) return;
  const uint32_t m = *count;
  const uint32_t base = blockIdx.x * SCAN_TILE;
  if (base >= m) {
    if (m == 0 && blockIdx.x == 0 && threadIdx.x == 0 && total_out) *total_out = 0;
    return;
  }
  // The tile's offset: the sums of the tiles before it (a few thousand words
  // at most), added by this block itself.
  uint32_t pre = 0;
  for (uint32_t j = threadIdx.x; j < blockIdx.x; j += SCAN_THREADS) pre += partials[j];
  uint32_t offset;
  block_exclusive_scan(pre, &offset);
  // Thread t owns items base + t*ITEMS .. +ITEMS (blocked, so order is kept).
  uint32_t v[SCAN_ITEMS];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint32_t i = base + threadIdx.x * SCAN_ITEMS + k;
    v[k] = i < m ? in[i] : 0;
    s += v[k];
  }
  uint32_t total;
  uint32_t ex = block_exclusive_scan(s, &total) + offset;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    uint32_t i = base + threadIdx.x * SCAN_ITEMS + k;
    if (i < m) out[i] = ex;
    if (i == m - 1 && total_out) *total_out = ex + v[k];
    ex += v[k];
  }
}

hipError_t scan_exclusive(const uint32_t* in, uint32_t* out, uint32_t n_max, const uint32_t* count_dev,
                          uint32_t* partials, uint32_t* total_out, const uint32_t* skip, hipStream_t stream) {
  uint32_t nb = ceil_div(n_max > 0 ? n_max : 1, SCAN_TILE);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, stream, in, count_dev, partials, skip);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(SCAN_THREADS), 0, stream, in, out, count_dev, partials,
                     total_out, skip);
  return hipGetLastError();
}

// ---------------------------------------------------------------- radix sort
// Stable LSD radix sort of (u64 key, u32 value) by 8-bit digits (sort.zig:55-91
// is a stable LSD radix sort; a stable sort of the same keys has one result).

// Clears `words` counters (a kernel rather than a memset, so that a captured
// graph of the sort holds no memset node).
__global__ __launch_bounds__(256) void k_zero_words(uint32_t* __restrict__ p, uint32_t words) {
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < words; i += gridDim.x * 256) p[i] = 0;
}

__global__ __launch_bounds__(RS_THREADS) void k_rs_hist(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ count,
                                                        uint32_t* __restrict__ hist, uint32_t ntiles, int shift) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t m = *count;
  const uint32_t base = blockIdx.x * RS_TILE;
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint32_t k = base + j * RS_THREADS + threadIdx.x;
    if (k < m) atomicAdd(&h[(keys[k] >> shift) & 255], 1u);
  }
  __syncthreads();
  hist[threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Digit d's tile offsets within the digit (exclusive scan of its tile
// counts) and the digit's total in glob[d], one block per digit (the digit's
// first slot is added by the scatter: no atomics on the totals).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scan(uint32_t* __restrict__ hist, uint32_t* __restrict__ glob,
                                                        uint32_t ntiles) {
  __shared__ uint32_t wsum[RS_THREADS / 64];
  __shared__ uint32_t carry;
  const uint32_t d = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  uint32_t* row = hist + (size_t)d * ntiles;
  for (uint32_t b0 = 0; b0 < ntiles; b0 += RS_THREADS) {
    const uint32_t i = b0 + threadIdx.x;
    const uint32_t x = i < ntiles ? row[i] : 0;
    uint32_t inc = x;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
      uint32_t t = __shfl_up(inc, k, 64);
      if (lane >= k) inc += t;
    }
    __syncthreads();
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t before = carry;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    if (i < ntiles) row[i] = before + inc - x;
    __syncthreads();
    if (threadIdx.x == RS_THREADS - 1) carry = before + inc;
    __syncthreads();
  }
  if (threadIdx.x == 0) glob[d] = carry;
}

// Each wave ranks its RS_TILE / 4 consecutive keys in RS_ITEMS slots of 64
// with ballot matching (stable: slot order, then lane order), the waves are
// stacked in order, and the tile is placed in LDS in digit order;
// consecutive threads then write consecutive keys of one digit's run to its
// scanned offset.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint64_t* __restrict__ kin,
                                                           const uint32_t* __restrict__ vin,
                                                           uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           const uint32_t* __restrict__ hist,
                                                           const uint32_t* __restrict__ glob,
                                                           const uint32_t* __restrict__ count, uint32_t ntiles,
                                                           int shift) {
  __shared__ uint32_t dbase[256];       // first output slot of each digit
  __shared__ uint32_t wcount[RS_THREADS / 64][256];
  __shared__ uint32_t wbase[RS_THREADS / 64][256];
  __shared__ uint32_t tstart[256];      // first slot of each digit in the tile
  __shared__ uint32_t wsum[RS_THREADS / 64], gsum[RS_THREADS / 64];
  __shared__ uint64_t skey[RS_TILE];    // the tile in digit order
  __shared__ uint32_t sval[RS_TILE];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = threadIdx.x; d < 256 * (RS_THREADS / 64); d += RS_THREADS) (&wcount[0][0])[d] = 0;
  __syncthreads();
  const uint32_t m = *count;
  const uint32_t base = blockIdx.x * RS_TILE + wave * (RS_TILE / (RS_THREADS / 64));
  const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  uint64_t key[RS_ITEMS];
  uint32_t val[RS_ITEMS], rank[RS_ITEMS];
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint32_t k = base + j * 64 + lane;
    bool valid = k < m;
    key[j] = valid ? kin[k] : 0;
    val[j] = valid ? vin[k] : 0;
    uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      bool bit = (d >> b) & 1u;
      uint64_t B = __ballot(bit);
      peers &= bit ? B : ~B;
    }
    uint32_t r = __popcll(peers & lt);
    uint32_t old = 0;
    if (valid) old = wcount[wave][d];
    __builtin_amdgcn_wave_barrier();  // every lane's read precedes the leaders' write
    if (valid && r == 0) wcount[wave][d] = old + (uint32_t)__popcll(peers);
    __builtin_amdgcn_wave_barrier();
    rank[j] = old + r;
  }
  __syncthreads();
  {
    // thread d: the waves' bases within digit d, then the digit's first slot
    // in the tile (exclusive scan of the tile's digit counts)
    uint32_t b = 0;
    const uint32_t d = threadIdx.x;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; ++w) {
      wbase[w][d] = b;
      b += wcount[w][d];
    }
    // two exclusive scans over the digits: the tile's counts (each digit's
    // first slot in the tile) and the digit totals (its first output slot)
    const uint32_t gt = glob[d];
    uint32_t inc = b, ginc = gt;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
      const uint32_t t = __shfl_up(inc, k, 64), gu = __shfl_up(ginc, k, 64);
      if (lane >= k) {
        inc += t;
        ginc += gu;
      }
    }
    if (lane == 63) {
      wsum[wave] = inc;
      gsum[wave] = ginc;
    }
    __syncthreads();
    uint32_t before = 0, gbefore = 0;
#pragma unroll
    for (int w = 0; w < RS_THREADS / 64; ++w)
      if (w < wave) {
        before += wsum[w];
        gbefore += gsum[w];
      }
    tstart[d] = before + inc - b;
    dbase[d] = gbefore + ginc - gt;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    uint32_t k = base + j * 64 + lane;
    if (k < m) {
      uint32_t d = (uint32_t)(key[j] >> shift) & 255u;
      const uint32_t slot = tstart[d] + wbase[wave][d] + rank[j];
      skey[slot] = key[j];
      sval[slot] = val[j];
    }
  }
  __syncthreads();
  const uint32_t tile0 = blockIdx.x * RS_TILE;  // tiles past the count hold no key
  const uint32_t tile_n = tile0 >= m ? 0 : (m - tile0 < (uint32_t)RS_TILE ? m - tile0 : (uint32_t)RS_TILE);
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    const uint32_t t = j * RS_THREADS + threadIdx.x;
    if (t < tile_n) {
      const uint64_t kk = skey[t];
      const uint32_t d = (uint32_t)(kk >> shift) & 255u;
      const uint32_t pos = dbase[d] + hist[d * ntiles + blockIdx.x] + (t - tstart[d]);
      kout[pos] = kk;
      vout[pos] = sval[t];
    }
  }
}

// ---------------------------------------------------------------- groups

__device__ inline int div_depth(uint64_t a, uint64_t b) {  // shared leading digits, 32 when equal
  uint64_t x = a ^ b;
  return x == 0 ? KEY_LEVELS : (int)(__clzll(x) >> 1);
}

// First p >= from with keys[p] != key (keys sorted, keys[from-1] == key).
__device__ inline uint32_t group_end(const uint64_t* keys, uint32_t m, uint32_t from, uint64_t key) {
  uint32_t lo = from, step = 1;
  while (true) {
    uint32_t q = lo + step - 1;
    if (q >= m || keys[q] != key) break;
    lo = q + 1;
    step <<= 1;
  }
  uint32_t hi = lo + step - 1 < m ? lo + step - 1 : m;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] != key) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// First p >= from with keys[p] >> shift != prefix (0 < shift < 64).
__device__ inline uint32_t prefix_end(const uint64_t* keys, uint32_t m, uint32_t from, uint64_t prefix, int shift) {
  uint32_t lo = from, step = 1;
  while (true) {
    uint32_t q = lo + step - 1;
    if (q >= m || (keys[q] >> shift) != prefix) break;
    lo = q + 1;
    step <<= 1;
  }
  uint32_t hi = lo + step - 1 < m ? lo + step - 1 : m;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if ((keys[mid] >> shift) != prefix) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// First p in [lo, hi) whose digit at `dsh` is >= q (digits are sorted there).
__device__ inline uint32_t digit_bound(const uint64_t* keys, uint32_t lo, uint32_t hi, int dsh, uint32_t q) {
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (((uint32_t)(keys[mid] >> dsh) & 3u) >= q) hi = mid; else lo = mid + 1;
  }
  return lo;
}

// First position of the run of equal keys that contains k.
__device__ inline uint32_t group_start(const uint64_t* keys, uint32_t k) {
  const uint64_t key = keys[k];
  uint32_t lo = 0, hi = k;
  while (lo < hi) {
    uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Marks group starts and groups whose points are not all coincident.
// The sort of the 64-bit Morton keys runs its radix passes on the high 32
// bits only; keys whose high halves are equal (points within 2^-16 of the
// extent of each other, few in practice) are then ordered by their low halves
// with a stable insertion sort, one thread per run. A run longer than
// LOW_SORT_MAX raises meta->long_group and the host sorts all 64 bits again.
// The result equals the stable sort by the whole key either way.
constexpr uint32_t LOW_SORT_MAX = 256;
__global__ __launch_bounds__(EW_THREADS) void k_sort_low_halves(uint64_t* __restrict__ keys,
                                                                uint32_t* __restrict__ perm,
                                                                TreeMeta* __restrict__ meta) {
  const uint32_t m = meta->m;
  const uint32_t p = blockIdx.x * EW_THREADS + threadIdx.x;
  if (p >= m) return;
  const uint32_t hi = (uint32_t)(keys[p] >> 32);
  if (p > 0 && (uint32_t)(keys[p - 1] >> 32) == hi) return;  // not the first of its run
  uint32_t q = p + 1;
  while (q < m && (uint32_t)(keys[q] >> 32) == hi) {
    if (q - p >= LOW_SORT_MAX) {
      atomicOr(&meta->long_group, 1u);
      return;
    }
    ++q;
  }
  for (uint32_t k = p + 1; k < q; ++k) {
    const uint64_t key = keys[k];
    const uint32_t v = perm[k];
    uint32_t h = k;
    while (h > p && keys[h - 1] > key) {
      keys[h] = keys[h - 1];
      perm[h] = perm[h - 1];
      --h;
    }
    keys[h] = key;
    perm[h] = v;
  }
}

__global__ __launch_bounds__(EW_THREADS) void k_group_a(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ perm,
                                                        const double* __restrict__ xs, const double* __restrict__ ys,
                                                        const TreeMeta* __restrict__ meta,
                                                        uint32_t* __restrict__ start, uint32_t* __restrict__ mixed) {
  uint32_t k = blockIdx.x * EW_THREADS + threadIdx.x;
  if (k >= meta->m) return;
  uint64_t key = keys[k];
  bool s = k == 0 || keys[k - 1] != key;
  start[k] = s ? 1u : 0u;
  if (!s) {
    uint32_t p = perm[k], q = perm[k - 1];
    if (xs[p] != xs[q] || ys[p] != ys[q]) atomicOr(&mixed[group_start(keys, k)], 1u);
  }
}

__global__ __launch_bounds__(EW_THREADS) void k_group_b(const uint64_t* __restrict__ keys,
                                                        const uint32_t* __restrict__ start,
                                                        const uint32_t* __restrict__ mixed,
                                                        TreeMeta* __restrict__ meta, uint32_t* __restrict__ count,
                                                        int32_t* __restrict__ dleft, uint32_t* __restrict__ gend,
                                                        uint32_t* __restrict__ leafdepth,
                                                        uint32_t* __restrict__ mixed_list) {
  const uint32_t m = meta->m;
  const uint32_t k = blockIdx.x * EW_THREADS + threadIdx.x;
  __shared__ uint32_t bmax;
  if (threadIdx.x == 0) bmax = 0;
  __syncthreads();
  if (k < m && !start[k]) count[k] = 0;
  if (k < m && start[k]) {
  const uint64_t key = keys[k];
  const uint32_t e = group_end(keys, m, k + 1, key);
  const int dl = k == 0 ? -1 : div_depth(keys[k - 1], key);
  const int dr = e == m ? -1 : div_depth(key, keys[e]);
  uint32_t L, c;
  if (mixed[k]) {
    L = KEY_LEVELS;
    c = (uint32_t)(KEY_LEVELS - dl);
    uint32_t slot = atomicAdd(&meta->mixed, 1u);
    mixed_list[slot] = k;
  } else {
    L = (uint32_t)(1 + (dl > dr ? dl : dr));
    c = (uint32_t)((int)L - dl);
  }
  count[k] = c;
  dleft[k] = dl;
  gend[k] = e;
  leafdepth[k] = L;
  atomicMax(&bmax, L);
  }
  __syncthreads();
  // One block's atomic on the shared word only when it could raise it:
  // atomics on one address serialize, and the depth stops rising after the
  // first blocks (a stale read only costs an atomic).
  if (threadIdx.x == 0 && bmax &&
      bmax > __hip_atomic_load(&meta->depth, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(&meta->depth, bmax);
}

// ---------------------------------------------------------------- below key resolution
// The serial split of tree.zig:572-647 for frames at depth >= 32: a stable
// partition by d3's midpoint test, one mixed group after another, on one
// thread. WRITE = false counts cells (and leaves perm partitioned, which is
// idempotent, tree.zig:446-448); WRITE = true emits them.

__device__ inline unsigned quadrant(double x, double y, double xm, double ym) {
  return (unsigned)(y >= ym) << 1 | (unsigned)(x >= xm);
}

template <bool WRITE>
__global__ void k_deep(const uint64_t* __restrict__ keys, uint32_t* __restrict__ perm,
                       uint32_t* __restrict__ perm_tmp, const double* __restrict__ xs,
                       const double* __restrict__ ys, TreeMeta* __restrict__ meta,
                       const uint32_t* __restrict__ mixed_list, const int32_t* __restrict__ dleft,
                       const uint32_t* __restrict__ gend, uint32_t* __restrict__ count,
                       const uint32_t* __restrict__ base, Frame* __restrict__ stack, Cell* __restrict__ cells,
                       uint32_t* __restrict__ cdepth) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t groups = meta->mixed;
  uint32_t maxd = 0;
  for (uint32_t g = 0; g < groups; ++g) {
    const uint32_t s = mixed_list[g];
    const uint32_t e = gend[s];
    const int dl = dleft[s];
    const uint64_t key = keys[s];
    double x0 = meta->extent[0], y0 = meta->extent[1], x1 = meta->extent[2], y1 = meta->extent[3];
    for (int d = 0; d < KEY_LEVELS; ++d) {
      uint32_t dg = (uint32_t)(key >> (62 - 2 * d)) & 3u;
      double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
      if (dg & 1) x0 = xm; else x1 = xm;
      if (dg & 2) y0 = ym; else y1 = ym;
    }
    uint32_t idx = WRITE ? base[s] + (uint32_t)(KEY_LEVELS - dl - 1) : 0;
    uint32_t cnt = 0, sp = 0;
    stack[sp++] = Frame{s, e, (uint32_t)KEY_LEVELS, NONE, 0, 0, x0, y0, x1, y1};
    while (sp > 0) {
      Frame f = stack[--sp];
      if (WRITE) {
        Cell c;
        c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
        c.first = f.lo;
        c.count = f.hi - f.lo;
        c.x0 = f.x0; c.y0 = f.y0; c.x1 = f.x1; c.y1 = f.y1;
        c.x = c.y = c.value = c.r = 0;
        cells[idx] = c;
        cdepth[idx] = f.depth;
        if (f.parent != NONE) cells[f.parent].child[f.slot] = idx;
      }
      const uint32_t me = idx++;
      ++cnt;
      if (f.depth > maxd) maxd = f.depth;
      if (f.depth >= MAX_DEPTH) continue;
      // allSame (tree.zig:437-444): the depth-32 frame holds distinct
      // positions by construction; deeper frames compare explicitly.
      if (f.depth > (uint32_t)KEY_LEVELS) {
        uint32_t p0 = perm[f.lo];
        bool same = true;
        for (uint32_t k = f.lo + 1; k < f.hi && same; ++k) {
          uint32_t p = perm[k];
          same = !(xs[p] != xs[p0] || ys[p] != ys[p0]);
        }
        if (same) continue;
      }
      // split (tree.zig:460-477): stable partition by the midpoint test.
      const double xm = (f.x0 + f.x1) / 2, ym = (f.y0 + f.y1) / 2;
      uint32_t counts[4] = {0, 0, 0, 0};
      for (uint32_t k = f.lo; k < f.hi; ++k) counts[quadrant(xs[perm[k]], ys[perm[k]], xm, ym)]++;
      uint32_t off[4], out[5];
      uint32_t sum = f.lo;
      for (int q = 0; q < 4; ++q) { off[q] = sum; out[q] = sum; sum += counts[q]; }
      out[4] = f.hi;
      for (uint32_t k = f.lo; k < f.hi; ++k) {
        uint32_t p = perm[k];
        perm_tmp[off[quadrant(xs[p], ys[p], xm, ym)]++] = p;
      }
      for (uint32_t k = f.lo; k < f.hi; ++k) perm[k] = perm_tmp[k];
      for (int q = 3; q >= 0; --q) {
        if (out[q] == out[q + 1]) continue;
        if (sp == DEEP_STACK) {
          meta->overflow = 1;
          return;
        }
        bool right = (q & 1) != 0, bottom = (q & 2) != 0;
        stack[sp++] = Frame{out[q], out[q + 1], f.depth + 1, me, (uint32_t)q, 0,
                            right ? xm : f.x0, bottom ? ym : f.y0, right ? f.x1 : xm, bottom ? f.y1 : ym};
      }
    }
    if (!WRITE) count[s] += cnt - 1;
  }
  if (maxd > meta->depth) meta->depth = maxd;
}

// Phase 1, one thread per key group: names the group's cells (depths
// (dleft, L], to depth 31 for a mixed group, whose depth-32 cell and subtree
// k_deep writes) by writing their group and depth at their preorder index.
__global__ __launch_bounds__(EW_THREADS) void k_assign(const uint32_t* __restrict__ start,
                                                       const uint32_t* __restrict__ mixed,
                                                       const TreeMeta* __restrict__ meta,
                                                       const int32_t* __restrict__ dleft,
                                                       const uint32_t* __restrict__ leafdepth,
                                                       const uint32_t* __restrict__ base,
                                                       uint32_t* __restrict__ cell_s, uint32_t* __restrict__ cell_d) {
  const uint32_t s = blockIdx.x * EW_THREADS + threadIdx.x;
  if (s >= meta->m || !start[s] || meta->stop) return;
  const int dl = dleft[s];
  const int lmax = mixed[s] ? KEY_LEVELS - 1 : (int)leafdepth[s];
  const uint32_t b = base[s];
  for (int d = dl + 1; d <= lmax; ++d) {
    const uint32_t idx = b + (uint32_t)(d - dl - 1);
    cell_s[idx] = s;
    cell_d[idx] = (uint32_t)d;
  }
}

// A slot in a list for each lane with `want`, one atomic per wave: the
// wave's count is added once and each lane takes its rank among the wanting
// lanes. Called by every active lane of the wave alike.
__device__ inline uint32_t wave_append(uint32_t* counter, bool want) {
  const uint64_t mask = __ballot(want);
  if (mask == 0) return 0;
  const uint32_t leader = (uint32_t)__ffsll((unsigned long long)mask) - 1;
  const uint32_t lane = __lane_id();
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = (uint32_t)__shfl((int)base, (int)leader);
  return base + (uint32_t)__popcll(mask & ((1ull << lane) - 1));
}

// Phase 2, one thread per cell: bounds by d3's midpoint walk along the key,
// the point range, and children 0..3 from the digit boundaries inside it.
// A cell's subtree ends where the group after its point range starts its
// cells (preorder is the order of (first point, depth), and that group's
// first cell is the one at the depth where it leaves the previous group's
// path): skip = base[hi] - idx, or len - idx for the last range. Cells whose
// subtree exceeds ACC_SUB go to `top`; their children within it, and a root
// within it, go to `cut` (the subtrees accumulate sums in LDS).
__global__ __launch_bounds__(EW_THREADS) void k_emit(const uint64_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ mixed,
                                                     TreeMeta* __restrict__ meta,
                                                     const uint32_t* __restrict__ gend,
                                                     const uint32_t* __restrict__ leafdepth,
                                                     const uint32_t* __restrict__ base,
                                                     const uint32_t* __restrict__ cell_s,
                                                     const uint32_t* __restrict__ cell_d, Cell* __restrict__ cells,
                                                     uint32_t* __restrict__ cdepth, uint32_t* __restrict__ cskip,
                                                     uint2* __restrict__ cut, uint32_t* __restrict__ top) {
  const uint32_t idx = blockIdx.x * EW_THREADS + threadIdx.x;
  if (idx >= meta->len || meta->stop) return;
  const uint32_t s = cell_s[idx];
  if (s == NONE) return;  // a cell below key resolution (k_deep)
  const int d = (int)cell_d[idx];
  const uint32_t m = meta->m;
  const uint64_t key = keys[s];
  const uint32_t e = gend[s];
  const bool internal = mixed[s] != 0 || d < (int)leafdepth[s];
  double x0 = meta->extent[0], y0 = meta->extent[1], x1 = meta->extent[2], y1 = meta->extent[3];
  for (int k = 0; k < d; ++k) {
    const uint32_t dg = (uint32_t)(key >> (62 - 2 * k)) & 3u;
    const double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
    if (dg & 1) x0 = xm; else x1 = xm;
    if (dg & 2) y0 = ym; else y1 = ym;
  }
  uint32_t hi;
  if (!internal) hi = e;
  else if (d == 0) hi = m;
  else hi = prefix_end(keys, m, e, key >> (64 - 2 * d), 64 - 2 * d);
  Cell c;
  c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
  c.first = s;
  c.count = hi - s;
  c.x0 = x0; c.y0 = y0; c.x1 = x1; c.y1 = y1;
  c.x = c.y = c.value = c.r = 0;
  const uint32_t len = meta->len;
  const uint32_t skip = (hi < m ? base[hi] : len) - idx;
  const bool is_top = skip > ACC_SUB;
  uint32_t child_skip[4] = {0, 0, 0, 0};
  if (internal) {
    const int dsh = 62 - 2 * d;
    uint32_t bnd[5];
    bnd[0] = s;
    bnd[4] = hi;
    for (uint32_t q = 1; q < 4; ++q) bnd[q] = digit_bound(keys, bnd[q - 1], hi, dsh, q);
    for (int q = 0; q < 4; ++q) {
      if (bnd[q] == bnd[q + 1]) continue;
      c.child[q] = bnd[q] == s ? idx + 1 : base[bnd[q]];
      child_skip[q] = (bnd[q + 1] < m ? base[bnd[q + 1]] : len) - c.child[q];
    }
  }
  cells[idx] = c;
  cdepth[idx] = (uint32_t)d;
  cskip[idx] = skip;
  // List appends, one atomic per wave and list.
  for (int q = 0; q < 4; ++q) {
    const bool want = is_top && c.child[q] != NONE && child_skip[q] <= ACC_SUB;
    const uint32_t slot = wave_append(&meta->ncut, want);
    if (want) cut[slot] = make_uint2(c.child[q], child_skip[q] | (uint32_t)(d + 1) << CUT_SIZE_BITS);
  }
  const bool small_root = d == 0 && !is_top;
  const uint32_t rslot = wave_append(&meta->ncut, small_root);
  if (small_root) cut[rslot] = make_uint2(idx, skip);  // the root, depth 0
  const uint32_t tslot = wave_append(&meta->ntop, is_top);
  if (is_top) top[tslot] = idx;
  if (__ballot(is_top)) {
    // The wave's deepest top cell, bit by bit (top cells lie above depth
    // KEY_LEVELS); the first active lane posts it.
    uint32_t dm = 0;
    for (int b = 5; b >= 0; --b) {
      const uint32_t cand = dm | (1u << b);
      if (__ballot(is_top && (uint32_t)d >= cand)) dm = cand;
    }
    if (__lane_id() == (uint32_t)__ffsll((unsigned long long)__ballot(true)) - 1) atomicMax(&meta->topdepth, dm);
  }
}

}  // namespace

// ---------------------------------------------------------------- host side

// Stable sort of (keys, vals)[0 .. *count) by the 8-bit digits first_digit ..
// first_digit + digits - 1. hist holds 256 x ceil(n_max / RS_TILE) counts,
// glob 256 x digits. After an odd number of passes the first n_max entries of
// the ping-pong buffers are copied back.
hipError_t sort_digits(uint32_t* hist, uint32_t* glob, uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp,
                       uint32_t* vals_tmp, uint32_t n_max, const uint32_t* count, int first_digit, int digits,
                       hipStream_t stream) {
  if (digits < 0 || digits > 8) HIPFAIL(hipErrorInvalidValue, "radix digits out of range");
  const uint32_t ntiles = ceil_div(n_max > 0 ? n_max : 1, RS_TILE);
  uint64_t *ka = keys, *kb = keys_tmp;
  uint32_t *va = vals, *vb = vals_tmp;
  for (int pass = 0; pass < digits; ++pass) {
    uint32_t* g = glob + 256 * pass;
    const int shift = 8 * (first_digit + pass);
    hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_THREADS), 0, stream, ka, count, hist, ntiles, shift);
    hipLaunchKernelGGL(k_rs_scan, dim3(256), dim3(RS_THREADS), 0, stream, hist, g, ntiles);
    hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_THREADS), 0, stream, ka, va, kb, vb, hist, g, count,
                       ntiles, shift);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (digits & 1) {
    HIPCHECK(hipMemcpyAsync(keys, ka, (size_t)n_max * sizeof(uint64_t), hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(vals, va, (size_t)n_max * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
  }
  return hipGetLastError();
}

hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t n_max, const uint32_t* count_dev, uint32_t* partials,
                    uint32_t* total_out, hipStream_t stream) {
  return scan_exclusive(in, out, n_max, count_dev, partials, total_out, nullptr, stream);
}

size_t sort_hist_words(uint32_t n_max) { return (size_t)256 * ceil_div(n_max > 0 ? n_max : 1, RS_TILE); }


template <class T>
static hipError_t grow(T*& p, size_t n) {
  if (p) HIPCHECK(hipFree(p));
  p = nullptr;
  if (n == 0) return hipSuccess;
  return hipMalloc(&p, n * sizeof(T));
}

namespace {

// Capturable build, once the cell count is known: the trees it leaves to the
// eager build (tree_build_exact) stop here.
__global__ void k_tree_gate(TreeMeta* __restrict__ meta, uint32_t cell_cap) {
  if (threadIdx.x != 0) return;
  if (!meta->valid || meta->mixed || meta->overflow || meta->long_group || meta->len > cell_cap) meta->stop = 1;
}

// cell_s[0 .. len) = NONE (the cells k_assign does not name), with the grid
// sized by the capacity.
__global__ __launch_bounds__(EW_THREADS) void k_fill_cells(uint32_t* __restrict__ cell_s,
                                                           const TreeMeta* __restrict__ meta) {
  if (meta->stop) return;
  const uint32_t len = meta->len;
  for (uint32_t i = blockIdx.x * EW_THREADS + threadIdx.x; i < len; i += gridDim.x * EW_THREADS) cell_s[i] = NONE;
}

}  // namespace

hipError_t tree_reserve(TreeBuffers& t, uint32_t n) {
  if (!t.meta) {
    HIPCHECK(hipMalloc(&t.meta, sizeof(TreeMeta)));
    HIPCHECK(hipMemset(t.meta, 0, sizeof(TreeMeta)));
    HIPCHECK(hipHostMalloc(&t.meta_host, sizeof(TreeMeta), hipHostMallocDefault));
    std::memset(t.meta_host, 0, sizeof(TreeMeta));
    HIPCHECK(hipMalloc(&t.box_partials, BOX_BLOCKS * sizeof(BoxPartial)));
    HIPCHECK(hipMalloc(&t.deep_stack, DEEP_STACK * sizeof(Frame)));
    HIPCHECK(hipMalloc(&t.dbins, (MAX_DEPTH + 64) * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&t.rs_glob, 8 * 256 * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&t.dstart, (MAX_DEPTH + 64) * sizeof(uint32_t)));
  }
  if (n <= t.cap) return hipSuccess;
  uint32_t cap = n < 1024 ? 1024 : n;
  HIPCHECK(grow(t.keys, cap));
  HIPCHECK(grow(t.keys_tmp, cap));
  HIPCHECK(grow(t.keys_all, cap));
  HIPCHECK(grow(t.perm, cap));
  HIPCHECK(grow(t.perm_tmp, cap));
  HIPCHECK(grow(t.rest, cap));
  HIPCHECK(grow(t.flags, cap));
  HIPCHECK(grow(t.offs, cap));
  HIPCHECK(grow(t.count, cap));
  HIPCHECK(grow(t.base, cap));
  HIPCHECK(grow(t.dleft, cap));
  HIPCHECK(grow(t.gend, cap));
  HIPCHECK(grow(t.leafdepth, cap));
  HIPCHECK(grow(t.mixed, cap));
  HIPCHECK(grow(t.mixed_list, cap));
  HIPCHECK(grow(t.hist, (size_t)256 * ceil_div(cap, RS_TILE)));
  HIPCHECK(grow(t.partials, (size_t)ceil_div(cap, SCAN_TILE) + 64));
  t.cap = cap;
  return tree_reserve_cells(t, 2 * cap + 64);
}

hipError_t tree_reserve_cells(TreeBuffers& t, uint32_t cells) {
  if (cells <= t.cell_cap) return hipSuccess;
  uint32_t cap = cells + cells / 4;
  HIPCHECK(grow(t.cells, cap));
  HIPCHECK(grow(t.tcells, cap));
  HIPCHECK(grow(t.cdepth, cap));
  HIPCHECK(grow(t.cskip, cap));
  HIPCHECK(grow(t.cut, cap));
  HIPCHECK(grow(t.top, cap));
  HIPCHECK(grow(t.cell_s, cap));
  HIPCHECK(grow(t.cell_d, cap));
  HIPCHECK(grow(t.dorder, cap));
  t.cell_cap = cap;
  return hipSuccess;
}

void tree_free(TreeBuffers& t) {
  void* ptrs[] = {t.keys, t.keys_tmp, t.keys_all, t.perm, t.perm_tmp, t.rest, t.flags, t.offs, t.count,
                  t.base, t.dleft, t.gend, t.leafdepth, t.mixed, t.mixed_list, t.hist, t.partials,
                  t.box_partials, t.deep_stack, t.cells, t.tcells, t.cdepth, t.cskip, t.cut, t.top, t.dbins, t.dstart, t.dorder,
                  t.rs_glob, t.cell_s, t.cell_d, t.meta};
  for (void* p : ptrs) if (p) (void)hipFree(p);
  if (t.meta_host) (void)hipHostFree(t.meta_host);
  t = TreeBuffers();
}

hipError_t tree_build_exact(TreeBuffers& t, const double* x, const double* y, uint32_t n, hipStream_t stream,
                            bool graph) {
  HIPCHECK(tree_reserve(t, n > 0 ? n : 1));
  const uint32_t nb = ceil_div(n > 0 ? n : 1, EW_THREADS);
  uint32_t box_blocks = ceil_div(n > 0 ? n : 1, BOX_THREADS * 4);
  if (box_blocks > (uint32_t)BOX_BLOCKS) box_blocks = BOX_BLOCKS;
  hipLaunchKernelGGL(k_box, dim3(box_blocks), dim3(BOX_THREADS), 0, stream, x, y, n,
                     (BoxPartial*)t.box_partials);
  hipLaunchKernelGGL(k_box_final, dim3(1), dim3(BOX_THREADS), 0, stream, n, box_blocks,
                     (const BoxPartial*)t.box_partials, t.meta);
  hipLaunchKernelGGL(k_keys, dim3(nb), dim3(EW_THREADS), 0, stream, x, y, n, t.meta, t.keys, t.perm, t.keys_all,
                     t.flags);
  // Compaction of NaN points; skipped on the device when m == n.
  const uint32_t* nptr = &t.meta->n;
  const uint32_t* skip = &t.meta->n;  // skip[0] = n, skip[1] = m
  HIPCHECK(scan_exclusive(t.flags, t.offs, n, nptr, t.partials, nullptr, skip, stream));
  hipLaunchKernelGGL(k_compact, dim3(nb), dim3(EW_THREADS), 0, stream, n, t.meta, t.keys_all, t.flags, t.offs,
                     t.keys, t.perm, t.rest);
  hipLaunchKernelGGL(k_rest_all, dim3(nb), dim3(EW_THREADS), 0, stream, n, t.meta, t.rest);
  // Radix sort over m = meta->m keys: 4 passes on the high halves (even: the
  // result ends in keys/perm), then the runs of equal high halves by their
  // low halves; all 8 passes when a run is too long for that.
  const uint32_t* mptr = &t.meta->m;
  HIPCHECK(sort_digits(t.hist, t.rs_glob, t.keys, t.perm, t.keys_tmp, t.perm_tmp, n > 0 ? n : 1, mptr, 4, 4,
                       stream));
  hipLaunchKernelGGL(k_sort_low_halves, dim3(nb), dim3(EW_THREADS), 0, stream, t.keys, t.perm, t.meta);
  const uint32_t zero_blocks = std::min(ceil_div(n > 0 ? n : 1, 256u), 4096u);
  if (!graph) {
    HIPCHECK(hipMemcpyAsync(&t.meta_host->long_group, &t.meta->long_group, sizeof(uint32_t), hipMemcpyDeviceToHost,
                            stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (t.meta_host->long_group)
      HIPCHECK(sort_digits(t.hist, t.rs_glob, t.keys, t.perm, t.keys_tmp, t.perm_tmp, n > 0 ? n : 1, mptr, 0, 8,
                           stream));
  }
  // Key groups.
  hipLaunchKernelGGL(k_zero_words, dim3(zero_blocks), dim3(256), 0, stream, t.mixed, n > 0 ? n : 1);
  hipLaunchKernelGGL(k_group_a, dim3(nb), dim3(EW_THREADS), 0, stream, t.keys, t.perm, x, y, t.meta, t.flags,
                     t.mixed);
  hipLaunchKernelGGL(k_group_b, dim3(nb), dim3(EW_THREADS), 0, stream, t.keys, t.flags, t.mixed, t.meta, t.count,
                     t.dleft, t.gend, t.leafdepth, t.mixed_list);
  hipLaunchKernelGGL(k_deep<false>, dim3(1), dim3(64), 0, stream, t.keys, t.perm, t.perm_tmp, x, y, t.meta,
                     t.mixed_list, t.dleft, t.gend, t.count, t.base, t.deep_stack, t.cells, t.cdepth);
  HIPCHECK(scan_exclusive(t.count, t.base, n, mptr, t.partials, &t.meta->len, nullptr, stream));
  HIPCHECK(hipGetLastError());
  if (graph) {
    const uint32_t cap = t.cell_cap;
    hipLaunchKernelGGL(k_tree_gate, dim3(1), dim3(64), 0, stream, t.meta, cap);
    hipLaunchKernelGGL(k_fill_cells, dim3(std::min(ceil_div(cap, (uint32_t)EW_THREADS), 4096u)), dim3(EW_THREADS), 0,
                       stream, t.cell_s, t.meta);
    hipLaunchKernelGGL(k_assign, dim3(nb), dim3(EW_THREADS), 0, stream, t.flags, t.mixed, t.meta, t.dleft,
                       t.leafdepth, t.base, t.cell_s, t.cell_d);
    hipLaunchKernelGGL(k_emit, dim3(ceil_div(cap, (uint32_t)EW_THREADS)), dim3(EW_THREADS), 0, stream, t.keys,
                       t.mixed, t.meta, t.gend, t.leafdepth, t.base, t.cell_s, t.cell_d, t.cells, t.cdepth, t.cskip,
                       t.cut, t.top);
    return hipGetLastError();
  }
  HIPCHECK(hipMemcpyAsync(t.meta_host, t.meta, sizeof(TreeMeta), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  if (t.meta_host->overflow) HIPFAIL(hipErrorOutOfMemory, "tree arrays overflowed");
  if (!t.meta_host->valid) {
    t.meta_host->len = 0;
    return hipSuccess;
  }
  HIPCHECK(tree_reserve_cells(t, t.meta_host->len));
  const uint32_t len = t.meta_host->len;
  HIPCHECK(hipMemsetAsync(t.cell_s, 0xff, (size_t)len * sizeof(uint32_t), stream));
  hipLaunchKernelGGL(k_assign, dim3(nb), dim3(EW_THREADS), 0, stream, t.flags, t.mixed, t.meta, t.dleft,
                     t.leafdepth, t.base, t.cell_s, t.cell_d);
  hipLaunchKernelGGL(k_emit, dim3(ceil_div(len, EW_THREADS)), dim3(EW_THREADS), 0, stream, t.keys, t.mixed, t.meta,
                     t.gend, t.leafdepth, t.base, t.cell_s, t.cell_d, t.cells, t.cdepth, t.cskip, t.cut, t.top);
  if (t.meta_host->mixed) {
    hipLaunchKernelGGL(k_deep<true>, dim3(1), dim3(64), 0, stream, t.keys, t.perm, t.perm_tmp, x, y, t.meta,
                       t.mixed_list, t.dleft, t.gend, t.count, t.base, t.deep_stack, t.cells, t.cdepth);
  }
  return hipGetLastError();
}


}  // namespace d3f
