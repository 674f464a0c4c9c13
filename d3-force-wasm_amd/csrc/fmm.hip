// Fast many-body on gfx950: the reference's adaptive FMM (src/core/fmm.zig)
// with its tree, geometry and interaction lists reproduced exactly in FP64,
// and the expansions, M2L and near field computed in FP32.
//
//  1. Keys and sort: every graph of the forest gets its box and frame, its
//     points 16-level Morton keys (tree.zig:346-364) with the graph id above
//     them, and one radix sort orders all graphs at once; then the
//     level-order cells (fmm.zig:685-711) one depth at a time: each
//     splitting cell finds its quadrant boundaries by binary search in its
//     key range, a scan of the child counts places the children contiguously
//     in the next level.
//  2. Gather in Morton order (fmm.zig:668-682), geometry (fmm.zig:721-754)
//     and P2M / M2M (fmm.zig:757-806) one depth at a time, bottom up. Bounds,
//     centres and radii are FP64 with the reference's operations, so the
//     opening test sees the same numbers.
//  3. The dual-tree interaction (fmm.zig:400-453) breadth first: one round
//     per pair depth expands the frontier of cell pairs into children pairs,
//     M2L pairs and P2P pairs (block-aggregated appends). The recursion's
//     decisions depend on the pair alone, so the lists hold exactly the
//     reference's pairs.
//  4. Both lists are grouped by target (count, scan, scatter), and each
//     target's group is ordered by source (in LDS, or in place for long
//     groups) before it is read, so every FP32 sum runs in one fixed order
//     whatever the schedule.
//  5. M2L on the matrix cores, one wave per target cell, 16 source cells per
//     product; L2L top down; then P2P and L2P fused, one wave per batch of 64
//     target points, long source groups split into chunks over several
//     blocks, and the sum is added to the FP64 velocity once per node.
//
// The launches are sized either from counts read back once per tick (the
// eager form) or from the buffers' capacities with every kernel reading its
// count on the device (the form a captured graph replays); both run the same
// kernels on the same data.
//
// Expansions are scaled by their cell's radius rho (M_k / rho^k, L_l rho^l),
// so every FP32 quantity stays near unit size at any depth. With w = 1/z0
// (z0 = centre_A - centre_B), a = rho_B w and b = rho_A w:
//   L^_l(A) += (-1)^l b^l w sum_k C(k+l, k) M^_k(B) a^k,
// which is the reference's T_{k+l} = (k+l)! / z0^{k+l+1} M2L (fmm.zig:561-610)
// with its 1/k! and (-1)^l / l! scalings (fmm.zig:629-657) folded in.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "common.h"

namespace d3f {
namespace {

constexpr int EWF = 256;
constexpr uint32_t SEG_MAX = 256;       // M2L sources of one target ordered in LDS (longer: in place)
constexpr uint32_t NEAR_CHUNK = 64;     // source leaves per near-field chunk
constexpr uint32_t NEAR_ALL = 0xFFFFFFFFu;
constexpr uint32_t LONG_IDS = 262144;   // cell ids per pass of the long-segment sort (32 KB of LDS bitmap)
constexpr uint32_t LONG_BLOCKS = 128;
constexpr double MIN_SEP2 = 1e-28;      // fmm.zig:49
constexpr int BN = 2 * FMM_MAX_ORDER + 1;

typedef float f32x4 __attribute__((ext_vector_type(4)));

__constant__ float c_binom[BN * BN];  // C(n, k), n, k <= 2 * FMM_MAX_ORDER

__device__ inline float binom(int n, int k) { return c_binom[n * BN + k]; }

// Zig's @min / @max on floats as the reference uses them: the first operand
// unless the second is strictly smaller (larger); a NaN first operand yields
// the second.
__device__ inline double zmin(double a, double b) { return (b < a || a != a) ? b : a; }
__device__ inline double zmax(double a, double b) { return (b > a || a != a) ? b : a; }

__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ inline float2 cfma(float2 a, float2 b, float2 c) {  // a * b + c
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}
__device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

// ---------------------------------------------------------------- build

// A chunk's box of its non-NaN points; equal values go to the lower node,
// as the serial scan keeps the first (tree.zig:237-244).
struct BoxPart {
  double x0, y0, x1, y1;
  uint32_t ix0, iy0, ix1, iy1, m, pad;
};
__device__ inline void bmin(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void bmax(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void bcombine(BoxPart& a, const BoxPart& b) {
  bmin(a.x0, a.ix0, b.x0, b.ix0);
  bmin(a.y0, a.iy0, b.y0, b.iy0);
  bmax(a.x1, a.ix1, b.x1, b.ix1);
  bmax(a.y1, a.iy1, b.y1, b.iy1);
  a.m += b.m;
}
__device__ inline BoxPart bempty() {
  const double inf = __builtin_huge_val();
  return BoxPart{inf, inf, -inf, -inf, NONE, NONE, NONE, NONE, 0, 0};
}

__global__ void k_forest_reset_overflow(FmmParams* __restrict__ fp) {
  if (threadIdx.x == 0) fp->cell_overflow = 0;
}

__global__ void k_forest_reset(FmmParams* __restrict__ fp, uint32_t n) {
  if (threadIdx.x == 0) {
    fp->n = n;
    fp->m_total = 0;
    fp->roots = 0;
    fp->declined = 0;
    fp->cell_overflow = 0;
  }
}

// One block per chunk of a graph's nodes (chunks never straddle graphs).
__global__ __launch_bounds__(EWF) void k_forest_box(const double* __restrict__ xs, const double* __restrict__ ys,
                                                    const uint32_t* __restrict__ chunk_lo,
                                                    const uint32_t* __restrict__ chunk_hi,
                                                    BoxPart* __restrict__ parts) {
  const uint32_t lo = chunk_lo[blockIdx.x], hi = chunk_hi[blockIdx.x];
  BoxPart p = bempty();
  for (uint32_t i = lo + threadIdx.x; i < hi; i += EWF) {
    const double x = xs[i], y = ys[i];
    if (x != x || y != y) continue;
    bmin(p.x0, p.ix0, x, i);
    bmin(p.y0, p.iy0, y, i);
    bmax(p.x1, p.ix1, x, i);
    bmax(p.y1, p.iy1, y, i);
    p.m++;
  }
  __shared__ BoxPart sh[EWF];
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int k = EWF / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) bcombine(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) parts[blockIdx.x] = sh[0];
}

// One wave per graph: its box from its chunks' boxes (the lanes combine
// every 64th chunk, then a tree in LDS; the combine is order-free, ties going
// to the lower node), the fast extent (tree.zig:357-359), the unit frame
// (fmm.zig:253-270) and the scaled distance limits (fmm.zig:272-281). A
// graph of fewer than min_n nodes is given no point (no tree).
__global__ __launch_bounds__(64) void k_forest_graphs(const BoxPart* __restrict__ parts,
                                                      const uint32_t* __restrict__ chunk_first, uint32_t graphs,
                                                      double dmin2, double dmax2, const uint64_t* __restrict__ seeds,
                                                      const uint32_t* __restrict__ node_off, uint32_t min_n,
                                                      FmmGraph* __restrict__ gp, FmmParams* __restrict__ fp) {
  __shared__ BoxPart sh[64];
  const uint32_t g = blockIdx.x;
  if (g >= graphs) return;
  BoxPart b = bempty();
  if (node_off[g + 1] - node_off[g] >= min_n)
    for (uint32_t c = chunk_first[g] + threadIdx.x; c < chunk_first[g + 1]; c += 64) bcombine(b, parts[c]);
  sh[threadIdx.x] = b;
  __syncthreads();
  for (int k = 32; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) bcombine(sh[threadIdx.x], sh[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  b = sh[0];
  FmmGraph r;
  memset(&r, 0, sizeof(r));
  const double nan = __builtin_nan(""), inf = __builtin_huge_val();
  r.box[0] = b.x0;
  r.box[1] = b.y0;
  r.box[2] = b.x1;
  r.box[3] = b.y1;
  for (int k = 0; k < 4; ++k) r.extent[k] = nan;
  r.m = b.m;
  r.seed = seeds[g];
  r.root = NONE;
  r.node_base = node_off[g];
  r.node_end = node_off[g + 1];
  if (b.m) {
    double side = zmax(b.x1 - b.x0, b.y1 - b.y0);
    if (!(side > 0)) side = 1;
    r.extent[0] = b.x0;
    r.extent[1] = b.y0;
    r.extent[2] = b.x0 + side;
    r.extent[3] = b.y0 + side;
    r.scale = 65536.0 / side;
    r.w = r.extent[2] - r.extent[0];
    r.ok = r.w > 0 && r.w < inf;
    r.inv_w = 1 / r.w;
    r.ox = (r.extent[0] + r.extent[2]) / 2;
    r.oy = (r.extent[1] + r.extent[3]) / 2;
    r.dmax_s = dmax2 < inf ? __builtin_sqrt(dmax2) / r.w : inf;
    r.dmin_s = dmin2 > 0 ? __builtin_sqrt(dmin2) / r.w : 0;
    r.ring = r.dmin_s > 0 || r.dmax_s < inf;
    if (!r.ok) atomicAdd(&fp->declined, 1u);
  }
  gp[g] = r;
  if (b.m) atomicAdd(&fp->m_total, b.m);
}

// Bits 0..15 of v moved to the even bit positions (tree.zig:688-695).
__device__ inline uint32_t spread16(uint32_t v) {
  v &= 0xffffu;
  v = (v | (v << 8)) & 0x00ff00ffu;
  v = (v | (v << 4)) & 0x0f0f0f0fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Fast keys (tree.zig:346-364) with the graph above the 32 Morton bits, so
// one sort orders every graph's points, graph by graph.
__global__ __launch_bounds__(EWF) void k_forest_keys(const double* __restrict__ xs, const double* __restrict__ ys,
                                                     uint32_t n, const uint32_t* __restrict__ node_graph,
                                                     const FmmGraph* __restrict__ gp,
                                                     const FmmParams* __restrict__ fp, uint64_t* __restrict__ keys,
                                                     uint32_t* __restrict__ perm, uint64_t* __restrict__ keys_all,
                                                     uint32_t* __restrict__ flags) {
  const uint32_t i = blockIdx.x * EWF + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = node_graph ? node_graph[i] : 0;
  const double x = xs[i], y = ys[i];
  const bool valid = !(x != x || y != y) && gp[g].m != 0;  // a graph without points: NaN or left out
  uint64_t key = 0;
  if (valid) {
    const FmmGraph& G = gp[g];
    double fx = floor((x - G.box[0]) * G.scale), fy = floor((y - G.box[1]) * G.scale);
    fx = (65535.0 < fx || fx != fx) ? 65535.0 : fx;
    fy = (65535.0 < fy || fy != fy) ? 65535.0 : fy;
    key = (uint64_t)g << 32 | (uint64_t)(spread16((uint32_t)fx) | spread16((uint32_t)fy) << 1);
  }
  if (fp->m_total == n) {
    keys[i] = key;
    perm[i] = i;
  } else {
    keys_all[i] = key;
    flags[i] = valid ? 1u : 0u;
  }
}

__global__ __launch_bounds__(EWF) void k_forest_compact(uint32_t n, const FmmParams* __restrict__ fp,
                                                        const uint64_t* __restrict__ keys_all,
                                                        const uint32_t* __restrict__ flags,
                                                        const uint32_t* __restrict__ offs, uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ perm) {
  const uint32_t i = blockIdx.x * EWF + threadIdx.x;
  if (i >= n || fp->m_total == n || !flags[i]) return;
  keys[offs[i]] = keys_all[i];
  perm[offs[i]] = i;
}

// One block: each graph's first point in Morton order (the sorted keys hold
// the graphs in order) and the index of its root cell for every graph with
// points and a finite extent, roots numbered in graph order (level 0 of the
// forest; k_cells_emit writes them).
__global__ __launch_bounds__(1024) void k_forest_roots(FmmGraph* __restrict__ gp, uint32_t graphs,
                                                       FmmParams* __restrict__ fp, TreeMeta* __restrict__ meta) {
  __shared__ uint32_t wsum[2][16];
  __shared__ uint32_t carry[2];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (threadIdx.x == 0) carry[0] = carry[1] = 0;
  __syncthreads();
  for (uint32_t g0 = 0; g0 < graphs; g0 += 1024) {
    const uint32_t g = g0 + threadIdx.x;
    const bool in = g < graphs;
    const uint32_t m = in ? gp[g].m : 0;
    const uint32_t root = in && m && gp[g].ok ? 1u : 0u;
    uint32_t v[2] = {m, root}, inc[2];
    for (int k = 0; k < 2; ++k) {
      inc[k] = v[k];
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inc[k], d, 64);
        if (lane >= d) inc[k] += t;
      }
      if (lane == 63) wsum[k][wave] = inc[k];
    }
    __syncthreads();
    uint32_t before[2] = {carry[0], carry[1]}, total[2] = {0, 0};
    for (int w = 0; w < 16; ++w)
      for (int k = 0; k < 2; ++k) {
        if (w < wave) before[k] += wsum[k][w];
        total[k] += wsum[k][w];
      }
    if (in) {
      const uint32_t first = before[0] + inc[0] - m, idx = before[1] + inc[1] - root;
      gp[g].point_off = first;
      gp[g].root = root ? idx : NONE;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      carry[0] += total[0];
      carry[1] += total[1];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    fp->roots = carry[1];
    // prepareFast leaves no exact tree behind (tree.zig:351-356).
    for (int k = 0; k < 4; ++k) meta->extent[k] = gp[0].extent[k];
    meta->len = 0;
    meta->valid = 0;
  }
}

// ---- cells (fmm.zig:685-711) from the sorted keys, every depth at once
//
// The reference builds level-order cells breadth first: a cell of more than
// FMM_LEAF points above depth FMM_LEVELS splits into its non-empty quadrants.
// With the keys sorted, the cells follow from per-point quantities:
//   b(k)  = the quad digits keys k - 1 and k share (-1 at a graph's first
//           point, where a new root begins); point k starts cells at depths
//           b(k) + 1 and deeper;
//   W(j)  = the digits keys j and j + FMM_LEAF share (-1 past the end or
//           across graphs): the depth-d cell holding k has more than FMM_LEAF
//           points exactly when W(j) >= d for some j in [k - FMM_LEAF, k];
//   L(k)  = min(FMM_LEVELS, 1 + max of those W(j)): the depth of the leaf
//           holding k.
// So point k starts one cell at every depth d with b(k) + 1 <= d <= L(k),
// and a cell's level-order index is the number of cells of its depth that
// start at earlier points. The build is three launches: per-point b and L
// with per-block counts and first breaks per depth (k_cells_mark); per depth
// a scan of the counts and a suffix minimum of the breaks over the blocks
// (k_cells_scan); then every cell written with its count, parent and first
// child (k_cells_emit). Children after the first add themselves to their
// parent's nchild in k_fmm_gather.
constexpr uint32_t CB_THREADS = 256, CB_PTS = 1024;  // points per block of the cell build (4 per thread)
constexpr uint32_t CB_DEPTHS = FMM_LEVELS + 1;       // depths 0 .. FMM_LEVELS
constexpr uint32_t CB_HALO = FMM_LEAF + 1;           // keys kept before the block's first point
constexpr uint32_t UP_SMALL = 64;                    // upward pass: cells of at most this many points get 8 lanes
// cells of more than UP_SMALL points starting in one block: they are
// disjoint, so at most CB_PTS / (UP_SMALL + 1) lie inside it and one more
// runs past its end, per depth
constexpr uint32_t CB_MIDS = CB_DEPTHS * (CB_PTS / (UP_SMALL + 1) + 1);

// Shared quad digits of two keys of one graph (FMM_LEVELS when equal), -1
// for keys of different graphs.
__device__ inline int key_lcp(uint64_t a, uint64_t b) {
  if ((a >> 32) != (b >> 32)) return -1;
  const uint32_t x = (uint32_t)a ^ (uint32_t)b;
  return x == 0 ? (int)FMM_LEVELS : (int)(__clz(x) >> 1);
}

// The key prefix that names a depth-d cell (the graph id above it).
__device__ inline uint64_t key_prefix(uint64_t key, uint32_t d) { return key >> (2 * (FMM_LEVELS - d)); }

__global__ __launch_bounds__(CB_THREADS) void k_cells_mark(const uint64_t* __restrict__ keys,
                                                           const FmmGraph* __restrict__ gp,
                                                           const FmmParams* __restrict__ fp, uint32_t nblk,
                                                           uint16_t* __restrict__ fl, uint32_t* __restrict__ blk_cnt,
                                                           uint32_t* __restrict__ blk_fb) {
  __shared__ uint64_t sk[CB_HALO + CB_PTS + FMM_LEAF];  // keys [base - CB_HALO, base + CB_PTS + FMM_LEAF)
  __shared__ int8_t sw[FMM_LEAF + CB_PTS];              // W(j), j in [base - FMM_LEAF, base + CB_PTS)
  __shared__ uint32_t cnt[CB_DEPTHS], fb[CB_DEPTHS];
  const uint32_t m = fp->m_total;
  const int64_t base = (int64_t)blockIdx.x * CB_PTS;
  for (uint32_t i = threadIdx.x; i < CB_HALO + CB_PTS + FMM_LEAF; i += CB_THREADS) {
    const int64_t k = base - CB_HALO + i;
    sk[i] = (k >= 0 && k < (int64_t)m) ? keys[k] : 0;
  }
  if (threadIdx.x < CB_DEPTHS) {
    cnt[threadIdx.x] = 0;
    fb[threadIdx.x] = NONE;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < FMM_LEAF + CB_PTS; i += CB_THREADS) {
    const int64_t j = base - FMM_LEAF + i;  // sk index of key j: i + 1
    sw[i] = (int8_t)((j >= 0 && j + FMM_LEAF < (int64_t)m) ? key_lcp(sk[i + 1], sk[i + 1 + FMM_LEAF]) : -1);
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t t = q * CB_THREADS + threadIdx.x;
    const int64_t k = base + t;
    uint32_t first = 1, leafp1 = 0;  // no cell
    if (k < (int64_t)m) {
      const uint64_t key = sk[CB_HALO + t];
      const int b = k == 0 ? -1 : key_lcp(sk[CB_HALO + t - 1], key);
      for (uint32_t d = (uint32_t)(b + 1); d < CB_DEPTHS; ++d) atomicMin(&fb[d], (uint32_t)k);
      if (gp[key >> 32].ok) {
        int mw = -1;
        for (uint32_t s = 0; s <= FMM_LEAF; ++s) mw = max(mw, (int)sw[t + s]);
        first = (uint32_t)(b + 1);
        leafp1 = (uint32_t)min((int)FMM_LEVELS, mw + 1) + 1;
        for (uint32_t d = first; d < leafp1; ++d) atomicAdd(&cnt[d], 1u);
      }
      fl[k] = (uint16_t)(first | leafp1 << 8);
    }
  }
  __syncthreads();
  if (threadIdx.x < CB_DEPTHS) {
    blk_cnt[(size_t)threadIdx.x * nblk + blockIdx.x] = cnt[threadIdx.x];
    blk_fb[(size_t)threadIdx.x * nblk + blockIdx.x] = fb[threadIdx.x];
  }
}

// Block d: the exclusive scan of depth d's per-block cell counts (the
// level-order rank of each block's first depth-d cell) and the suffix
// minimum of its first breaks (the first point at or after each block that
// ends a depth-d cell). Block 0 also clears the counts of the cell lists.
__global__ __launch_bounds__(1024) void k_cells_scan(const uint32_t* __restrict__ blk_cnt,
                                                     const uint32_t* __restrict__ blk_fb, uint32_t nblk,
                                                     uint32_t* __restrict__ blk_off, uint32_t* __restrict__ blk_nb,
                                                     FmmParams* __restrict__ fp, uint32_t cell_cap) {
  __shared__ uint32_t ws[16];
  __shared__ uint32_t carry;
  const uint32_t d = blockIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t* cnt = blk_cnt + (size_t)d * nblk;
  uint32_t* off = blk_off + (size_t)d * nblk;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (uint32_t b0 = 0; b0 < nblk; b0 += 1024) {
    const uint32_t b = b0 + threadIdx.x;
    const uint32_t v = b < nblk ? cnt[b] : 0;
    uint32_t inc = v;
    for (int s = 1; s < 64; s <<= 1) {
      const uint32_t u = __shfl_up(inc, s, 64);
      if (lane >= s) inc += u;
    }
    if (lane == 63) ws[wave] = inc;
    __syncthreads();
    uint32_t before = carry, total = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wave) before += ws[w];
      total += ws[w];
    }
    if (b < nblk) off[b] = before + inc - v;
    __syncthreads();
    if (threadIdx.x == 0) carry += total;
    __syncthreads();
  }
  if (threadIdx.x == 0) fp->level_n[d] = carry;
  // suffix minimum, right to left
  const uint32_t* fb = blk_fb + (size_t)d * nblk;
  uint32_t* nb = blk_nb + (size_t)d * nblk;
  if (threadIdx.x == 0) carry = NONE;
  __syncthreads();
  for (uint32_t e = nblk; e > 0;) {
    const uint32_t s0 = e > 1024 ? e - 1024 : 0;
    const uint32_t b = s0 + threadIdx.x;  // this pass covers [s0, e)
    uint32_t v = b < e ? fb[b] : NONE;
    for (int s = 1; s < 64; s <<= 1) {
      const uint32_t u = __shfl_down(v, s, 64);
      if (lane + s < 64) v = min(v, u);
    }
    if (lane == 0) ws[wave] = v;
    __syncthreads();
    uint32_t after = carry;
    for (int w = wave + 1; w < 16; ++w) after = min(after, ws[w]);
    v = min(v, after);
    if (b < e) nb[b] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t lo = carry;
      for (int w = 0; w < 16; ++w) lo = min(lo, ws[w]);
      carry = lo;
    }
    __syncthreads();
    e = s0;
  }
  if (d == 0 && threadIdx.x < CB_DEPTHS) fp->big_n[threadIdx.x] = 0;
  if (d == 0 && threadIdx.x == 0) {
    fp->level_n[FMM_LEVELS + 1] = 0;
    fp->mid_n = 0;
  }
}

// The cells. Block 0 publishes the level offsets; every block ranks its
// points' cells per depth with wave ballots (points in the order q, wave,
// lane) and writes them: count from the block's keys when the cell ends in
// the block, else from the next break after it; the parent from this
// thread's previous cell (a first child) or from the rank of the cell of the
// depth above holding the point (a later child); cells of more than
// FMM_DIRECT points also go to their depth's list for k_fmm_up_big, cells of
// UP_SMALL + 1 .. FMM_DIRECT points to the list for k_fmm_up_mid.
__global__ __launch_bounds__(CB_THREADS) void k_cells_emit(const uint64_t* __restrict__ keys,
                                                           const uint16_t* __restrict__ fl, uint32_t nblk,
                                                           const uint32_t* __restrict__ blk_off,
                                                           const uint32_t* __restrict__ blk_nb,
                                                           FmmParams* __restrict__ fp, FCell* __restrict__ cells,
                                                           uint32_t* __restrict__ parent, uint32_t cell_cap,
                                                           uint32_t* __restrict__ big, uint32_t big_cap,
                                                           uint32_t* __restrict__ mid, uint32_t mid_cap) {
  __shared__ uint64_t sk[CB_PTS];
  __shared__ uint32_t loff[CB_DEPTHS + 1];
  __shared__ uint32_t woff[CB_DEPTHS][16];  // per depth: exclusive offsets of the (q, wave) groups
  __shared__ uint32_t mids[CB_MIDS], n_mid, mid_base;  // the block's part of the mid-size list
  const uint32_t m = fp->m_total;
  const uint32_t base = blockIdx.x * CB_PTS;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1;
  if (threadIdx.x == 0) {
    uint32_t o = 0;
    for (uint32_t d = 0; d < CB_DEPTHS; ++d) {
      loff[d] = o;
      o += fp->level_n[d];
    }
    loff[CB_DEPTHS] = o;
    n_mid = 0;
    if (blockIdx.x == 0) {
      for (uint32_t d = 0; d <= CB_DEPTHS; ++d) fp->level_off[d] = loff[d];
      fp->level_off[FMM_LEVELS + 2] = o;
      if (o > cell_cap) fp->cell_overflow = 1;
    }
  }
  for (uint32_t i = threadIdx.x; i < CB_PTS; i += CB_THREADS) sk[i] = base + i < m ? keys[base + i] : 0;
  uint32_t first[4], leafp1[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint32_t k = base + q * CB_THREADS + threadIdx.x;
    const uint32_t v = k < m ? fl[k] : 1u;
    first[q] = v & 0xffu;
    leafp1[q] = v >> 8;
  }
  for (uint32_t d = 0; d < CB_DEPTHS; ++d)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint64_t bal = __ballot(first[q] <= d && d < leafp1[q]);
      if (lane == 0) woff[d][q * 4 + wave] = (uint32_t)__popcll(bal);
    }
  __syncthreads();
  if (threadIdx.x < CB_DEPTHS) {
    uint32_t o = blk_off[(size_t)threadIdx.x * nblk + blockIdx.x];
    for (int g = 0; g < 16; ++g) {
      const uint32_t c = woff[threadIdx.x][g];
      woff[threadIdx.x][g] = o;
      o += c;
    }
  }
  __syncthreads();
  const uint32_t total = loff[CB_DEPTHS];
  uint32_t prev_idx[4], prev_excl[4];
  for (uint32_t d = 0; d < CB_DEPTHS; ++d) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bool has = first[q] <= d && d < leafp1[q];
      const uint64_t bal = __ballot(has);
      const uint32_t excl = woff[d][q * 4 + wave] + (uint32_t)__popcll(bal & lt);  // depth-d cells before the point
      if (has) {
        const uint32_t t = q * CB_THREADS + threadIdx.x;
        const uint32_t k = base + t;
        const uint32_t idx = loff[d] + excl;
        const uint64_t key = sk[t];
        const uint64_t P = key_prefix(key, d);
        // the first point after k outside the cell: the block's keys, else
        // the first break at or after the next block
        uint32_t end;
        if (t + 1 >= CB_PTS || base + t + 1 >= m || key_prefix(sk[t + 1], d) != P) {
          end = k + 1;
          if (t + 1 >= CB_PTS && base + CB_PTS < m)
            end = blockIdx.x + 1 < nblk ? blk_nb[(size_t)d * nblk + blockIdx.x + 1] : NONE;
        } else {
          uint32_t a = t + 1, step = 1, b;
          for (;;) {
            b = a + step;
            if (b >= CB_PTS || base + b >= m || key_prefix(sk[b], d) != P) break;
            a = b;
            step <<= 1;
          }
          if (b > CB_PTS) b = CB_PTS;
          if (base + b > m) b = m - base;
          while (b - a > 1) {
            const uint32_t mid = a + (b - a) / 2;
            if (key_prefix(sk[mid], d) == P) a = mid; else b = mid;
          }
          end = base + b;
          if (b == CB_PTS && base + CB_PTS < m)  // the cell runs into the next block
            end = blockIdx.x + 1 < nblk ? blk_nb[(size_t)d * nblk + blockIdx.x + 1] : NONE;
        }
        if (end > m) end = m;
        uint32_t par = NONE;
        if (d > 0) par = d > first[q] ? prev_idx[q] : loff[d - 1] + prev_excl[q] - 1;
        const uint32_t split = d + 1 < leafp1[q] ? 1u : 0u;
        if (total <= cell_cap) {
          cells[idx] = FCell{0, 0, 0, k, end - k, 0, split, d, (uint32_t)(key >> 32)};
          parent[idx] = par;
          if (d > first[q]) cells[par].child0 = idx;  // this thread wrote the parent one depth above
          if (end - k > FMM_DIRECT) {
            const uint32_t s = atomicAdd(&fp->big_n[d], 1u);
            if (s < big_cap) big[(size_t)d * big_cap + s] = idx;
          } else if (end - k > UP_SMALL) {
            mids[atomicAdd(&n_mid, 1u)] = idx;
          }
        }
        prev_idx[q] = idx;
      }
      prev_excl[q] = excl + (has ? 1u : 0u);  // depth-d cells up to and including the point
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && n_mid) mid_base = atomicAdd(&fp->mid_n, n_mid);
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < n_mid; i += CB_THREADS)
    if (mid_base + i < mid_cap) mid[mid_base + i] = mids[i];
}


// fmm.zig:668-682: Morton-ordered raw positions, unit-frame positions and
// strength * alpha (NaN -> 0), each point in its graph's frame.
__global__ __launch_bounds__(EWF) void k_fmm_gather(const uint32_t* __restrict__ perm, const uint64_t* __restrict__ keys,
                                                    const double* __restrict__ xs, const double* __restrict__ ys,
                                                    const double* __restrict__ st, const FmmGraph* __restrict__ gp,
                                                    const FmmParams* __restrict__ fp, double alpha,
                                                    const double* __restrict__ alpha_dev, uint32_t n,
                                                    double* __restrict__ px, double* __restrict__ py,
                                                    double* __restrict__ ux, double* __restrict__ uy,
                                                    float* __restrict__ ps, FCell* __restrict__ cells,
                                                    const uint32_t* __restrict__ parent) {
  if (alpha_dev) alpha = *alpha_dev;
  const uint32_t j = blockIdx.x * EWF + threadIdx.x;
  if (j < n && j < fp->m_total) {
    const FmmGraph& G = gp[keys[j] >> 32];
    if (G.ok) {
      const uint32_t i = perm[j];
      const double x = xs[i], y = ys[i];
      px[j] = x;
      py[j] = y;
      const double s = st[i] * alpha;
      ps[j] = s != s ? 0.0f : (float)s;
      ux[j] = (x - G.ox) * G.inv_w;
      uy[j] = (y - G.oy) * G.inv_w;
    }
  }
  // k_cells_emit gave each splitting cell its first child; every later child
  // (one starting after its parent's first point) adds itself.
  const uint32_t nc = fp->ncg;
  for (uint32_t c = j; c < nc; c += gridDim.x * EWF) {
    const uint32_t p = parent[c];
    if (p != NONE && cells[p].first != cells[c].first) atomicAdd(&cells[p].nchild, 1u);
  }
}

// ---------------------------------------------------------------- upward

// Geometry (fmm.zig:721-754) of one cell: the bounding box of its points
// (a leaf) or of its children's boxes, centre = the box's midpoint, radius =
// half its diagonal, in FP64 with the reference's operations.
__device__ __forceinline__ void geom_finish(uint32_t c, double x0, double y0, double x1, double y1,
                                            FCell* __restrict__ cells, double* __restrict__ bbox) {
  bbox[4 * (size_t)c + 0] = x0;
  bbox[4 * (size_t)c + 1] = y0;
  bbox[4 * (size_t)c + 2] = x1;
  bbox[4 * (size_t)c + 3] = y1;
  const double dx = x1 - x0, dy = y1 - y0;
  cells[c].cx = (x0 + x1) / 2;
  cells[c].cy = (y0 + y1) / 2;
  cells[c].r = 0.5 * __builtin_sqrt(dx * dx + dy * dy);
}

// A large cell's geometry, one thread: its points' box for a leaf of the last
// depth, else the union of its children's boxes.
__device__ __forceinline__ void geom_cell(uint32_t c, FCell* __restrict__ cells, double* __restrict__ bbox,
                                          const double* __restrict__ ux, const double* __restrict__ uy) {
  const FCell cell = cells[c];
  const double inf = __builtin_huge_val();
  double x0 = inf, y0 = inf, x1 = -inf, y1 = -inf;
  if (cell.nchild == 0) {
    for (uint32_t k = cell.first; k < cell.first + cell.count; ++k) {
      x0 = zmin(x0, ux[k]);
      x1 = zmax(x1, ux[k]);
      y0 = zmin(y0, uy[k]);
      y1 = zmax(y1, uy[k]);
    }
  } else {
    for (uint32_t ch = cell.child0; ch < cell.child0 + cell.nchild; ++ch) {
      x0 = zmin(x0, bbox[4 * (size_t)ch + 0]);
      x1 = zmax(x1, bbox[4 * (size_t)ch + 2]);
      y0 = zmin(y0, bbox[4 * (size_t)ch + 1]);
      y1 = zmax(y1, bbox[4 * (size_t)ch + 3]);
    }
  }
  geom_finish(c, x0, y0, x1, y1, cells, bbox);
}

// A large cell's radius-scaled multipole (fmm.zig:757-806), one thread: P2M
// for a leaf of the last depth, else M2M from its children, whose multipoles
// and geometry are done.
template <int Q>
__device__ __forceinline__ void mp_cell(uint32_t c, const FCell* __restrict__ cells, const double* __restrict__ ux,
                                        const double* __restrict__ uy, const float* __restrict__ ps,
                                        float2* __restrict__ mexp) {
  const FCell cell = cells[c];
  const double cx = cell.cx, cy = cell.cy, r = cell.r;
  float2 m[Q];
#pragma unroll
  for (int k = 0; k < Q; ++k) m[k] = make_float2(0.f, 0.f);
  const double inv = r > 0 ? 1 / r : 0;
  if (cell.nchild == 0) {
    for (uint32_t k = cell.first; k < cell.first + cell.count; ++k) {
      const float s = ps[k];
      const float2 z = make_float2((float)((ux[k] - cx) * inv), (float)((uy[k] - cy) * inv));
      float2 zp = make_float2(s, 0.f);
      m[0].x += s;
#pragma unroll
      for (int o = 1; o < Q; ++o) {
        zp = cmul(zp, z);
        m[o].x += zp.x;
        m[o].y += zp.y;
      }
    }
  } else {
    for (uint32_t ch = cell.child0; ch < cell.child0 + cell.nchild; ++ch) {
      const FCell cc = cells[ch];
      const float tr = (float)(cc.r * inv);
      const float2 e = make_float2((float)((cc.cx - cx) * inv), (float)((cc.cy - cy) * inv));
      float2 mc[Q], ep[Q];
      float tk = 1.f;
      ep[0] = make_float2(1.f, 0.f);
#pragma unroll
      for (int k = 0; k < Q; ++k) {
        mc[k] = cscale(mexp[(size_t)ch * Q + k], tk);
        tk *= tr;
        if (k > 0) ep[k] = cmul(ep[k - 1], e);
      }
      // M^_n(P) += sum_k C(n, k) (rho_c / rho_P)^k M^_k(c) (d / rho_P)^(n-k)
#pragma unroll
      for (int nn = 0; nn < Q; ++nn) {
        float2 acc = make_float2(0.f, 0.f);
#pragma unroll
        for (int k = 0; k <= nn; ++k) acc = cfma(cscale(mc[k], binom(nn, k)), ep[nn - k], acc);
        m[nn].x += acc.x;
        m[nn].y += acc.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < Q; ++k) mexp[(size_t)c * Q + k] = m[k];
}

// A cell of up to FMM_DIRECT points straight from its points, by a group of
// G lanes (lane l takes points l, l + G, ...; lane results combined in a
// fixed butterfly): its geometry (the bounds of its points are the union of
// its children's: min and max are exact), or, with the geometry done, its
// multipole (the multipole of its points is what M2M from its children sums
// to).
template <int G>
__device__ __forceinline__ void geom_direct(uint32_t c, uint32_t first, uint32_t count, int lane,
                                            FCell* __restrict__ cells, double* __restrict__ bbox,
                                            const double* __restrict__ ux, const double* __restrict__ uy) {
  const double inf = __builtin_huge_val();
  double x0 = inf, y0 = inf, x1 = -inf, y1 = -inf;
  for (uint32_t k = first + lane; k < first + count; k += G) {
    const double u = ux[k], v = uy[k];
    x0 = zmin(x0, u);
    x1 = zmax(x1, u);
    y0 = zmin(y0, v);
    y1 = zmax(y1, v);
  }
#pragma unroll
  for (int s = G / 2; s > 0; s >>= 1) {
    x0 = zmin(x0, __shfl_xor(x0, s, 64));
    x1 = zmax(x1, __shfl_xor(x1, s, 64));
    y0 = zmin(y0, __shfl_xor(y0, s, 64));
    y1 = zmax(y1, __shfl_xor(y1, s, 64));
  }
  if (lane == 0) geom_finish(c, x0, y0, x1, y1, cells, bbox);
}

template <int Q, int G>
__device__ __forceinline__ void mp_direct(uint32_t c, uint32_t first, uint32_t count, int lane, double cx, double cy,
                                          double r, const double* __restrict__ ux, const double* __restrict__ uy,
                                          const float* __restrict__ ps, float2* __restrict__ mexp) {
  const double inv = r > 0 ? 1 / r : 0;
  float2 mm[Q];
#pragma unroll
  for (int o = 0; o < Q; ++o) mm[o] = make_float2(0.f, 0.f);
  for (uint32_t k = first + lane; k < first + count; k += G) {
    const float s = ps[k];
    const float2 z = make_float2((float)((ux[k] - cx) * inv), (float)((uy[k] - cy) * inv));
    float2 zp = make_float2(s, 0.f);
    mm[0].x += s;
#pragma unroll
    for (int o = 1; o < Q; ++o) {
      zp = cmul(zp, z);
      mm[o].x += zp.x;
      mm[o].y += zp.y;
    }
  }
#pragma unroll
  for (int o = 0; o < Q; ++o)
#pragma unroll
    for (int s = G / 2; s > 0; s >>= 1) {
      mm[o].x += __shfl_xor(mm[o].x, s, 64);
      mm[o].y += __shfl_xor(mm[o].y, s, 64);
    }
#pragma unroll
  for (int o = 0; o < Q; ++o)
    if (o % G == lane) mexp[(size_t)c * Q + o] = mm[o];
}

// The upward pass in two halves: geometry (what the dual traversal needs)
// and multipoles (what M2L needs), each by the cells' sizes: up to UP_SMALL
// points 8 lanes a cell, 8 cells a wave (every cell is visited and the larger
// ones skipped); UP_SMALL + 1 .. FMM_DIRECT points (listed by k_cells_emit) a
// wave each; the larger ones (at most m / FMM_DIRECT a depth, listed per
// depth) by one workgroup, deepest depth first, from their children.
// MP = false runs the geometry, MP = true the multipoles.
template <int Q, bool MP>
__global__ __launch_bounds__(256) void k_fmm_up_small(FCell* __restrict__ cells, double* __restrict__ bbox,
                                                      const double* __restrict__ ux, const double* __restrict__ uy,
                                                      const float* __restrict__ ps, float2* __restrict__ mexp,
                                                      const FmmParams* __restrict__ fp) {
  if (fp->abort) return;
  const uint32_t nc = fp->ncg;
  const int lane = threadIdx.x & 7;
  const uint32_t groups = gridDim.x * 32;
  for (uint32_t c0 = blockIdx.x * 32; c0 < nc; c0 += groups) {
    const uint32_t c = c0 + (threadIdx.x >> 3);  // the groups of a wave stay together
    if (c >= nc) continue;
    const FCell cell = cells[c];
    if (cell.count > UP_SMALL) continue;
    if (MP) mp_direct<Q, 8>(c, cell.first, cell.count, lane, cell.cx, cell.cy, cell.r, ux, uy, ps, mexp);
    else geom_direct<8>(c, cell.first, cell.count, lane, cells, bbox, ux, uy);
  }
}

template <int Q, bool MP>
__global__ __launch_bounds__(256) void k_fmm_up_mid(FCell* __restrict__ cells, double* __restrict__ bbox,
                                                    const double* __restrict__ ux, const double* __restrict__ uy,
                                                    const float* __restrict__ ps, float2* __restrict__ mexp,
                                                    const uint32_t* __restrict__ mid, uint32_t mid_cap,
                                                    const FmmParams* __restrict__ fp) {
  if (fp->abort) return;
  const uint32_t cnt = min(fp->mid_n, mid_cap);
  for (uint32_t i = blockIdx.x * 4 + (threadIdx.x >> 6); i < cnt; i += gridDim.x * 4) {
    const uint32_t c = mid[i];
    const FCell cell = cells[c];
    if (MP) mp_direct<Q, 64>(c, cell.first, cell.count, threadIdx.x & 63, cell.cx, cell.cy, cell.r, ux, uy, ps, mexp);
    else geom_direct<64>(c, cell.first, cell.count, threadIdx.x & 63, cells, bbox, ux, uy);
  }
}

template <int Q, bool MP>
__global__ __launch_bounds__(1024) void k_fmm_up_big(FCell* __restrict__ cells, double* __restrict__ bbox,
                                                     const double* __restrict__ ux, const double* __restrict__ uy,
                                                     const float* __restrict__ ps, float2* __restrict__ mexp,
                                                     const uint32_t* __restrict__ big, uint32_t big_cap,
                                                     const FmmParams* __restrict__ fp) {
  __shared__ uint32_t cnt[FMM_LEVELS + 1];
  if (fp->abort) return;
  if (threadIdx.x <= FMM_LEVELS) cnt[threadIdx.x] = min(fp->big_n[threadIdx.x], big_cap);
  __syncthreads();
  for (int d = (int)FMM_LEVELS; d >= 0; --d) {
    if (cnt[d] == 0) continue;
    for (uint32_t i = threadIdx.x; i < cnt[d]; i += 1024) {
      const uint32_t c = big[(size_t)d * big_cap + i];
      if (MP) mp_cell<Q>(c, cells, ux, uy, ps, mexp);
      else geom_cell(c, cells, bbox, ux, uy);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- dual tree

struct MacArgs {
  double mac_far;  // 2 kappa - 1 (fmm.zig:266)
  uint32_t nbits;
  const FmmGraph* graphs;
};

constexpr int STEP_PER = 4;                   // pairs per thread per tile
constexpr int STEP_TILE = EWF * STEP_PER;     // pairs per block per tile

// Inclusive scan of v over the block (EWF threads); wsum holds EWF / 64 words.
__device__ inline uint32_t block_inclusive(uint32_t v, uint32_t* wsum, uint32_t& total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = __shfl_up(inc, d, 64);
    if (lane >= d) inc += t;
  }
  if (lane == 63) wsum[wave] = inc;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < EWF / 64; ++w) {
    const uint32_t x = wsum[w];
    if (w < wave) before += x;
    total += x;
  }
  __syncthreads();
  return before + inc;
}

// fmm.zig:400-453 for one pair (A, B): kind 0 none, 1 M2L, 2 P2P; push 0
// none, 1 children x children, 2 (A, children of B), 3 (children of A, B),
// the children being c0 .. c0 + nch; pairs counts the P2P point pairs.
__device__ __forceinline__ void dual_pair(const FCell* __restrict__ cells, uint32_t A, uint32_t B, const MacArgs& mac,
                                          uint32_t& kind, uint32_t& push, uint32_t& c0, uint32_t& nch,
                                          unsigned long long& pairs) {
  const FCell ca = cells[A];
  if (A == B) {
    if (ca.nchild == 0) {
      kind = 2;
      pairs += (unsigned long long)ca.count * ca.count;
    } else {
      push = 1;
      c0 = ca.child0;
      nch = ca.nchild;
    }
    return;
  }
  const FCell cb = cells[B];
  const double ra = ca.r, rb = cb.r;
  const double dx = ca.cx - cb.cx, dy = ca.cy - cb.cy;
  const double d2 = dx * dx + dy * dy;
  const double rmax = zmax(ra, rb), rmin = zmin(ra, rb);
  const double thr = mac.mac_far * rmax + rmin;
  bool done = false;
  if (d2 > MIN_SEP2 && d2 >= thr * thr) {
    bool separated = true;
    const FmmGraph& G = mac.graphs[ca.graph];
    if (G.ring) {
      const double gap = __builtin_sqrt(d2) - ra - rb;
      if (gap >= G.dmax_s) done = true;
      else if (gap + 2 * (ra + rb) > G.dmax_s || gap < G.dmin_s) separated = false;
    }
    if (!done && separated) {
      kind = 1;
      done = true;
    }
  }
  if (!done) {
    const bool la = ca.nchild == 0, lb = cb.nchild == 0;
    if (la && lb) {
      kind = 2;
      pairs += (unsigned long long)ca.count * cb.count;
    } else if (la || (!lb && rb > ra)) {
      push = 2;
      c0 = cb.child0;
      nch = cb.nchild;
    } else {
      push = 3;
      c0 = ca.child0;
      nch = ca.nchild;
    }
  }
}

// One round of fmm.zig:400-453 over the frontier of round step (its size at fp->front[step * FRONT_STRIDE]). A block
// classifies a tile of STEP_TILE pairs, scans its output counts and reserves
// room in each list with one atomic per list and tile. A frontier that
// overflowed its buffer voids the pass (it is redone after the buffers grow,
// or the tick aborts): its counts include pairs that were never written, so
// the rounds after it read nothing.
__global__ __launch_bounds__(EWF) void k_dual_step(const FCell* __restrict__ cells, FmmParams* __restrict__ fp,
                                                   uint32_t step, const uint2* __restrict__ in,
                                                   uint2* __restrict__ out, uint32_t front_cap,
                                                   uint64_t* __restrict__ m2l, uint32_t m2l_cap,
                                                   uint64_t* __restrict__ p2p, uint32_t p2p_cap,
                                                   uint32_t* __restrict__ m2l_cnt, uint32_t* __restrict__ p2p_cnt,
                                                   MacArgs mac) {
  __shared__ uint32_t wsum[3][EWF / 64];
  __shared__ uint32_t base_sh[3];
  __shared__ unsigned long long pairs_sh[EWF / 64];
  if (fp->abort || (fp->list_overflow & 1u)) return;
  uint32_t cnt = fp->front[step * FRONT_STRIDE];
  if (cnt > front_cap) cnt = front_cap;
  unsigned long long block_pairs = 0;  // thread 0: the block's P2P point pairs, added once at the end
  for (uint32_t tile = blockIdx.x * STEP_TILE; tile < cnt; tile += gridDim.x * STEP_TILE) {
    // kind: 0 none, 1 M2L, 2 P2P; push: 0 none, 1 children x children,
    // 2 (A, children of B), 3 (children of A, B)
    uint32_t pa[STEP_PER], pb[STEP_PER], pc0[STEP_PER], code[STEP_PER];
    uint32_t n_front = 0, n_m2l = 0, n_p2p = 0;
    unsigned long long pairs = 0;
#pragma unroll
    for (int q = 0; q < STEP_PER; ++q) {
      const uint32_t idx = tile + q * EWF + threadIdx.x;
      uint32_t A = 0, B = 0, kind = 0, push = 0, c0 = 0, nch = 0;
      if (idx < cnt) {
        const uint2 pr = in[idx];
        A = pr.x;
        B = pr.y;
        dual_pair(cells, A, B, mac, kind, push, c0, nch, pairs);
      }
      pa[q] = A;
      pb[q] = B;
      pc0[q] = c0;
      code[q] = kind | push << 2 | nch << 4;
      n_front += push == 1 ? nch * nch : (push ? nch : 0);
      n_m2l += kind == 1;
      n_p2p += kind == 2;
    }
    uint32_t t_front, t_m2l, t_p2p;
    uint32_t o_front = block_inclusive(n_front, wsum[0], t_front) - n_front;
    uint32_t o_m2l = block_inclusive(n_m2l, wsum[1], t_m2l) - n_m2l;
    uint32_t o_p2p = block_inclusive(n_p2p, wsum[2], t_p2p) - n_p2p;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) pairs += __shfl_xor(pairs, d, 64);
    if ((threadIdx.x & 63) == 0) pairs_sh[threadIdx.x >> 6] = pairs;
    if (threadIdx.x == 0) {
      base_sh[0] = t_front ? atomicAdd(&fp->front[(step + 1) * FRONT_STRIDE], t_front) : 0;
      base_sh[1] = t_m2l ? atomicAdd(&fp->m2l, t_m2l) : 0;
      base_sh[2] = t_p2p ? atomicAdd(&fp->p2p, t_p2p) : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int w = 0; w < EWF / 64; ++w) block_pairs += pairs_sh[w];
    o_front += base_sh[0];
    o_m2l += base_sh[1];
    o_p2p += base_sh[2];
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < STEP_PER; ++q) {
      const uint32_t kind = code[q] & 3, push = (code[q] >> 2) & 3, nch = code[q] >> 4;
      const uint32_t A = pa[q], B = pb[q], c0 = pc0[q];
      const uint64_t key = (uint64_t)A << mac.nbits | B;
      // Per-target counts, one atomic per run of lanes holding the same A
      // (the pairs a split pushes are adjacent, so such runs are common).
      const uint32_t prev = __shfl_up(A, 1, 64);
      const bool head = lane == 0 || prev != A;
      const uint64_t heads = __ballot(head), is_m2l = __ballot(kind == 1), is_p2p = __ballot(kind == 2);
      if (head) {
        const uint64_t above = heads & ~((2ull << lane) - 1ull);  // heads after this lane
        const uint64_t upto = above ? (above & (0ull - above)) - 1ull : ~0ull;  // lanes before the next head
        const uint64_t run = upto & ~((1ull << lane) - 1ull);
        const uint32_t nm = (uint32_t)__popcll(is_m2l & run), np = (uint32_t)__popcll(is_p2p & run);
        if (nm) atomicAdd(&m2l_cnt[A], nm);
        if (np) atomicAdd(&p2p_cnt[A], np);
      }
      if (kind == 1) {
        if (o_m2l < m2l_cap) m2l[o_m2l] = key;
        else atomicOr(&fp->list_overflow, 2u);
        ++o_m2l;
      } else if (kind == 2) {
        if (o_p2p < p2p_cap) p2p[o_p2p] = key;
        else atomicOr(&fp->list_overflow, 4u);
        ++o_p2p;
      }
      if (push) {
        const uint32_t np = push == 1 ? nch * nch : nch;
        if (o_front + np > front_cap) {
          atomicOr(&fp->list_overflow, 1u);
        } else if (push == 1) {
          for (uint32_t i = 0; i < nch; ++i)
            for (uint32_t j = 0; j < nch; ++j) out[o_front + i * nch + j] = make_uint2(c0 + i, c0 + j);
        } else if (push == 2) {
          for (uint32_t j = 0; j < nch; ++j) out[o_front + j] = make_uint2(A, c0 + j);
        } else {
          for (uint32_t i = 0; i < nch; ++i) out[o_front + i] = make_uint2(c0 + i, B);
        }
        o_front += np;
      }
    }
  }
  if (threadIdx.x == 0 && block_pairs) atomicAdd(&fp->pairs, block_pairs);
}

// The first frontier: every root paired with itself; the list counters
// cleared.
__global__ __launch_bounds__(EWF) void k_fmm_lists_reset(FmmParams* __restrict__ fp, uint2* __restrict__ front,
                                                         uint32_t front_cap) {
  if (fp->abort) return;
  const uint32_t roots = fp->roots <= front_cap ? fp->roots : 0;
  const uint32_t t = blockIdx.x * EWF + threadIdx.x;
  if (t < roots) front[t] = make_uint2(t, t);
  if (blockIdx.x == 0) {
    if (threadIdx.x < FMM_MAX_STEPS + 2) fp->front[threadIdx.x * FRONT_STRIDE] = threadIdx.x == 0 ? roots : 0u;
    if (threadIdx.x == 0) {
      fp->m2l = 0;
      fp->p2p = 0;
      fp->list_overflow = 0;
      fp->long_segs = 0;
      fp->pairs = 0;
      fp->near_items = 0;
      fp->near_slots = 0;
      fp->near_combines = 0;
    }
  }
}

// Interaction pairs grouped by target: the dual-tree rounds count each
// target's pairs, a scan places the groups, and a scatter through per-target
// cursors fills them. Within a target the scatter's order is arbitrary; the
// kernels that read a group first order it by source (sort_segment), so
// every sum runs in one fixed order.
__global__ __launch_bounds__(EWF) void k_target_scatter(const uint64_t* __restrict__ list,
                                                        const uint32_t* __restrict__ count, uint32_t nbits,
                                                        const uint32_t* __restrict__ off, uint32_t* __restrict__ cursor,
                                                        uint32_t* __restrict__ src, const FmmParams* __restrict__ fp) {
  if (fp->abort) return;
  // Consecutive keys often share a target: one cursor atomic per run of
  // lanes with the same target, each lane placed by its rank in the run.
  const uint32_t n = *count, lane = threadIdx.x & 63;
  for (uint32_t k0 = blockIdx.x * EWF; k0 < n; k0 += gridDim.x * EWF) {
    const uint32_t k = k0 + threadIdx.x;
    const bool valid = k < n;
    const uint64_t key = valid ? list[k] : 0;
    const uint32_t a = valid ? (uint32_t)(key >> nbits) : 0xFFFFFFFFu;  // no cell has that index
    const uint32_t prev = __shfl_up(a, 1, 64);
    const uint64_t heads = __ballot(lane == 0 || prev != a);
    const uint64_t below = (2ull << lane) - 1ull;  // lanes 0 .. lane (all 64 at lane 63)
    const uint32_t head = 63u - (uint32_t)__builtin_clzll(heads & below);
    const uint64_t above = heads & ~below;
    uint32_t base = 0;
    if (valid && head == lane) base = atomicAdd(&cursor[a], (above ? (uint32_t)__builtin_ctzll(above) : 64u) - lane);
    base = __shfl(base, (int)head, 64);
    if (valid) src[off[a] + base + (lane - head)] = (uint32_t)(key & (((uint64_t)1 << nbits) - 1));
  }
}

// One wave loads the sources of window w (SEG_MAX at a time) of a target
// with k sources into out[0 .. min(SEG_MAX, k - w)): ordered by source when
// the target has at most SEG_MAX of them, in scatter order otherwise.
__device__ inline uint32_t load_window(const uint32_t* __restrict__ src, uint32_t off, uint32_t k, uint32_t w,
                                       uint32_t (&in)[SEG_MAX], uint32_t (&out)[SEG_MAX]);

// One wave orders the k <= SEG_MAX distinct sources src[off .. off + k) by
// rank into out[0 .. k).
template <uint32_t CAP>
__device__ inline void sort_segment(const uint32_t* __restrict__ src, uint32_t off, uint32_t k, uint32_t (&in)[CAP],
                                    uint32_t (&out)[CAP]) {
  const int lane = threadIdx.x & 63;
  for (uint32_t q = lane; q < k; q += 64) in[q] = src[off + q];
  __syncthreads();
  for (uint32_t q = lane; q < k; q += 64) {
    const uint32_t b = in[q];
    uint32_t r = 0;
    for (uint32_t t = 0; t < k; ++t) r += in[t] < b;
    out[r] = b;
  }
  __syncthreads();
}

__device__ inline uint32_t load_window(const uint32_t* __restrict__ src, uint32_t off, uint32_t k, uint32_t w,
                                       uint32_t (&in)[SEG_MAX], uint32_t (&out)[SEG_MAX]) {
  if (k <= SEG_MAX) {
    sort_segment(src, off, k, in, out);
    return k;
  }
  const uint32_t wk = k - w < SEG_MAX ? k - w : SEG_MAX;
  __syncthreads();
  for (uint32_t q = threadIdx.x & 63; q < wk; q += 64) out[q] = src[off + w + q];
  __syncthreads();
  return wk;
}

// ---------------------------------------------------------------- M2L

// M2L (fmm.zig:561-610) on the FP32 matrix cores. With w = 1/z0,
// a = rho_B w and b = rho_A w, the radius-scaled translation is
//   L^_l(A) += (-1)^l b^l w sum_k C(k + l, k) a^k M^_k(B),
// i.e. U_k = a^k M^_k (per pair), V = P U with the constant Pascal matrix
// P[l][k] = C(k + l, k), then L^_l += (-1)^l b^l w V_l. The middle step is
// one batched GEMM whose left operand every pair shares: one wave per target
// takes its sources (ordered by source) 16 at a time as the 16 columns of
// v_mfma_f32_16x16x4f32 (real and imaginary parts in two products), rows
// 16 t + i of P in row tiles of 16 and k in steps of 4. Lane L supplies
// P[16 t + L % 16][4 s + L / 16] and U of pair L % 16 at k = 4 s + L / 16,
// and receives V rows 16 t + 4 (L / 16) + r of pair L % 16, so the
// post-scaling uses the lane's own pair. Each lane sums its pairs in source
// order; the 16 lanes of a row group are then added by a fixed butterfly.
template <int Q>
__device__ __forceinline__ void m2l_target(uint32_t A, const FCell* __restrict__ cells,
                                           const uint32_t* __restrict__ srcs, const uint32_t* __restrict__ offs,
                                           const uint32_t* __restrict__ cnts, const float2* __restrict__ mexp,
                                           float2* __restrict__ lexp) {
  constexpr int RT = (Q + 15) / 16;  // row tiles of local coefficients
  constexpr int KS = (Q + 3) / 4;    // k steps of multipole coefficients
  __shared__ uint32_t seg_in[SEG_MAX], seg[SEG_MAX];
  const int lane = threadIdx.x;
  const int col = lane & 15, kq = lane >> 4;
  const uint32_t K = cnts[A], off = offs[A];
  float pa[RT][KS];
#pragma unroll
  for (int t = 0; t < RT; ++t)
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int l = 16 * t + col, k = 4 * s + kq;
      pa[t][s] = (l < Q && k < Q) ? binom(k + l, k) : 0.f;
    }
  f32x4 accr[RT], acci[RT];
#pragma unroll
  for (int t = 0; t < RT; ++t) {
    accr[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    acci[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  if (K > 0) {
    const FCell ca = cells[A];
    for (uint32_t w0 = 0; w0 < K; w0 += SEG_MAX) {
      const uint32_t wk = load_window(srcs, off, K, w0, seg_in, seg);
      for (uint32_t base = 0; base < wk; base += 16) {
        const uint32_t sp = base + col;
        float2 w = make_float2(0.f, 0.f), a = w, b = w;
        uint32_t B = 0;
        if (sp < wk) {
          B = seg[sp];
          const FCell cb = cells[B];
          const double dx = ca.cx - cb.cx, dy = ca.cy - cb.cy;
          const double inv = 1 / (dx * dx + dy * dy);
          const double wr = dx * inv, wi = -dy * inv;
          w = make_float2((float)wr, (float)wi);
          a = make_float2((float)(cb.r * wr), (float)(cb.r * wi));
          b = make_float2((float)(ca.r * wr), (float)(ca.r * wi));
        }
        const bool valid = sp < wk;
        float2 ak = make_float2(1.f, 0.f);
        for (int q = 0; q < kq; ++q) ak = cmul(ak, a);
        const float2 a2 = cmul(a, a), a4 = cmul(a2, a2);
        f32x4 vr[RT], vi[RT];
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          vr[t] = f32x4{0.f, 0.f, 0.f, 0.f};
          vi[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const int k = 4 * s + kq;
          const float2 u = (valid && k < Q) ? cmul(mexp[(size_t)B * Q + k], ak) : make_float2(0.f, 0.f);
          ak = cmul(ak, a4);
#pragma unroll
          for (int t = 0; t < RT; ++t) {
            vr[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[t][s], u.x, vr[t], 0, 0, 0);
            vi[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[t][s], u.y, vi[t], 0, 0, 0);
          }
        }
        // coefficient (-b)^l w for rows l = 16 t + 4 kq + r
        const float2 nb = make_float2(-b.x, -b.y);
        const float2 nb2 = cmul(nb, nb), nb4 = cmul(nb2, nb2);
        const float2 nb8 = cmul(nb4, nb4), nb16 = cmul(nb8, nb8);
        float2 c = w;
        for (int q = 0; q < kq; ++q) c = cmul(c, nb4);
#pragma unroll
        for (int t = 0; t < RT; ++t) {
          float2 cr = c;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            accr[t][r] = fmaf(cr.x, vr[t][r], fmaf(-cr.y, vi[t][r], accr[t][r]));
            acci[t][r] = fmaf(cr.x, vi[t][r], fmaf(cr.y, vr[t][r], acci[t][r]));
            cr = cmul(cr, nb);
          }
          c = cmul(c, nb16);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int d = 8; d > 0; d >>= 1) {
          accr[t][r] += __shfl_xor(accr[t][r], d, 64);
          acci[t][r] += __shfl_xor(acci[t][r], d, 64);
        }
  }
  if (col == 0) {
#pragma unroll
    for (int t = 0; t < RT; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int l = 16 * t + 4 * kq + r;
        if (l < Q) lexp[(size_t)A * Q + l] = make_float2(accr[t][r], acci[t][r]);
      }
  }
}

template <int Q>
__global__ __launch_bounds__(64) void k_m2l(const FCell* __restrict__ cells, const uint32_t* __restrict__ srcs,
                                            const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cnts,
                                            const float2* __restrict__ mexp, float2* __restrict__ lexp,
                                            const FmmParams* __restrict__ fp) {
  if (fp->abort) return;
  for (uint32_t A = blockIdx.x; A < fp->ncg; A += gridDim.x) m2l_target<Q>(A, cells, srcs, offs, cnts, mexp, lexp);
}

// ---------------------------------------------------------------- downward

// L2L (fmm.zig:823-849) straight from every ancestor into each leaf, one
// thread per leaf: the local expansion a leaf evaluates is its own M2L sum
// plus each ancestor's M2L sum shifted to the leaf's centre, which is what
// the reference's level-by-level L2L chain adds up to (shifting a polynomial
// is exact). With d the offset of the leaf's centre from the ancestor's,
//   L^_m(leaf) += (rho_leaf / rho_A)^m sum_{l >= m} C(l, m) L^_l(A) (d / rho_A)^(l-m).
// Only leaves are written, so the ancestors' sums are read unchanged.
template <int Q>
__global__ __launch_bounds__(EWF) void k_fmm_down_leaf(const FCell* __restrict__ cells,
                                                       const uint32_t* __restrict__ parent,
                                                       const uint32_t* __restrict__ m2l_cnt,
                                                       float2* __restrict__ lexp, const FmmParams* __restrict__ fp) {
  if (fp->abort) return;
  const uint32_t nc = fp->ncg;
  for (uint32_t c = blockIdx.x * EWF + threadIdx.x; c < nc; c += gridDim.x * EWF) {
    const FCell cc = cells[c];
    if (cc.nchild != 0) continue;
    float2 acc[Q];
#pragma unroll
    for (int k = 0; k < Q; ++k) acc[k] = make_float2(0.f, 0.f);
    bool any = false;
    for (uint32_t a = parent[c]; a != NONE; a = parent[a]) {
      if (m2l_cnt[a] == 0) continue;  // no M2L sum of its own
      any = true;
      const FCell ca = cells[a];
      const double inv = ca.r > 0 ? 1 / ca.r : 0;
      const float tr = (float)(cc.r * inv);
      const float2 e = make_float2((float)((cc.cx - ca.cx) * inv), (float)((cc.cy - ca.cy) * inv));
      float2 lp[Q], ep[Q];
#pragma unroll
      for (int l = 0; l < Q; ++l) lp[l] = lexp[(size_t)a * Q + l];
      ep[0] = make_float2(1.f, 0.f);
#pragma unroll
      for (int k = 1; k < Q; ++k) ep[k] = cmul(ep[k - 1], e);
      float tm = 1.f;
#pragma unroll
      for (int mm = 0; mm < Q; ++mm) {
        float2 s = make_float2(0.f, 0.f);
#pragma unroll
        for (int l = mm; l < Q; ++l) s = cfma(cscale(lp[l], binom(l, mm)), ep[l - mm], s);
        acc[mm].x = fmaf(s.x, tm, acc[mm].x);
        acc[mm].y = fmaf(s.y, tm, acc[mm].y);
        tm *= tr;
      }
    }
    if (!any) continue;
#pragma unroll
    for (int k = 0; k < Q; ++k) {
      float2 v = lexp[(size_t)c * Q + k];
      v.x += acc[k].x;
      v.y += acc[k].y;
      lexp[(size_t)c * Q + k] = v;
    }
  }
}

// ---------------------------------------------------------------- near field + L2P

// L2P (fmm.zig:853-886) at one target and the FP64 velocity update with its
// near-field sum.
template <int Q>
__device__ __forceinline__ void near_finish(const FCell& ca, uint32_t A, uint32_t ti, float2 near,
                                            const float2* __restrict__ lexp, const double* __restrict__ ux,
                                            const double* __restrict__ uy, const uint32_t* __restrict__ perm,
                                            double* __restrict__ vxs, double* __restrict__ vys, double inv_w) {
  const double inv_r = ca.r > 0 ? 1 / ca.r : 0;
  const float2 z = make_float2((float)((ux[ti] - ca.cx) * inv_r), (float)((uy[ti] - ca.cy) * inv_r));
  float2 f = lexp[(size_t)A * Q + Q - 1];
#pragma unroll
  for (int l = Q - 2; l >= 0; --l) f = cfma(f, z, lexp[(size_t)A * Q + l]);
  const uint32_t i = perm[ti];
  vxs[i] += (double)near.x - (double)f.x * inv_w;
  vys[i] += (double)near.y + (double)f.y * inv_w;
}

// Near field of up to 64 targets of leaf A, starting at target tb, on one
// wave. A target's sources (its P2P segment, ordered by source leaf) are cut
// into chunks of NEAR_CHUNK leaves; within a chunk, S lanes per target split
// the source points and their sums are added in lane order; the chunk sums
// are added in chunk order. chunk == NEAR_ALL runs every chunk and finishes
// the targets (L2P, velocity); otherwise the one chunk's sums go to `out`
// (one float2 per target) and k_near_combine adds them later in the same
// order, so a split target's result is bit-identical to an unsplit one.
// Each source leaf is staged in LDS as FP32 offsets from an anchor (the
// target leaf's first point, FP64), so two points with equal FP64
// coordinates have an exactly zero FP32 delta. The pair term is d3's
// (fmm.zig:529-553): skip beyond dmax, hash jiggle on a zero delta,
// sqrt(dmin2 * l) below dmin.
template <int Q>
__device__ __forceinline__ void near_batch(uint32_t A, uint32_t tb, uint32_t chunk, float2* __restrict__ out,
                                           const FCell* __restrict__ cells, const uint32_t* __restrict__ srcs,
                                           const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cnts,
                                           const float2* __restrict__ lexp, const double* __restrict__ px,
                                           const double* __restrict__ py, const float* __restrict__ ps,
                                           const double* __restrict__ ux, const double* __restrict__ uy,
                                           const uint32_t* __restrict__ perm, double* __restrict__ vxs,
                                           double* __restrict__ vys, const FmmGraph* __restrict__ gp, float dmin2,
                                           float dmax2) {
  const FCell ca = cells[A];
  const int lane = threadIdx.x;
  // at most NEAR_CHUNK sources at a time: a longer segment is ordered in
  // place beforehand and read chunk by chunk
  __shared__ uint32_t seg_in[NEAR_CHUNK], seg[NEAR_CHUNK];
  __shared__ float4 src[64];
  __shared__ float2 part[64];
  const uint32_t K = cnts[A], off = offs[A];
  const uint32_t nch = K > NEAR_CHUNK ? (K + NEAR_CHUNK - 1) / NEAR_CHUNK : 1;
  const uint32_t c0 = chunk == NEAR_ALL ? 0 : chunk, c1 = chunk == NEAR_ALL ? nch : chunk + 1;
  const double ax = px[ca.first], ay = py[ca.first];
  const uint64_t seed = gp[ca.graph].seed;
  const uint32_t base = gp[ca.graph].node_base;  // the hash takes indices within the graph
  const uint32_t T = ca.count - tb < 64 ? ca.count - tb : 64;
  const uint32_t S = 64 / T;
  const uint32_t t = lane / S, sub = lane % S;
  const bool active = t < T;
  const uint32_t ti = ca.first + tb + (active ? t : 0);
  const float xi = (float)(px[ti] - ax), yi = (float)(py[ti] - ay);
  const uint32_t il = perm[ti] - base;
  float2 total = make_float2(0.f, 0.f);
  for (uint32_t c = c0; c < c1; ++c) {
    uint32_t wk;
    if (K <= NEAR_CHUNK) {
      sort_segment(srcs, off, K, seg_in, seg);
      wk = K;
    } else {  // ordered in place by k_sort_long
      wk = K - c * NEAR_CHUNK < NEAR_CHUNK ? K - c * NEAR_CHUNK : NEAR_CHUNK;
      __syncthreads();
      if ((uint32_t)lane < wk) seg[lane] = srcs[off + c * NEAR_CHUNK + lane];
      __syncthreads();
    }
    float accx = 0.f, accy = 0.f;
    // The sources in 64-point steps (leaf s of the segment, points p0 ..):
    // each step's points are loaded into registers while the step before it
    // is computed, then staged in LDS.
    uint32_t s = 0, p0 = 0, f0 = 0, cnt = 0;
    if (wk) {
      const uint32_t B = seg[0];
      f0 = cells[B].first;
      cnt = cells[B].count;
    }
    double nx = 0, ny = 0;
    float nsr = 0.f;
    if (wk && (uint32_t)lane < cnt) {
      nx = px[f0 + lane];
      ny = py[f0 + lane];
      nsr = ps[f0 + lane];
    }
    while (s < wk) {
      const uint32_t np = cnt - p0 < 64 ? cnt - p0 : 64;
      if ((uint32_t)lane < np)
        src[lane] = make_float4((float)(nx - ax), (float)(ny - ay), nsr, __uint_as_float(f0 + p0 + lane));
      // the next step, its loads in flight during this one
      uint32_t s2 = s, q0 = p0 + 64, g0 = f0, cnt2 = cnt;
      if (q0 >= cnt) {
        s2 = s + 1;
        q0 = 0;
        if (s2 < wk) {
          const uint32_t B = seg[s2];
          g0 = cells[B].first;
          cnt2 = cells[B].count;
        }
      }
      if (s2 < wk && q0 + (uint32_t)lane < cnt2) {
        nx = px[g0 + q0 + lane];
        ny = py[g0 + q0 + lane];
        nsr = ps[g0 + q0 + lane];
      }
      {
        __syncthreads();
        if (active) {
          // Branch-free pair step: a skipped pair (the target itself, or
          // beyond distanceMax) adds dx * 0; the rare jiggle case runs only
          // when a lane of the wave needs it; distanceMin's square root is
          // the hardware one (FP32 like the rest of the near field).
          for (uint32_t q = sub; q < np; q += S) {
            const float4 cq = src[q];
            const uint32_t j = f0 + p0 + q;
            float dx = cq.x - xi, dy = cq.y - yi;
            float l = fmaf(dx, dx, dy * dy);
            const bool ok = j != ti && l < dmax2;
            if (__ballot(ok && (dx == 0.f || dy == 0.f)) != 0) {
              if (ok && (dx == 0.f || dy == 0.f)) {
                const uint32_t jl = perm[j] - base;
                if (dx == 0.f) {
                  dx = (float)hash_jiggle(seed, il, jl, 0);
                  l = fmaf(dx, dx, l);
                }
                if (dy == 0.f) {
                  dy = (float)hash_jiggle(seed, il, jl, 1);
                  l = fmaf(dy, dy, l);
                }
              }
            }
            if (l < dmin2) l = __builtin_amdgcn_sqrtf(dmin2 * l);
            const float wj = ok ? cq.z * __builtin_amdgcn_rcpf(l) : 0.f;
            accx = fmaf(dx, wj, accx);
            accy = fmaf(dy, wj, accy);
          }
        }
        __syncthreads();
      }
      s = s2;
      p0 = q0;
      f0 = g0;
      cnt = cnt2;
    }
    part[lane] = make_float2(accx, accy);
    __syncthreads();
    if (active && sub == 0) {
      float2 p = part[lane];
      for (uint32_t q = 1; q < S; ++q) {
        p.x += part[lane + q].x;
        p.y += part[lane + q].y;
      }
      if (c == c0) {
        total = p;
      } else {
        total.x += p.x;
        total.y += p.y;
      }
    }
    __syncthreads();
  }
  if (!(active && sub == 0)) return;
  if (chunk != NEAR_ALL) {
    out[t] = total;
    return;
  }
  near_finish<Q>(ca, A, ti, total, lexp, ux, uy, perm, vxs, vys, gp[ca.graph].inv_w);
}

// Targets of one list whose segments are longer than the in-LDS sort takes
// (M2L: SEG_MAX) or longer than one near-field chunk (P2P).
__global__ __launch_bounds__(EWF) void k_long_list(const uint32_t* __restrict__ ncp,
                                                   const uint32_t* __restrict__ m2l_cnt,
                                                   const uint32_t* __restrict__ p2p_cnt, FmmParams* __restrict__ fp,
                                                   uint32_t* __restrict__ list) {
  if (fp->abort) return;
  for (uint32_t c = blockIdx.x * EWF + threadIdx.x; c < *ncp; c += gridDim.x * EWF) {
    if (p2p_cnt[c] > NEAR_CHUNK) list[atomicAdd(&fp->long_segs, 1u)] = c | 0x80000000u;
    if (m2l_cnt[c] > SEG_MAX) list[atomicAdd(&fp->long_segs, 1u)] = c;
  }
}

// Orders each long segment by source in place. A target's sources are
// distinct cell ids, so a bitmap over the ids orders them: the segment is
// copied aside (into the pair list, which is no longer read), its ids are
// set in an LDS bitmap covering LONG_IDS ids per pass, and the set bits are
// written back in id order (each thread owns 32 consecutive words; a block
// scan of their counts gives the output positions).
__global__ __launch_bounds__(256) void k_sort_long(const uint32_t* __restrict__ list,
                                                   const FmmParams* __restrict__ fp, const uint32_t* __restrict__ ncp,
                                                   uint32_t* __restrict__ m2l_src, const uint32_t* __restrict__ m2l_off,
                                                   const uint32_t* __restrict__ m2l_cnt, uint32_t* __restrict__ m2l_tmp,
                                                   uint32_t* __restrict__ p2p_src, const uint32_t* __restrict__ p2p_off,
                                                   const uint32_t* __restrict__ p2p_cnt,
                                                   uint32_t* __restrict__ p2p_tmp) {
  constexpr uint32_t WORDS = LONG_IDS / 32, PER = WORDS / 256;
  __shared__ uint32_t bits[WORDS];
  __shared__ uint32_t scan[256];
  const uint32_t tid = threadIdx.x;
  if (fp->abort) return;
  const uint32_t nseg = fp->long_segs, nc = *ncp;
  for (uint32_t k = blockIdx.x; k < nseg; k += gridDim.x) {
    const uint32_t e = list[k];
    const bool near = e >> 31;
    const uint32_t A = e & 0x7FFFFFFFu;
    uint32_t* src = near ? p2p_src : m2l_src;
    uint32_t* tmp = near ? p2p_tmp : m2l_tmp;
    const uint32_t off = (near ? p2p_off : m2l_off)[A], K = (near ? p2p_cnt : m2l_cnt)[A];
    __syncthreads();
    for (uint32_t q = tid; q < K; q += 256) tmp[off + q] = src[off + q];
    __syncthreads();
    uint32_t done = 0;
    for (uint32_t lo = 0; lo < nc; lo += LONG_IDS) {
      for (uint32_t w = tid; w < WORDS; w += 256) bits[w] = 0;
      __syncthreads();
      for (uint32_t q = tid; q < K; q += 256) {
        const uint32_t b = tmp[off + q] - lo;
        if (b < LONG_IDS) atomicOr(&bits[b >> 5], 1u << (b & 31));
      }
      __syncthreads();
      uint32_t mine = 0;
      for (uint32_t w = 0; w < PER; ++w) mine += __popc(bits[tid * PER + w]);
      scan[tid] = mine;
      __syncthreads();
      for (uint32_t d = 1; d < 256; d <<= 1) {
        const uint32_t v = tid >= d ? scan[tid - d] : 0u;
        __syncthreads();
        scan[tid] += v;
        __syncthreads();
      }
      uint32_t pos = off + done + scan[tid] - mine;
      for (uint32_t w = 0; w < PER; ++w) {
        uint32_t word = bits[tid * PER + w];
        while (word) {
          const uint32_t bit = __ffs(word) - 1;
          word &= word - 1;
          src[pos++] = lo + (tid * PER + w) * 32 + bit;
        }
      }
      done += scan[255];
      __syncthreads();
    }
  }
}

// Near-field work beyond each leaf's first 64 targets and its first
// unsplit pass: for every (leaf, target batch), either one item that runs
// every chunk, or (segments of several chunks, while partial-sum slots last)
// one item per chunk plus a combine. split0[leaf] says whether the leaf's
// first batch was split (its own block then skips it).
__device__ inline void near_items_cell(uint32_t c, const FCell* __restrict__ cells, const uint32_t* __restrict__ cnts,
                                       FmmParams* __restrict__ fp, uint4* __restrict__ items,
                                       uint4* __restrict__ combines, uint32_t* __restrict__ split0,
                                       uint32_t slot_cap) {
  const FCell cell = cells[c];
  if (cell.nchild != 0) return;
  const uint32_t K = cnts[c];
  const uint32_t nch = K > NEAR_CHUNK ? (K + NEAR_CHUNK - 1) / NEAR_CHUNK : 1;
  uint32_t first = 0;
  for (uint32_t tb = 0; tb < cell.count; tb += 64) {
    bool split = false;
    if (nch > 1) {
      const uint32_t s0 = atomicAdd(&fp->near_slots, nch);
      if (s0 <= slot_cap && nch <= slot_cap - s0) {
        const uint32_t k = atomicAdd(&fp->near_items, nch);
        for (uint32_t q = 0; q < nch; ++q) items[k + q] = make_uint4(c, tb, q, s0 + q);
        combines[atomicAdd(&fp->near_combines, 1u)] = make_uint4(c, tb, s0, nch);
        split = true;
      }
    }
    if (tb == 0)
      first = split;
    else if (!split)
      items[atomicAdd(&fp->near_items, 1u)] = make_uint4(c, tb, NEAR_ALL, 0);
  }
  split0[c] = first;
}

__global__ __launch_bounds__(EWF) void k_near_items(const FCell* __restrict__ cells, const uint32_t* __restrict__ ncp,
                                                    const uint32_t* __restrict__ cnts, FmmParams* __restrict__ fp,
                                                    uint4* __restrict__ items, uint4* __restrict__ combines,
                                                    uint32_t* __restrict__ split0, uint32_t slot_cap) {
  if (fp->abort) return;
  for (uint32_t c = blockIdx.x * EWF + threadIdx.x; c < *ncp; c += gridDim.x * EWF)
    near_items_cell(c, cells, cnts, fp, items, combines, split0, slot_cap);
}


// Blocks [0, extra) share the items (launched first, so the longest work
// starts first); the other blocks run the first 64 targets of each leaf
// (blocks extra + c, extra + c + (gridDim - extra), ...) unless that batch
// was split into chunk items.
template <int Q>
__global__ __launch_bounds__(64) void k_near(const FCell* __restrict__ cells, const uint32_t* __restrict__ srcs,
                                             const uint32_t* __restrict__ offs, const uint32_t* __restrict__ cnts,
                                             const FmmParams* __restrict__ fp, const float2* __restrict__ lexp,
                                             const double* __restrict__ px, const double* __restrict__ py,
                                             const float* __restrict__ ps, const double* __restrict__ ux,
                                             const double* __restrict__ uy, const uint32_t* __restrict__ perm,
                                             double* __restrict__ vxs, double* __restrict__ vys,
                                             const FmmGraph* __restrict__ gp, float dmin2, float dmax2,
                                             const uint4* __restrict__ items, const uint32_t* __restrict__ split0,
                                             float2* __restrict__ partial, uint32_t extra) {
  if (fp->abort) return;
  if (blockIdx.x < extra) {
    const uint32_t count = fp->near_items;
    for (uint32_t k = blockIdx.x; k < count; k += extra) {
      const uint4 it = items[k];
      near_batch<Q>(it.x, it.y, it.z, partial + (size_t)it.w * 64, cells, srcs, offs, cnts, lexp, px, py, ps, ux, uy,
                    perm, vxs, vys, gp, dmin2, dmax2);
    }
    return;
  }
  const uint32_t nc = fp->ncg;
  for (uint32_t A = blockIdx.x - extra; A < nc; A += gridDim.x - extra) {
    const FCell ca = cells[A];
    if (ca.nchild != 0 || ca.count == 0 || split0[A]) continue;
    near_batch<Q>(A, 0, NEAR_ALL, nullptr, cells, srcs, offs, cnts, lexp, px, py, ps, ux, uy, perm, vxs, vys, gp, dmin2,
                  dmax2);
  }
}

// Chunk sums of a split (leaf, batch) added in chunk order, then L2P and the
// velocity update, as near_batch does for an unsplit batch.
template <int Q>
__global__ __launch_bounds__(64) void k_near_combine(const FCell* __restrict__ cells, const FmmParams* __restrict__ fp,
                                                     const uint4* __restrict__ combines,
                                                     const float2* __restrict__ partial,
                                                     const float2* __restrict__ lexp, const double* __restrict__ ux,
                                                     const double* __restrict__ uy, const uint32_t* __restrict__ perm,
                                                     double* __restrict__ vxs, double* __restrict__ vys,
                                                     const FmmGraph* __restrict__ gp) {
  if (fp->abort) return;
  const uint32_t count = fp->near_combines;
  const uint32_t t = threadIdx.x;
  for (uint32_t k = blockIdx.x; k < count; k += gridDim.x) {
    const uint4 it = combines[k];
    const FCell ca = cells[it.x];
    const uint32_t T = ca.count - it.y < 64 ? ca.count - it.y : 64;
    if (t >= T) continue;
    float2 total = partial[(size_t)it.z * 64 + t];
    for (uint32_t c = 1; c < it.w; ++c) {
      const float2 p = partial[(size_t)(it.z + c) * 64 + t];
      total.x += p.x;
      total.y += p.y;
    }
    near_finish<Q>(ca, it.x, ca.first + it.y + t, total, lexp, ux, uy, perm, vxs, vys, gp[ca.graph].inv_w);
  }
}

// Device-side checks of a tick whose sizes the host does not read (a captured
// graph): after the build (stage 0) the cells must fit every per-cell buffer;
// after the traversal (stage 1) the lists must fit and the rounds must have
// drained the recursion. A failed check sets abort, which every later FMM
// kernel tests first, and zeroes the guarded cell count ncg; the host reads
// the flags after the tick and redoes it eagerly with larger buffers.
__global__ void k_fmm_guard(FmmParams* __restrict__ fp, uint32_t stage, uint32_t cell_room, uint32_t front_cap,
                            uint32_t last_step) {
  if (threadIdx.x != 0) return;
  const uint32_t nc = fp->level_off[FMM_LEVELS + 1];
  if (stage == 0) {  // last_step: the depths built (cells at that depth would need more)
    fp->abort = (fp->cell_overflow || nc > cell_room || fp->roots > front_cap ||
                 (last_step < FMM_LEVELS && fp->level_n[last_step] != 0))
                    ? 1u
                    : 0u;
  } else if (fp->list_overflow || fp->front[last_step * FRONT_STRIDE] != 0) {
    fp->abort |= 2u;
  }
  fp->ncg = fp->abort ? 0u : nc;
}

__global__ __launch_bounds__(256) void k_fmm_zero(uint32_t* __restrict__ p, const uint32_t* __restrict__ count) {
  const uint32_t n = *count;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) p[i] = 0;
}

// ---------------------------------------------------------------- direct sum

// FP64 reference force for the error measurement: d3's pair term summed over
// every other node in index order, one block per sampled target
// (fmm.zig:946-971, test/fast/error-test.js:44-58): coincident pairs and
// NaN terms are skipped, l < dmin2 -> sqrt(dmin2 * l), l >= dmax2 skipped.
__global__ __launch_bounds__(256) void k_direct_sum(const double* __restrict__ xs, const double* __restrict__ ys,
                                                    const double* __restrict__ st, uint32_t n,
                                                    const uint32_t* __restrict__ sample, double alpha, double dmin2,
                                                    double dmax2, double* __restrict__ out) {
  const uint32_t i = sample[blockIdx.x];
  const double xi = xs[i], yi = ys[i];
  double vx = 0, vy = 0;
  for (uint32_t j = threadIdx.x; j < n; j += 256) {
    if (j == i) continue;
    const double dx = xs[j] - xi, dy = ys[j] - yi;
    double l = dx * dx + dy * dy;
    if (!(l > 0) || !(l < dmax2)) continue;
    const double s = st[j] * alpha;
    if (s != s) continue;
    if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
    const double w = s / l;
    vx += dx * w;
    vy += dy * w;
  }
  __shared__ double sx[256], sy[256];
  sx[threadIdx.x] = vx;
  sy[threadIdx.x] = vy;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      sx[threadIdx.x] += sx[threadIdx.x + k];
      sy[threadIdx.x] += sy[threadIdx.x + k];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = sx[0];
    out[2 * blockIdx.x + 1] = sy[0];
  }
}

// ---------------------------------------------------------------- host

template <class T>
hipError_t grow(T*& p, size_t n) {
  if (p) HIPCHECK(hipFree(p));
  p = nullptr;
  return hipMalloc((void**)&p, (n > 0 ? n : 1) * sizeof(T));
}

// The cell array and its per-cell parents (contents not kept).
hipError_t grow_cells(FmmBuffers& f, uint32_t cap) {
  HIPCHECK(grow(f.cells, cap));
  HIPCHECK(grow(f.parent, cap));
  f.cell_cap = cap;
  return hipSuccess;
}

hipError_t upload_binomials() {
  static bool done = false;
  if (done) return hipSuccess;
  std::vector<double> b((size_t)BN * BN, 0.0);
  for (int n = 0; n < BN; ++n) {
    b[(size_t)n * BN] = 1;
    for (int k = 1; k <= n; ++k) b[(size_t)n * BN + k] = b[(size_t)(n - 1) * BN + k - 1] + (k < n ? b[(size_t)(n - 1) * BN + k] : 0.0);
  }
  std::vector<float> f(b.begin(), b.end());
  HIPCHECK(hipMemcpyToSymbol(HIP_SYMBOL(c_binom), f.data(), f.size() * sizeof(float)));
  done = true;
  return hipSuccess;
}

hipError_t reserve_points(FmmBuffers& f, uint32_t n) {
  if (!f.params) {
    HIPCHECK(hipMalloc((void**)&f.params, sizeof(FmmParams)));
    HIPCHECK(hipHostMalloc((void**)&f.params_host, sizeof(FmmParams), hipHostMallocDefault));
  }
  if (n <= f.point_cap) return hipSuccess;
  const uint32_t cap = n < 4096 ? 4096 : n;
  HIPCHECK(grow(f.px, cap));
  HIPCHECK(grow(f.py, cap));
  HIPCHECK(grow(f.ux, cap));
  HIPCHECK(grow(f.uy, cap));
  HIPCHECK(grow(f.ps, cap));
  HIPCHECK(grow(f.fl, cap));
  const uint32_t blk = ceil_div(cap, CB_PTS);
  HIPCHECK(grow(f.blk_cnt, (size_t)CB_DEPTHS * blk));
  HIPCHECK(grow(f.blk_fb, (size_t)CB_DEPTHS * blk));
  HIPCHECK(grow(f.blk_off, (size_t)CB_DEPTHS * blk));
  HIPCHECK(grow(f.blk_nb, (size_t)CB_DEPTHS * blk));
  f.blk_cap = blk;
  f.big_cap = cap / FMM_DIRECT + 1;
  HIPCHECK(grow(f.big, (size_t)CB_DEPTHS * f.big_cap));
  f.mid_cap = (uint32_t)((uint64_t)CB_DEPTHS * cap / (UP_SMALL + 1)) + 64;  // at most cap / (UP_SMALL + 1) a depth
  HIPCHECK(grow(f.mid, f.mid_cap));
  f.point_cap = cap;
  if (f.cell_cap < cap / 4 + 4096) HIPCHECK(grow_cells(f, cap / 4 + 4096));
  return hipSuccess;
}

// Per-cell arrays and the first list capacities, kept across runs.
hipError_t reserve_cells(FmmBuffers& f, uint32_t nc, uint32_t q) {
  if (nc > f.cellarr_cap) {
    const uint32_t cap = nc + nc / 4 + 1024;
    HIPCHECK(grow(f.bbox, (size_t)4 * cap));
    HIPCHECK(grow(f.m2l_off, cap));
    HIPCHECK(grow(f.m2l_cnt, cap));
    HIPCHECK(grow(f.p2p_off, cap));
    HIPCHECK(grow(f.p2p_cnt, cap));
    HIPCHECK(grow(f.cursor, cap));
    HIPCHECK(grow(f.scan_partials, (size_t)cap / 4096 + 64));
    f.cellarr_cap = cap;
  }
  if ((size_t)nc * q > f.exp_cap) {
    const size_t cap = (size_t)(nc + nc / 4 + 1024) * q;
    HIPCHECK(grow(f.mexp, 2 * cap));
    HIPCHECK(grow(f.lexp, 2 * cap));
    f.exp_cap = cap;
  }
  const uint32_t want_front = std::max<uint32_t>(1u << 16, 16 * nc);
  if (want_front > f.front_cap) {
    HIPCHECK(grow(f.front_a, want_front));
    HIPCHECK(grow(f.front_b, want_front));
    f.front_cap = want_front;
  }
  const uint32_t want_list = std::max<uint32_t>(1u << 16, 40 * nc);
  if (want_list > f.m2l_cap) {
    HIPCHECK(grow(f.m2l, want_list));
    HIPCHECK(grow(f.m2l_src, want_list));
    HIPCHECK(grow(f.p2p, want_list));
    HIPCHECK(grow(f.p2p_src, want_list));
    f.m2l_cap = f.p2p_cap = want_list;
  }
  return hipSuccess;
}

// Grids of the capturable form (grid-stride kernels over device counts).
constexpr uint32_t CAP_CELL_BLOCKS = 4096;  // per-cell and per-pair elementwise kernels
constexpr uint32_t CAP_TARGET_BLOCKS = 65536;  // one wave per target, looping (many, for load balance)
constexpr uint32_t CAP_NEAR_BLOCKS = 1u << 22;   // the near field's blocks (one a cell, up to this)

template <int Q>
hipError_t run_order(FmmBuffers& f, TreeBuffers& t, const FmmArgs& a, hipStream_t stream) {
  const FmmParams& P = *f.params_host;
  const bool g = a.graph;
  // The eager form sizes its launches from the counts read back after the
  // build; the capturable form from the capacities, every kernel reading its
  // count on the device (both run the same kernels on the same data).
  const uint32_t nc = g ? f.cell_cap : P.level_off[FMM_LEVELS + 1];
  // the capturable form: every cell is above depth graph_levels (k_fmm_guard)
  uint32_t maxd = f.graph_levels > 0 ? f.graph_levels - 1 : 0;
  if (!g) {
    maxd = 0;
    for (uint32_t d = 0; d <= FMM_LEVELS; ++d)
      if (P.level_n[d]) maxd = d;
    if (nc > f.cell_cap || nc > f.cellarr_cap || (size_t)nc * Q > f.exp_cap)
      HIPFAIL(hipErrorInvalidValue, "cells beyond the reserved arrays");
  }
  // Per-cell launches: count bounds the cells (the capacity in the
  // capturable form, whose kernels loop over the device's count).
  auto grid = [&](uint32_t count, uint32_t per, uint32_t cap_blocks) -> uint32_t {
    const uint32_t b = std::max(1u, ceil_div(count, per));
    return g ? std::min(cap_blocks, b) : b;
  };
  // Per-pair launches: the capturable form cannot bound the pairs by the
  // last eager count, so its blocks loop over them.
  auto grid_pairs = [&](uint32_t count, uint32_t per, uint32_t cap_blocks) -> uint32_t {
    return g ? cap_blocks : std::max(1u, ceil_div(count, per));
  };
  // upward: the geometry first (the dual traversal reads it), then the
  // multipoles, which M2L reads: with an auxiliary stream (the eager form)
  // they run there beside the dual traversal, joined before M2L. Cells of up
  // to FMM_DIRECT points from their points (8 lanes a cell up to UP_SMALL
  // points, else a wave; the capturable form's level_blocks blocks loop over
  // them), the larger ones from their children in one workgroup.
  const uint32_t small_blocks = std::max(1u, std::min(g ? f.level_blocks : 1u << 20, ceil_div(nc, 32)));
  const uint32_t mid_blocks = std::max(1u, std::min(f.level_blocks, ceil_div(nc, 16)));
  hipLaunchKernelGGL((k_fmm_up_small<Q, false>), dim3(small_blocks), dim3(256), 0, stream, f.cells, f.bbox, f.ux,
                     f.uy, f.ps, (float2*)f.mexp, f.params);
  hipLaunchKernelGGL((k_fmm_up_mid<Q, false>), dim3(mid_blocks), dim3(256), 0, stream, f.cells, f.bbox, f.ux, f.uy,
                     f.ps, (float2*)f.mexp, f.mid, f.mid_cap, f.params);
  hipLaunchKernelGGL((k_fmm_up_big<Q, false>), dim3(1), dim3(1024), 0, stream, f.cells, f.bbox, f.ux, f.uy, f.ps,
                     (float2*)f.mexp, f.big, f.big_cap, f.params);
  const bool aux = a.aux != nullptr;  // captured too: the fork and join become graph edges
  hipStream_t mps = stream;
  if (aux) {
    if (!f.ev_geom) HIPCHECK(hipEventCreateWithFlags(&f.ev_geom, hipEventDisableTiming));
    if (!f.ev_mp) HIPCHECK(hipEventCreateWithFlags(&f.ev_mp, hipEventDisableTiming));
    HIPCHECK(hipEventRecord(f.ev_geom, stream));
    HIPCHECK(hipStreamWaitEvent(a.aux, f.ev_geom, 0));
    mps = a.aux;
  }
  hipLaunchKernelGGL((k_fmm_up_small<Q, true>), dim3(small_blocks), dim3(256), 0, mps, f.cells, f.bbox, f.ux, f.uy,
                     f.ps, (float2*)f.mexp, f.params);
  hipLaunchKernelGGL((k_fmm_up_mid<Q, true>), dim3(mid_blocks), dim3(256), 0, mps, f.cells, f.bbox, f.ux, f.uy,
                     f.ps, (float2*)f.mexp, f.mid, f.mid_cap, f.params);
  hipLaunchKernelGGL((k_fmm_up_big<Q, true>), dim3(1), dim3(1024), 0, mps, f.cells, f.bbox, f.ux, f.uy, f.ps,
                     (float2*)f.mexp, f.big, f.big_cap, f.params);
  if (aux) HIPCHECK(hipEventRecord(f.ev_mp, a.aux));
  // interaction lists (pair keys: target << nbits | source)
  uint32_t nbits = 1;
  while (((uint64_t)1 << nbits) < f.cell_cap) ++nbits;
  MacArgs mac;
  double kappa = a.opening;
  if (!(kappa > 1)) kappa = 1.4;  // fmm.zig:262-263
  mac.mac_far = 2 * kappa - 1;
  mac.nbits = nbits;
  mac.graphs = f.graphs;
  const uint32_t steps = 2 * maxd + 1;  // a traversal not drained by then aborts the replay (k_fmm_guard)
  const uint32_t root_blocks = ceil_div(a.graphs > 0 ? a.graphs : 1, EWF);  // one root per graph at most
  // per-target list counts, taken by the rounds (cleared by kernels: no memset node)
  const uint32_t* ncp = &f.params->ncg;
  for (;;) {
    hipLaunchKernelGGL(k_fmm_zero, dim3(grid(nc, 256, CAP_CELL_BLOCKS)), dim3(256), 0, stream, f.m2l_cnt, ncp);
    hipLaunchKernelGGL(k_fmm_zero, dim3(grid(nc, 256, CAP_CELL_BLOCKS)), dim3(256), 0, stream, f.p2p_cnt, ncp);
    hipLaunchKernelGGL(k_fmm_lists_reset, dim3(root_blocks), dim3(EWF), 0, stream, f.params, f.front_a, f.front_cap);
    uint2 *in = f.front_a, *out = f.front_b;
    for (uint32_t s = 0; s < steps; ++s) {
      hipLaunchKernelGGL(k_dual_step, dim3(f.dual_blocks), dim3(EWF), 0, stream, f.cells, f.params, s, in, out,
                         f.front_cap, f.m2l, f.m2l_cap, f.p2p, f.p2p_cap, f.m2l_cnt, f.p2p_cnt, mac);
      std::swap(in, out);
    }
    if (g) break;
    HIPCHECK(hipMemcpyAsync(f.params_host, f.params, sizeof(FmmParams), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (!P.list_overflow) {
      if (P.front[steps * FRONT_STRIDE] != 0) HIPFAIL(hipErrorUnknown, "dual traversal not drained");  // the rounds must drain the recursion
      break;
    }
    uint32_t fmax = 0;
    for (uint32_t s = 0; s <= steps; ++s) fmax = std::max(fmax, P.front[s * FRONT_STRIDE]);
    if (P.list_overflow & 1u) {
      const uint32_t c = std::max(2 * f.front_cap, fmax + fmax / 4);
      HIPCHECK(grow(f.front_a, c));
      HIPCHECK(grow(f.front_b, c));
      f.front_cap = c;
    }
    if (P.list_overflow & 6u) {
      const uint32_t c = std::max(P.m2l, P.p2p);
      const uint32_t cap = c + c / 4 + 1024;
      if (cap > f.m2l_cap) {
        HIPCHECK(grow(f.m2l, cap));
        HIPCHECK(grow(f.m2l_src, cap));
        HIPCHECK(grow(f.p2p, cap));
        HIPCHECK(grow(f.p2p_src, cap));
        f.m2l_cap = f.p2p_cap = cap;
      }
    }
  }
  hipLaunchKernelGGL(k_fmm_guard, dim3(1), dim3(64), 0, stream, f.params, 1u, 0u, 0u, steps);
  if (a.built) HIPCHECK(hipEventRecord(a.built, stream));  // the dual traversal is done
  if (!g) {
    // long-segment list, near-field items and partial-sum slots for these lists
    const uint32_t long_need = P.p2p / NEAR_CHUNK + P.m2l / SEG_MAX + 2;
    if (long_need > f.long_cap) {
      const uint32_t c = long_need + long_need / 4;
      HIPCHECK(grow(f.long_list, c));
      f.long_cap = c;
    }
    const uint32_t slot_need = 2 * (P.p2p / NEAR_CHUNK) + 1024;
    if (slot_need > f.slot_cap) {
      const uint32_t c = slot_need + slot_need / 4;
      HIPCHECK(grow(f.near_partial, (size_t)c * 64));
      HIPCHECK(grow(f.near_combines, c / 2 + 1));
      f.slot_cap = c;
    }
    const uint32_t item_need = f.slot_cap + P.m_total / 64 + 1;
    if (item_need > f.item_cap) {
      HIPCHECK(grow(f.near_items, item_need));
      f.item_cap = item_need;
    }
  }
  // lists grouped by target: the rounds' counts scanned, then a scatter
  const uint32_t scan_max = g ? f.cellarr_cap : nc;
  HIPCHECK(scan_u32(f.m2l_cnt, f.m2l_off, scan_max, ncp, f.scan_partials, nullptr, stream));
  HIPCHECK(scan_u32(f.p2p_cnt, f.p2p_off, scan_max, ncp, f.scan_partials, nullptr, stream));
  if (g || P.m2l) {
    hipLaunchKernelGGL(k_fmm_zero, dim3(grid(nc, 256, CAP_CELL_BLOCKS)), dim3(256), 0, stream, f.cursor, ncp);
    hipLaunchKernelGGL(k_target_scatter, dim3(grid_pairs(P.m2l, EWF, CAP_CELL_BLOCKS)), dim3(EWF), 0, stream, f.m2l,
                       &f.params->m2l, nbits, f.m2l_off, f.cursor, f.m2l_src, f.params);
  }
  if (g || P.p2p) {
    hipLaunchKernelGGL(k_fmm_zero, dim3(grid(nc, 256, CAP_CELL_BLOCKS)), dim3(256), 0, stream, f.cursor, ncp);
    hipLaunchKernelGGL(k_target_scatter, dim3(grid_pairs(P.p2p, EWF, CAP_CELL_BLOCKS)), dim3(EWF), 0, stream, f.p2p,
                       &f.params->p2p, nbits, f.p2p_off, f.cursor, f.p2p_src, f.params);
  }
  // segments too long for the in-LDS order, ordered in place
  hipLaunchKernelGGL(k_long_list, dim3(grid(nc, EWF, CAP_CELL_BLOCKS)), dim3(EWF), 0, stream, ncp, f.m2l_cnt,
                     f.p2p_cnt, f.params, f.long_list);
  hipLaunchKernelGGL(k_sort_long, dim3(LONG_BLOCKS), dim3(256), 0, stream, f.long_list, f.params, ncp, f.m2l_src,
                     f.m2l_off, f.m2l_cnt, (uint32_t*)f.m2l, f.p2p_src, f.p2p_off, f.p2p_cnt, (uint32_t*)f.p2p);
  // far field
  if (aux) HIPCHECK(hipStreamWaitEvent(stream, f.ev_mp, 0));  // the multipoles are done
  hipLaunchKernelGGL(k_m2l<Q>, dim3(grid(nc, 1, CAP_TARGET_BLOCKS)), dim3(64), 0, stream, f.cells, f.m2l_src,
                     f.m2l_off, f.m2l_cnt, (const float2*)f.mexp, (float2*)f.lexp, f.params);
  hipLaunchKernelGGL(k_fmm_down_leaf<Q>, dim3(grid(nc, EWF, CAP_CELL_BLOCKS)), dim3(EWF), 0, stream, f.cells, f.parent,
                     f.m2l_cnt, (float2*)f.lexp, f.params);
  // near field and evaluation
  const float dmin2 = (float)a.dmin2, dmax2 = (float)a.dmax2;
  hipLaunchKernelGGL(k_near_items, dim3(grid(nc, EWF, CAP_CELL_BLOCKS)), dim3(EWF), 0, stream, f.cells, ncp,
                     f.p2p_cnt, f.params, f.near_items, f.near_combines, f.cursor, f.slot_cap);
  const uint32_t extra = g ? f.near_extra : (P.m_total / 64 < f.near_extra ? P.m_total / 64 + 1 : f.near_extra);
  // one block a cell, as eagerly; captured, for the cells of the eager tick
  // before the capture and an eighth more (the blocks loop over any others)
  const uint32_t seen = P.level_off[FMM_LEVELS + 1];
  const uint32_t near_blocks =
      g ? std::min(std::min(f.cell_cap, CAP_NEAR_BLOCKS), std::max(4096u, seen + seen / 8 + 1024)) : grid(nc, 1, 0);
  hipLaunchKernelGGL(k_near<Q>, dim3(extra + near_blocks), dim3(64), 0, stream, f.cells,
                     f.p2p_src, f.p2p_off, f.p2p_cnt, f.params, (const float2*)f.lexp, f.px, f.py, f.ps, f.ux, f.uy,
                     t.perm, a.vx, a.vy, f.graphs, dmin2, dmax2, f.near_items, f.cursor, f.near_partial, extra);
  hipLaunchKernelGGL(k_near_combine<Q>, dim3(extra), dim3(64), 0, stream, f.cells, f.params, f.near_combines,
                     f.near_partial, (const float2*)f.lexp, f.ux, f.uy, t.perm, a.vx, a.vy, f.graphs);
  return hipGetLastError();
}

using OrderFn = hipError_t (*)(FmmBuffers&, TreeBuffers&, const FmmArgs&, hipStream_t);

template <int... Qs>
struct OrderTable {
  static constexpr OrderFn fns[] = {&run_order<Qs>...};
};
template <int... Qs>
constexpr OrderFn OrderTable<Qs...>::fns[];

using Orders = OrderTable<2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25>;

}  // namespace

void fmm_free(FmmBuffers& f) {
  void* ptrs[] = {f.graphs, f.seeds, f.node_off_dev, f.chunk_lo, f.chunk_hi, f.chunk_first, f.box_parts, f.cells, f.parent, f.bbox, f.fl, f.blk_cnt, f.blk_fb, f.blk_off, f.blk_nb, f.big, f.mid, f.mexp, f.lexp, f.px, f.py, f.ux, f.uy, f.ps,
                  f.front_a, f.front_b, f.long_list, f.near_items, f.near_combines, f.near_partial, f.m2l, f.p2p, f.m2l_src, f.p2p_src, f.m2l_off, f.m2l_cnt, f.p2p_off,
                  f.p2p_cnt, f.cursor, f.scan_partials, f.params};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (f.params_host) (void)hipHostFree(f.params_host);
  if (f.ev_geom) (void)hipEventDestroy(f.ev_geom);
  if (f.ev_mp) (void)hipEventDestroy(f.ev_mp);
  f = FmmBuffers();
}

hipError_t direct_sum(const double* x, const double* y, const double* strength, uint32_t n, const uint32_t* sample,
                      uint32_t k, double alpha, double dmin2, double dmax2, double* out, hipStream_t stream) {
  if (k == 0) return hipSuccess;
  hipLaunchKernelGGL(k_direct_sum, dim3(k), dim3(256), 0, stream, x, y, strength, n, sample, alpha, dmin2, dmax2, out);
  return hipGetLastError();
}

// Chunks of at most FOREST_CHUNK nodes that never straddle a graph, for the
// per-graph boxes; rebuilt when the forest changes.
constexpr uint32_t FOREST_CHUNK = 2048;

hipError_t reserve_forest(FmmBuffers& f, const FmmArgs& a, hipStream_t stream) {
  const uint32_t G = a.graphs;
  if (G > f.graph_cap) {
    const uint32_t cap = G + G / 4 + 16;
    HIPCHECK(grow(f.graphs, cap));
    HIPCHECK(grow(f.seeds, cap));
    HIPCHECK(grow(f.chunk_first, cap + 1));
    HIPCHECK(grow(f.node_off_dev, cap + 1));
    f.graph_cap = cap;
  }
  const bool same = f.node_off_host.size() == (size_t)G + 1 &&
                    std::equal(f.node_off_host.begin(), f.node_off_host.end(), a.node_off);
  if (!same) {
    std::vector<uint32_t> lo, hi, first(G + 1);
    for (uint32_t g = 0; g < G; ++g) {
      first[g] = (uint32_t)lo.size();
      for (uint32_t b = a.node_off[g]; b < a.node_off[g + 1]; b += FOREST_CHUNK) {
        lo.push_back(b);
        hi.push_back(std::min(b + FOREST_CHUNK, a.node_off[g + 1]));
      }
    }
    first[G] = (uint32_t)lo.size();
    const uint32_t C = (uint32_t)lo.size();
    if (C > f.chunk_cap) {
      const uint32_t cap = C + C / 4 + 16;
      HIPCHECK(grow(f.chunk_lo, cap));
      HIPCHECK(grow(f.chunk_hi, cap));
      if (f.box_parts) HIPCHECK(hipFree(f.box_parts));
      f.box_parts = nullptr;
      HIPCHECK(hipMalloc(&f.box_parts, (size_t)cap * sizeof(BoxPart)));
      f.chunk_cap = cap;
    }
    if (C) {
      HIPCHECK(hipMemcpyAsync(f.chunk_lo, lo.data(), C * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipMemcpyAsync(f.chunk_hi, hi.data(), C * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    }
    HIPCHECK(hipMemcpyAsync(f.chunk_first, first.data(), (G + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(f.node_off_dev, a.node_off, (G + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipStreamSynchronize(stream));  // the staging vectors end here
    f.node_off_host.assign(a.node_off, a.node_off + G + 1);
    f.chunks = C;
  }
  if (!a.seeds_dev && a.seeds)
    HIPCHECK(hipMemcpyAsync(f.seeds, a.seeds, G * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
  return hipSuccess;
}

hipError_t fmm_run(FmmBuffers& f, TreeBuffers& t, const FmmArgs& a, hipStream_t stream) {
  HIPCHECK(upload_binomials());
  const uint32_t n = a.n;
  uint32_t p = a.order;  // fmm.zig:258-261
  if (p < 1 || p > FMM_MAX_ORDER) p = 8;
  if (a.graphs == 0 || a.node_off[a.graphs] != n) HIPFAIL(hipErrorInvalidValue, "graph offsets do not cover n");
  HIPCHECK(tree_reserve(t, n > 0 ? n : 1));
  HIPCHECK(reserve_points(f, n));
  HIPCHECK(reserve_forest(f, a, stream));
  if (f.cell_cap < a.graphs + 4096) HIPCHECK(grow_cells(f, a.graphs + 4096));
  const uint32_t nb = ceil_div(n > 0 ? n : 1, EWF);
  // Boxes, frames and keys of every graph; NaN points compacted out; one sort.
  hipLaunchKernelGGL(k_forest_reset, dim3(1), dim3(64), 0, stream, f.params, n);
  if (f.chunks)
    hipLaunchKernelGGL(k_forest_box, dim3(f.chunks), dim3(EWF), 0, stream, a.x, a.y, f.chunk_lo, f.chunk_hi,
                       (BoxPart*)f.box_parts);
  hipLaunchKernelGGL(k_forest_graphs, dim3(a.graphs), dim3(64), 0, stream, (const BoxPart*)f.box_parts,
                     f.chunk_first, a.graphs, a.dmin2, a.dmax2, a.seeds_dev ? a.seeds_dev : f.seeds, f.node_off_dev,
                     a.min_graph_n, f.graphs, f.params);
  hipLaunchKernelGGL(k_forest_keys, dim3(nb), dim3(EWF), 0, stream, a.x, a.y, n, a.node_graph, f.graphs, f.params,
                     t.keys, t.perm, t.keys_all, t.flags);
  HIPCHECK(scan_u32(t.flags, t.offs, n > 0 ? n : 1, &f.params->n, t.partials, nullptr, stream));
  hipLaunchKernelGGL(k_forest_compact, dim3(nb), dim3(EWF), 0, stream, n, f.params, t.keys_all, t.flags, t.offs,
                     t.keys, t.perm);
  int digits = 4;  // 32 Morton bits (tree.zig:363), then the graph's bytes
  for (uint32_t v = a.graphs - 1; v > 0; v >>= 8) ++digits;
  HIPCHECK(sort_digits(t.hist, t.rs_glob, t.keys, t.perm, t.keys_tmp, t.perm_tmp, n > 0 ? n : 1, &f.params->m_total,
                       0, digits, stream));
  // Cells of every depth in three launches; grown and rebuilt if they
  // outgrow the buffer. The capturable form's loops cover the last eager
  // tree's depths plus two (a deeper tree aborts the replay).
  const uint32_t d_end = std::min(FMM_LEVELS, f.build_levels);
  f.graph_levels = d_end;
  const uint32_t nblk = ceil_div(n > 0 ? n : 1, CB_PTS);
  if (nblk > f.blk_cap) HIPFAIL(hipErrorInvalidValue, "cell build blocks beyond the reserved arrays");
  hipLaunchKernelGGL(k_forest_roots, dim3(1), dim3(1024), 0, stream, f.graphs, a.graphs, f.params, t.meta);
  for (;;) {
    hipLaunchKernelGGL(k_cells_mark, dim3(nblk), dim3(CB_THREADS), 0, stream, t.keys, f.graphs, f.params, nblk, f.fl,
                       f.blk_cnt, f.blk_fb);
    hipLaunchKernelGGL(k_cells_scan, dim3(CB_DEPTHS), dim3(1024), 0, stream, f.blk_cnt, f.blk_fb, nblk, f.blk_off,
                       f.blk_nb, f.params, f.cell_cap);
    hipLaunchKernelGGL(k_cells_emit, dim3(nblk), dim3(CB_THREADS), 0, stream, t.keys, f.fl, nblk, f.blk_off,
                       f.blk_nb, f.params, f.cells, f.parent, f.cell_cap, f.big, f.big_cap, f.mid, f.mid_cap);
    if (a.graph) break;  // sizes checked on the device (k_fmm_guard)
    HIPCHECK(hipMemcpyAsync(f.params_host, f.params, sizeof(FmmParams), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (!f.params_host->cell_overflow) break;
    HIPCHECK(grow_cells(f, std::max(2 * f.cell_cap, f.params_host->level_off[FMM_LEVELS + 1] + 4096)));
    hipLaunchKernelGGL(k_forest_reset_overflow, dim3(1), dim3(64), 0, stream, f.params);
  }
  if (!a.graph) {
    const FmmParams& P = *f.params_host;
    uint32_t deepest = 0;
    for (uint32_t d = 0; d <= FMM_LEVELS; ++d)
      if (P.level_n[d]) deepest = d;
    f.build_levels = std::min(FMM_LEVELS, deepest + 2);
    // fmm.zig:255-257: one graph with a non-finite extent runs the exact kernel.
    if (a.graphs == 1 && P.declined) return hipErrorNotSupported;
    if (P.roots == 0) return hipSuccess;  // no point in any tree (fmm.zig:251-252)
    HIPCHECK(reserve_cells(f, P.level_off[FMM_LEVELS + 1], p + 1));
  }
  uint32_t room = std::min<uint32_t>(std::min(f.cell_cap, f.cellarr_cap), (uint32_t)(f.exp_cap / (p + 1)));
  if (a.graph && a.cell_limit && a.cell_limit < room) room = a.cell_limit;
  hipLaunchKernelGGL(k_fmm_guard, dim3(1), dim3(64), 0, stream, f.params, 0u, room, f.front_cap,
                     a.graph ? d_end : FMM_LEVELS);
  hipLaunchKernelGGL(k_fmm_gather, dim3(nb), dim3(EWF), 0, stream, t.perm, t.keys, a.x, a.y, a.strength, f.graphs,
                     f.params, a.alpha, a.alpha_dev, n, f.px, f.py, f.ux, f.uy, f.ps, f.cells, f.parent);
  return Orders::fns[p - 1](f, t, a, stream);
}

hipError_t fmm_reserve_capacity(FmmBuffers& f, TreeBuffers& t, const FmmArgs& a, uint32_t cells, hipStream_t stream) {
  uint32_t p = a.order;
  if (p < 1 || p > FMM_MAX_ORDER) p = 8;
  HIPCHECK(upload_binomials());
  HIPCHECK(tree_reserve(t, a.n > 0 ? a.n : 1));
  HIPCHECK(reserve_points(f, a.n));
  HIPCHECK(reserve_forest(f, a, stream));
  if (f.cell_cap < cells) HIPCHECK(grow_cells(f, cells));
  HIPCHECK(reserve_cells(f, f.cell_cap, p + 1));
  const uint32_t long_need = f.p2p_cap / NEAR_CHUNK + f.m2l_cap / SEG_MAX + 2;
  if (long_need > f.long_cap) {
    HIPCHECK(grow(f.long_list, long_need));
    f.long_cap = long_need;
  }
  const uint32_t slot_need = 2 * (f.p2p_cap / NEAR_CHUNK) + 1024;
  if (slot_need > f.slot_cap) {
    HIPCHECK(grow(f.near_partial, (size_t)slot_need * 64));
    HIPCHECK(grow(f.near_combines, slot_need / 2 + 1));
    f.slot_cap = slot_need;
  }
  const uint32_t item_need = f.slot_cap + f.point_cap / 64 + 1;
  if (item_need > f.item_cap) {
    HIPCHECK(grow(f.near_items, item_need));
    f.item_cap = item_need;
  }
  return hipSuccess;
}

}  // namespace d3f
