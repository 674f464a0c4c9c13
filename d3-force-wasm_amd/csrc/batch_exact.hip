// Exact mode for the small graphs of a batch, on gfx950: a graph below the
// fast-mode crossover (n < 2000, root.zig:246-248) ticks with the exact
// kernels in the reference, so the batch runs its manyBody and link exactly
// as the reference's exact tick does (many_body.zig, link.zig), one 256-thread
// workgroup per graph, everything the graph needs in one launch:
//
//  1. box of the non-NaN points (first of equal values kept) and d3's cover
//     (tree.zig:148-187, 230-248);
//  2. 32-level midpoint keys (tree.zig:252-284), NaN points left out in index
//     order; a stable sort of (key, node) in LDS (bitonic, the node index
//     breaking ties, which is the stable order);
//  3. run boundaries (tree.zig:392-433) and the preorder emit of the cells on
//     one thread (tree.zig:437-647): children 0..3, a leaf when all points of
//     a cell coincide, a stable midpoint partition below depth 32;
//  4. accumulate level by level from the deepest (many_body.zig:101-135),
//     writing the traversal records with each cell's subtree size;
//  5. the Barnes-Hut walk of every node (many_body.zig:176-263), one node per
//     lane and no stack (cells are in visit preorder: a cell's first child is
//     the next cell, its subtree ends `skip` cells on); a walk that meets a
//     zero delta stops, and those nodes are redone on one thread in ascending
//     node order drawing from the graph's random source (many_body.zig:77-83);
//  6. the Gauss-Seidel link (link.zig:8-40) a level of the link schedule at a
//     time (links of a level share no node; k_small_schedule), on one thread
//     in link order when a link would draw.
//
// Random draws follow the facade in fast precision (core.js:213-231): one
// draw for the manyBody seed, manyBody's jiggle draws, one draw for the link
// seed, link's jiggle draws; the random source is d3's lcg (lcg.js), one
// state per graph.
#include "common.h"

namespace d3f {
namespace {

constexpr uint32_t SX_MAX = 2048;  // sort capacity: graphs below 2000 nodes
constexpr int SX_THREADS = 256;
constexpr int SX_STACK = 192;      // emit frames held in LDS (trees up to ~60 levels)
constexpr uint32_t SX_CELLS = 3 * SX_MAX + 128;  // cells per graph (3 n + 128)
constexpr uint32_t SX_DEPTHS = 64;               // depth bins of the accumulate

// The emit's cell: bounds, children, its points in perm and its depth.
struct SCell {
  double x0, y0, x1, y1;
  uint32_t child[4];
  uint32_t first, count, depth, pad;
};
static_assert(sizeof(SCell) == 64, "small-graph cell");

struct SFrame {
  double x0, y0, x1, y1;
  uint32_t parent;
  uint16_t lo, hi, depth;
  uint8_t slot, sorted;
};

__device__ inline double lcg_next(uint32_t& s) {  // lcg.js: s = (a s + c) mod 2^32, s / 2^32
  s = 1664525u * s + 1013904223u;
  return (double)s / 4294967296.0;
}

__device__ inline double jiggle_of(double r) { return (r - 0.5) * 1e-6; }  // rng.zig:7-11

__device__ inline bool truthy(double v) { return v != 0 && v == v; }

__device__ inline void take_min(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void take_max(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}

// d3-quadtree cover (tree.zig:148-187).
__device__ void cover_point(double e[4], double x, double y) {
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

__device__ inline unsigned quadrant_of(double x, double y, double xm, double ym) {  // tree.zig:697-699
  return (unsigned)(y >= ym) << 1 | (unsigned)(x >= xm);
}

// One node's walk (many_body.zig:176-263). DRAWS: 0 stops at the first point
// where d3 would call random (returns false), 1 draws from the graph's lcg.
template <int DRAWS>
__device__ bool walk(uint32_t i, double nx, double ny, double& vx, double& vy, const TCell* __restrict__ tc,
                     uint32_t len, const uint16_t* __restrict__ perm, uint32_t node0, const double* __restrict__ s,
                     double dmin2, double dmax2, double alpha, uint32_t* lcg) {
  uint32_t c = 0;
  while (c < len) {
    const TCell& r = tc[c];
    const double value = r.value;
    const uint32_t skip = r.skip;
    uint32_t next = c + (skip > 0 ? skip : 1);
    if (truthy(value)) {
      double dx = r.qx - nx, dy = r.qy - ny;
      double l = dx * dx + dy * dy;
      if (r.w2t < l) {
        if (l < dmax2) {
          if (dx == 0 || dy == 0) {
            if (DRAWS == 0) return false;
            if (dx == 0) { dx = jiggle_of(lcg_next(*lcg)); l += dx * dx; }
            if (dy == 0) { dy = jiggle_of(lcg_next(*lcg)); l += dy * dy; }
          }
          if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
          vx += dx * value * alpha / l;
          vy += dy * value * alpha / l;
        }
      } else if (skip > 1) {
        next = c + 1;
      } else if (!(l >= dmax2)) {
        const uint32_t first = r.first, cnt = r.count, head = r.head;
        if (head != i || cnt > 1) {
          if (dx == 0 || dy == 0) {
            if (DRAWS == 0) return false;
            if (dx == 0) { dx = jiggle_of(lcg_next(*lcg)); l += dx * dx; }
            if (dy == 0) { dy = jiggle_of(lcg_next(*lcg)); l += dy * dy; }
          }
          if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
        }
        for (uint32_t k = first + cnt; k > first;) {
          --k;
          const uint32_t j = node0 + perm[k];
          if (j != i) {
            const double wj = s[j] * alpha / l;
            vx += dx * wj;
            vy += dy * wj;
          }
        }
      }
    }
    c = next;
  }
  return true;
}

// A walk of the first pass, several in flight per lane so that their cell
// loads overlap: a walk that meets a point where d3 would call random stops
// there (drew) and is redone with the draws.
constexpr int SX_WALKS = 2;
struct SWalk {
  double nx, ny, vx, vy;
  uint32_t i, c;
  bool drew;
};

__device__ inline double f64_at(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// One visit of the first pass (walk without draws) at cell w.c, whose first 48
// bytes are g0 g1 g2; moves w.c on.
__device__ inline void step(SWalk& w, const uint4& g0, const uint4& g1, const uint4& g2, uint32_t len,
                            const uint16_t* __restrict__ perm, uint32_t node0, const double* __restrict__ s,
                            double dmin2, double dmax2, double alpha) {
  const double value = f64_at(g1.x, g1.y);
  const uint32_t skip = g2.x;
  uint32_t next = w.c + (skip > 0 ? skip : 1);
  if (truthy(value)) {
    const double dx = f64_at(g0.x, g0.y) - w.nx, dy = f64_at(g0.z, g0.w) - w.ny;
    double l = dx * dx + dy * dy;
    if (f64_at(g1.z, g1.w) < l) {
      if (l < dmax2) {
        if (dx == 0 || dy == 0) {
          w.drew = true;
          next = len;
        } else {
          if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
          w.vx += dx * value * alpha / l;
          w.vy += dy * value * alpha / l;
        }
      }
    } else if (skip > 1) {
      next = w.c + 1;
    } else if (!(l >= dmax2)) {
      const uint32_t first = g2.y, cnt = g2.z, head = g2.w;
      bool go = true;
      if (head != w.i || cnt > 1) {
        if (dx == 0 || dy == 0) {
          w.drew = true;
          next = len;
          go = false;
        } else if (l < dmin2) {
          l = __builtin_sqrt(dmin2 * l);
        }
      }
      if (go)
        for (uint32_t k = first + cnt; k > first;) {
          --k;
          const uint32_t j = node0 + perm[k];
          if (j != w.i) {
            const double wj = s[j] * alpha / l;
            w.vx += dx * wj;
            w.vy += dy * wj;
          }
        }
    }
  }
  w.c = next;
}

__global__ __launch_bounds__(SX_THREADS) void k_small_exact(
    const SmallJob* __restrict__ jobs, const double* __restrict__ xs, const double* __restrict__ ys,
    const double* __restrict__ vxs, const double* __restrict__ vys, const double* __restrict__ charge,
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt, const double* __restrict__ lstrength,
    const double* __restrict__ ldistance, const double* __restrict__ lbias, const SmallLink* __restrict__ lrec,
    const uint32_t* __restrict__ levels, const double* __restrict__ alpha_dev, double theta2, double dmin2, double dmax2, double link_iterations, uint32_t* __restrict__ lcg_state,
    SCell* __restrict__ cells_all, TCell* __restrict__ tcells_all, double* __restrict__ vout_x,
    double* __restrict__ vout_y, uint32_t* __restrict__ error) {
  // LDS: what the tree build needs, reused by the link pass for the
  // velocities (the build's arrays are dead by then)
  __shared__ union {
    struct {
      uint64_t keys[SX_MAX];
      uint16_t ptmp[SX_MAX], diff[SX_MAX];
      SFrame stack[SX_STACK];
    } b;
    struct {
      uint16_t bycell[SX_CELLS];  // cells grouped by depth (accumulate)
    } a;
    struct {
      double vx[SX_MAX], vy[SX_MAX];
    } l;
  } u;
  uint64_t* keys = u.b.keys;
  uint16_t* ptmp = u.b.ptmp;
  uint16_t* diff = u.b.diff;
  SFrame* stack = u.b.stack;
  uint16_t* bycell = u.a.bycell;
  __shared__ uint16_t perm[SX_MAX];
  __shared__ uint32_t flagged[SX_MAX / 32];
  __shared__ uint32_t dstart[SX_DEPTHS + 2];
  __shared__ double ext[4];
  __shared__ uint32_t red_i[4 * SX_THREADS / 64];
  __shared__ double red_v[4 * SX_THREADS / 64];
  __shared__ uint32_t wsum[SX_THREADS / 64 + 1];
  __shared__ uint32_t meta[4];  // m, len, max depth, overflow

  const SmallJob job = jobs[blockIdx.x];
  const uint32_t node0 = job.node0, n = job.n, tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  SCell* cells = cells_all + job.cell0;
  TCell* tc = tcells_all + job.cell0;
  const double alpha = *alpha_dev;

  // 1. box (tree.zig:316-340): first of equal values kept
  {
    const double inf = __builtin_huge_val();
    double b[4] = {inf, inf, -inf, -inf};
    uint32_t bi[4] = {NONE, NONE, NONE, NONE};
    for (uint32_t k = tid; k < n; k += SX_THREADS) {
      const double x = xs[node0 + k], y = ys[node0 + k];
      if (x != x || y != y) continue;
      take_min(b[0], bi[0], x, k);
      take_min(b[1], bi[1], y, k);
      take_max(b[2], bi[2], x, k);
      take_max(b[3], bi[3], y, k);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      for (int d = 32; d > 0; d >>= 1) {
        const double ov = __shfl_xor(b[q], d, 64);
        const uint32_t oi = __shfl_xor(bi[q], d, 64);
        if (q < 2) take_min(b[q], bi[q], ov, oi); else take_max(b[q], bi[q], ov, oi);
      }
      if (lane == 0) {
        red_v[wave * 4 + q] = b[q];
        red_i[wave * 4 + q] = bi[q];
      }
    }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < SX_THREADS / 64; ++w)
        for (int q = 0; q < 4; ++q) {
          if (q < 2) take_min(red_v[q], red_i[q], red_v[w * 4 + q], red_i[w * 4 + q]);
          else take_max(red_v[q], red_i[q], red_v[w * 4 + q], red_i[w * 4 + q]);
        }
      const double nan = __builtin_nan("");
      double e[4] = {nan, nan, nan, nan};
      if (!(red_v[0] > red_v[2] || red_v[1] > red_v[3])) {
        cover_point(e, red_v[0], red_v[1]);
        cover_point(e, red_v[2], red_v[3]);
      }
      for (int q = 0; q < 4; ++q) ext[q] = e[q];
      meta[3] = 0;
    }
    __syncthreads();
  }
  const bool valid = ext[0] == ext[0];

  // 2. keys of the finite points, compacted in index order; then the sort
  uint32_t m = 0;
  if (valid) {
    constexpr uint32_t PER = SX_MAX / SX_THREADS;  // consecutive nodes per thread
    uint32_t cnt = 0;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t k = tid * PER + u;
      if (k < n) cnt += (xs[node0 + k] == xs[node0 + k] && ys[node0 + k] == ys[node0 + k]) ? 1u : 0u;
    }
    uint32_t inc = cnt;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(inc, d, 64);
      if (lane >= d) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    if (tid == 0) {
      uint32_t run = 0;
      for (int w = 0; w < SX_THREADS / 64; ++w) {
        const uint32_t v = wsum[w];
        wsum[w] = run;
        run += v;
      }
      wsum[SX_THREADS / 64] = run;
    }
    __syncthreads();
    uint32_t o = wsum[wave] + inc - cnt;
    m = wsum[SX_THREADS / 64];
    const double e[4] = {ext[0], ext[1], ext[2], ext[3]};
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t k = tid * PER + u;
      if (k >= n) continue;
      const double x = xs[node0 + k], y = ys[node0 + k];
      if (x != x || y != y) continue;
      double x0 = e[0], y0 = e[1], x1 = e[2], y1 = e[3];
      uint64_t key = 0;
      for (int lev = 0; lev < KEY_LEVELS; ++lev) {
        const double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
        const bool right = x >= xm;
        if (right) x0 = xm; else x1 = xm;
        const bool bottom = y >= ym;
        if (bottom) y0 = ym; else y1 = ym;
        key = key << 2 | (uint64_t)bottom << 1 | (uint64_t)right;
      }
      keys[o] = key;
      perm[o] = (uint16_t)k;
      ++o;
    }
    uint32_t size = 1;
    while (size < m) size <<= 1;
    for (uint32_t k = m + tid; k < size; k += SX_THREADS) {
      keys[k] = ~0ull;
      perm[k] = 0xffff;
    }
    __syncthreads();
    // bitonic sort of (key, node): the node breaks ties, so equal keys keep
    // their index order (the stable LSD sort's result, sort.zig:55-91)
    for (uint32_t kk = 2; kk <= size; kk <<= 1) {
      for (uint32_t j = kk >> 1; j > 0; j >>= 1) {
        for (uint32_t t = tid; t < size / 2; t += SX_THREADS) {
          const uint32_t a = 2 * t - (t & (j - 1)), b = a + j;
          const bool up = (a & kk) == 0;
          const uint64_t ka = keys[a], kb = keys[b];
          const uint16_t pa = perm[a], pb = perm[b];
          const bool gt = ka > kb || (ka == kb && pa > pb);
          if (gt == up) {
            keys[a] = kb;
            keys[b] = ka;
            perm[a] = pb;
            perm[b] = pa;
          }
        }
        __syncthreads();
      }
    }
    // 3a. run boundaries: diff[k] = positions <= k differing from their predecessor
    uint32_t dc = 0;
    uint32_t dv[PER];
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t k = tid * PER + u;
      uint32_t d = 0;
      if (k > 0 && k < m) {
        const uint32_t p = node0 + perm[k], q = node0 + perm[k - 1];
        d = (xs[p] != xs[q] || ys[p] != ys[q]) ? 1u : 0u;
      }
      dv[u] = d;
      dc += d;
    }
    inc = dc;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(inc, d, 64);
      if (lane >= d) inc += t;
    }
    __syncthreads();
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    if (tid == 0) {
      uint32_t run = 0;
      for (int w = 0; w < SX_THREADS / 64; ++w) {
        const uint32_t v = wsum[w];
        wsum[w] = run;
        run += v;
      }
    }
    __syncthreads();
    uint32_t run = wsum[wave] + inc - dc;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t k = tid * PER + u;
      run += dv[u];
      if (k < m) diff[k] = (uint16_t)run;
    }
    __syncthreads();
    // 3b. the preorder emit (tree.zig:437-647). Below 32 levels a cell is a
    // key prefix: position k starts the cells of depths first(k)..D(k), where
    // first(k) is one more than the levels that keys k - 1 and k share (0 at
    // k = 0) and D(k) the depth from which k's cell holds coincident points
    // only (a leaf). Preorder takes the positions in order and a position's
    // cells from the shallowest, so a cell's index is the count of the earlier
    // positions' cells plus its depth offset. When points share all 32 key
    // levels without coinciding they are split below the keys (a midpoint
    // partition), and the whole emit runs on one thread in preorder.
    auto shared_levels = [&](uint32_t j) -> int {  // levels keys[j - 1] and keys[j] share
      const uint64_t d = keys[j - 1] ^ keys[j];
      return d == 0 ? KEY_LEVELS : (int)(__builtin_clzll(d) >> 1);
    };
    int deep = 0;
    int first_k[PER], leaf_k[PER];
    uint32_t cells_here = 0;
#pragma unroll
    for (uint32_t u = 0; u < PER; ++u) {
      const uint32_t k = tid * PER + u;
      first_k[u] = 1;
      leaf_k[u] = 0;
      if (k >= m) continue;
      const int c = k == 0 ? -1 : shared_levels(k);
      if (c == KEY_LEVELS && diff[k] != diff[k - 1]) deep = 1;
      const uint16_t dk = diff[k];
      uint32_t lo = 0, hi = k;  // the run of points coinciding with k: [lo, rh)
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (diff[mid] < dk) lo = mid + 1; else hi = mid;
      }
      const uint32_t rl = lo;
      lo = k + 1;
      hi = m;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (diff[mid] <= dk) lo = mid + 1; else hi = mid;
      }
      const uint32_t rh = lo;
      int top = -1;
      if (rl > 0) top = max(top, shared_levels(rl));
      if (rh < m) top = max(top, shared_levels(rh));
      first_k[u] = c + 1;
      leaf_k[u] = top + 1;
      if (leaf_k[u] >= first_k[u]) cells_here += (uint32_t)(leaf_k[u] - first_k[u] + 1);
    }
    deep = __syncthreads_or(deep);
    if (!deep) {
      // exclusive scan of the cells per position; a position's first cell index in ptmp
      uint32_t incl = cells_here;
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
      }
      if (lane == 63) wsum[wave] = incl;
      if (tid == 0) meta[2] = 0;
      __syncthreads();
      if (tid == 0) {
        uint32_t run = 0;
        for (int w = 0; w < SX_THREADS / 64; ++w) {
          const uint32_t v = wsum[w];
          wsum[w] = run;
          run += v;
        }
        wsum[SX_THREADS / 64] = run;
      }
      __syncthreads();
      const uint32_t total = wsum[SX_THREADS / 64];
      uint32_t base = wsum[wave] + incl - cells_here;
      int deepest = 0;
#pragma unroll
      for (uint32_t u = 0; u < PER; ++u) {
        const uint32_t k = tid * PER + u;
        if (k < m) ptmp[k] = (uint16_t)base;
        if (k < m && leaf_k[u] >= first_k[u]) base += (uint32_t)(leaf_k[u] - first_k[u] + 1);
        if (k < m) deepest = max(deepest, leaf_k[u]);
      }
      atomicMax(&meta[2], (uint32_t)deepest);
      const bool over = total > job.cell_cap;
      if (!over) {
#pragma unroll
        for (uint32_t u = 0; u < PER; ++u) {
          const uint32_t k = tid * PER + u;
          if (k >= m || leaf_k[u] < first_k[u]) continue;
          const uint64_t key = keys[k];
          double x0 = ext[0], y0 = ext[1], x1 = ext[2], y1 = ext[3];
          uint32_t idx = ptmp[k];
          for (int d = 0; d <= leaf_k[u]; ++d) {
            if (d >= first_k[u]) {
              uint32_t hi = m;  // end of the cell: the first later key with another d-level prefix
              if (d > 0) {
                const unsigned sh = 64 - 2 * d;
                const uint64_t pre = key >> sh;
                uint32_t lo = k + 1;
                while (lo < hi) {
                  const uint32_t mid = (lo + hi) >> 1;
                  if ((keys[mid] >> sh) == pre) lo = mid + 1; else hi = mid;
                }
              }
              SCell c;
              c.x0 = x0; c.y0 = y0; c.x1 = x1; c.y1 = y1;
              c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
              c.first = k;
              c.count = hi - k;
              c.depth = (uint32_t)d;
              c.pad = 0;
              cells[idx++] = c;
            }
            if (d == leaf_k[u]) break;
            // the child on key's path: the emit's midpoint split
            const unsigned q = (unsigned)(key >> (62 - 2 * d)) & 3u;
            const double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
            if (q & 1) x0 = xm; else x1 = xm;
            if (q & 2) y0 = ym; else y1 = ym;
          }
        }
      }
      __syncthreads();
      if (!over) {
        // each cell below the root links itself into its parent's slot
#pragma unroll
        for (uint32_t u = 0; u < PER; ++u) {
          const uint32_t k = tid * PER + u;
          if (k >= m || leaf_k[u] < first_k[u]) continue;
          const uint64_t key = keys[k];
          const uint32_t base_k = ptmp[k];
          for (int d = max(first_k[u], 1); d <= leaf_k[u]; ++d) {
            const uint32_t idx = base_k + (uint32_t)(d - first_k[u]);
            uint32_t parent;
            if (d > first_k[u]) {
              parent = idx - 1;
            } else if (d == 1) {
              parent = 0;
            } else {  // the parent starts at the first position of k's (d - 1)-level prefix
              const unsigned sh = 64 - 2 * (d - 1);
              const uint64_t pre = key >> sh;
              uint32_t lo = 0, hi = k;
              while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if ((keys[mid] >> sh) < pre) lo = mid + 1; else hi = mid;
              }
              const int pf = lo == 0 ? 0 : shared_levels(lo) + 1;
              parent = ptmp[lo] + (uint32_t)(d - 1 - pf);
            }
            cells[parent].child[(key >> (64 - 2 * d)) & 3u] = idx;
          }
        }
      }
      if (tid == 0) {
        meta[0] = m;
        meta[1] = total;
        meta[3] = over ? 1u : 0u;
        if (over) atomicOr(error, 1u);
      }
      __syncthreads();
    }
    if (deep && tid == 0) {
      uint32_t sp = 0, idx = 0, maxd = 0;
      bool over = false;
      stack[sp++] = SFrame{ext[0], ext[1], ext[2], ext[3], NONE, 0, (uint16_t)m, 0, 0, 1};
      while (sp > 0 && !over) {
        const SFrame f = stack[--sp];
        if (idx >= job.cell_cap || f.depth >= SX_DEPTHS) {
          over = true;
          break;
        }
        SCell c;
        c.x0 = f.x0; c.y0 = f.y0; c.x1 = f.x1; c.y1 = f.y1;
        c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
        c.first = f.lo;
        c.count = (uint32_t)f.hi - f.lo;
        c.depth = f.depth;
        c.pad = 0;
        cells[idx] = c;
        if (f.parent != NONE) cells[f.parent].child[f.slot] = idx;
        const uint32_t me = idx++;
        if (f.depth > maxd) maxd = f.depth;
        if (f.depth >= MAX_DEPTH) continue;
        // allSame (tree.zig:437-444)
        bool same;
        if (f.sorted) {
          same = diff[f.hi - 1] == diff[f.lo];
        } else {
          same = true;
          const uint32_t p0 = node0 + perm[f.lo];
          for (uint32_t k = f.lo + 1; k < f.hi && same; ++k) {
            const uint32_t p = node0 + perm[k];
            same = !(xs[p] != xs[p0] || ys[p] != ys[p0]);
          }
        }
        if (same) continue;
        // split (tree.zig:449-479)
        uint32_t out[5];
        out[0] = f.lo;
        out[4] = f.hi;
        const bool sorted = f.sorted && f.depth < (uint32_t)KEY_LEVELS;
        const double xm = (f.x0 + f.x1) / 2, ym = (f.y0 + f.y1) / 2;
        if (sorted) {
          const unsigned shift = 62 - 2 * f.depth;
          for (uint32_t q = 1; q < 4; ++q) {  // first position whose digit is >= q
            uint32_t lo = out[q - 1], hi = f.hi;
            while (lo < hi) {
              const uint32_t mid = (lo + hi) >> 1;
              if (((keys[mid] >> shift) & 3) < q) lo = mid + 1; else hi = mid;
            }
            out[q] = lo;
          }
        } else {
          uint32_t counts[4] = {0, 0, 0, 0};
          for (uint32_t k = f.lo; k < f.hi; ++k) {
            const uint32_t p = node0 + perm[k];
            counts[quadrant_of(xs[p], ys[p], xm, ym)]++;
          }
          uint32_t off[4], sum = f.lo;
          for (int q = 0; q < 4; ++q) { off[q] = sum; out[q] = sum; sum += counts[q]; }
          for (uint32_t k = f.lo; k < f.hi; ++k) {
            const uint32_t p = node0 + perm[k];
            ptmp[off[quadrant_of(xs[p], ys[p], xm, ym)]++] = perm[k];
          }
          for (uint32_t k = f.lo; k < f.hi; ++k) perm[k] = ptmp[k];
        }
        for (int q = 3; q >= 0; --q) {
          if (out[q] == out[q + 1]) continue;
          if (sp == (uint32_t)SX_STACK) {
            over = true;
            break;
          }
          const bool right = (q & 1) != 0, bottom = (q & 2) != 0;
          stack[sp++] = SFrame{right ? xm : f.x0, bottom ? ym : f.y0, right ? f.x1 : xm, bottom ? f.y1 : ym, me,
                               (uint16_t)out[q], (uint16_t)out[q + 1], (uint16_t)(f.depth + 1), (uint8_t)q,
                               (uint8_t)(sorted ? 1 : 0)};
        }
      }
      meta[0] = m;
      meta[1] = idx;
      meta[2] = maxd;
      meta[3] = over ? 1u : 0u;
      if (over) atomicOr(error, 1u);
    }
    __syncthreads();
  } else if (tid == 0) {
    meta[0] = 0;
    meta[1] = 0;
    meta[2] = 0;
  }
  __syncthreads();
  if (meta[3]) {  // the graph's tree outgrew this kernel: the tick fails, velocities kept
    for (uint32_t k = tid; k < n; k += SX_THREADS) {
      vout_x[node0 + k] = vxs[node0 + k];
      vout_y[node0 + k] = vys[node0 + k];
    }
    return;
  }
  const uint32_t len = meta[1];
  m = meta[0];

  // 4. accumulate, deepest level first (many_body.zig:101-135): cells
  // grouped by depth (counting sort in LDS), one level per step
  for (uint32_t d = tid; d < SX_DEPTHS + 2; d += SX_THREADS) dstart[d] = 0;
  __syncthreads();
  for (uint32_t c = tid; c < len; c += SX_THREADS) atomicAdd(&dstart[cells[c].depth + 1], 1u);
  __syncthreads();
  if (tid == 0)
    for (uint32_t d = 1; d < SX_DEPTHS + 2; ++d) dstart[d] += dstart[d - 1];
  __syncthreads();
  for (uint32_t c = tid; c < len; c += SX_THREADS) {
    const uint32_t slot = atomicAdd(&dstart[cells[c].depth], 1u);
    bycell[slot] = (uint16_t)c;
  }
  __syncthreads();
  // dstart[d] now ends depth d's cells; depth d starts at dstart[d - 1] (0 for d = 0)
  for (int d = (int)meta[2]; d >= 0 && len > 0; --d) {
    const uint32_t lo = d > 0 ? dstart[d - 1] : 0, hi = dstart[d];
    for (uint32_t q = lo + tid; q < hi; q += SX_THREADS) {
      const uint32_t c = bycell[q];
      const SCell cell = cells[c];
      TCell t;
      const double w = cell.x1 - cell.x0;
      t.w2t = w * w / theta2;
      t.first = cell.first;
      t.count = cell.count;
      t.pad[0] = t.pad[1] = t.pad[2] = 0;
      if (cell.child[0] == NONE && cell.child[1] == NONE && cell.child[2] == NONE && cell.child[3] == NONE) {
        const uint32_t head = node0 + perm[cell.first + cell.count - 1];
        double strength = 0;
        for (uint32_t j = cell.first + cell.count; j > cell.first;) {
          --j;
          strength += charge[node0 + perm[j]];
        }
        t.qx = xs[head];
        t.qy = ys[head];
        t.value = strength;
        t.skip = 1;
        t.head = head;
        t.flags = TC_LEAF | (truthy(strength) ? TC_REACH : 0u);
      } else {
        double st = 0, weight = 0, x = 0, y = 0;
        bool reach = false;
        uint32_t skip = 1;
        for (int q = 0; q < 4; ++q) {
          const uint32_t ch = cell.child[q];
          if (ch == NONE) continue;
          const TCell& k = tc[ch];
          skip += k.skip;
          const double v = k.value;
          const double wv = fabs(v);
          if (wv > 0) {
            st += v;
            weight += wv;
            x += wv * k.qx;
            y += wv * k.qy;
          }
          reach = reach || (k.flags & TC_REACH);
        }
        t.qx = x / weight;
        t.qy = y / weight;
        t.value = st;
        t.skip = skip;
        t.head = NONE;
        t.flags = (reach && truthy(st)) ? TC_REACH : 0u;
      }
      tc[c] = t;
    }
    __syncthreads();
  }

  // 5. the walks: nodes in Morton order, SX_WALKS per lane at a time; a walk
  // that would draw stops and its node is flagged
  for (uint32_t k = tid; k < SX_MAX / 32; k += SX_THREADS) flagged[k] = 0;
  __syncthreads();
  for (uint32_t k0 = tid; k0 < m; k0 += SX_THREADS * SX_WALKS) {
    SWalk w[SX_WALKS];
#pragma unroll
    for (int q = 0; q < SX_WALKS; ++q) {
      const uint32_t slot = k0 + q * SX_THREADS;
      const bool active = slot < m;
      SWalk& a = w[q];
      a.i = node0 + (active ? perm[slot] : 0u);
      a.nx = active ? xs[a.i] : 0;
      a.ny = active ? ys[a.i] : 0;
      a.vx = active ? vxs[a.i] : 0;
      a.vy = active ? vys[a.i] : 0;
      a.c = active ? 0 : len;
      a.drew = false;
    }
    for (;;) {
      bool any = false;
#pragma unroll
      for (int q = 0; q < SX_WALKS; ++q) any = any || w[q].c < len;
      if (!any) break;
      uint4 g[SX_WALKS][3];
#pragma unroll
      for (int q = 0; q < SX_WALKS; ++q) {
        const uint4* r = reinterpret_cast<const uint4*>(tc + (w[q].c < len ? w[q].c : 0));
        g[q][0] = r[0];
        g[q][1] = r[1];
        g[q][2] = r[2];
      }
#pragma unroll
      for (int q = 0; q < SX_WALKS; ++q)
        if (w[q].c < len) step(w[q], g[q][0], g[q][1], g[q][2], len, perm, node0, charge, dmin2, dmax2, alpha);
    }
#pragma unroll
    for (int q = 0; q < SX_WALKS; ++q) {
      if (k0 + q * SX_THREADS >= m) continue;
      const uint32_t local = w[q].i - node0;
      if (!w[q].drew) {
        vout_x[w[q].i] = w[q].vx;
        vout_y[w[q].i] = w[q].vy;
      } else {
        atomicOr(&flagged[local >> 5], 1u << (local & 31));
      }
    }
  }
  // nodes with a NaN coordinate reach every truthy leaf (v becomes NaN when
  // the root reaches one); with no tree they keep their velocity
  const bool reach = len > 0 && (tc[0].flags & TC_REACH);
  for (uint32_t k = tid; k < n; k += SX_THREADS) {
    const uint32_t i = node0 + k;
    const double x = xs[i], y = ys[i];
    if (x == x && y == y) continue;
    const double nan = __builtin_nan("");
    vout_x[i] = reach ? vxs[i] + nan : vxs[i];
    vout_y[i] = reach ? vys[i] + nan : vys[i];
  }
  __syncthreads();

  // 6. on one thread: the seed draw, the flagged nodes in ascending order,
  // the link seed draw
  __shared__ uint32_t lcg_sh;
  if (tid == 0) {
    uint32_t lcg = lcg_state[job.graph];
    (void)lcg_next(lcg);  // manyBody's seed (core.js:218-222)
    for (uint32_t w = 0; w < (n + 31) / 32; ++w) {
      uint32_t bits = flagged[w];
      while (bits) {
        const uint32_t b = __builtin_ctz(bits);
        bits &= bits - 1;
        const uint32_t i = node0 + w * 32 + b;
        double vx = vxs[i], vy = vys[i];
        walk<1>(i, xs[i], ys[i], vx, vy, tc, len, perm, node0, charge, dmin2, dmax2, alpha, &lcg);
        vout_x[i] = vx;
        vout_y[i] = vy;
      }
    }
    (void)lcg_next(lcg);  // link's seed
    lcg_sh = lcg;
  }
  __syncthreads();
  if (job.m == 0 || !(0 < link_iterations)) {
    if (tid == 0) lcg_state[job.graph] = lcg_sh;
    return;
  }

  // 7. the link (link.zig:8-40) on the velocities in LDS, a level of the
  // schedule at a time: a level's links share no node and read what the
  // earlier levels wrote, as in link order. A link that would draw (a zero or
  // NaN difference) stops this, and the passes are redone in link order on
  // one thread from the velocities before the link.
  double* lvx = u.l.vx;
  double* lvy = u.l.vy;
  for (uint32_t k = tid; k < n; k += SX_THREADS) {
    lvx[k] = vout_x[node0 + k];
    lvy[k] = vout_y[node0 + k];
  }
  __syncthreads();
  const SmallLink* recs = lrec + job.rec0;
  const uint32_t* lev = levels + job.lev0;
  const double* gx = xs + node0;
  const double* gy = ys + node0;
  bool draws = false;
  for (double it = 0; it < link_iterations && !draws; it += 1) {
    for (uint32_t level = 0; level < job.nlev; ++level) {
      const uint32_t hi = lev[level + 1];
      int stop = 0;
      for (uint32_t q = lev[level] + tid; q < hi; q += SX_THREADS) {
        const SmallLink r = recs[q];
        const uint32_t a = r.a, b = r.b;
        double x = gx[b] + lvx[b] - gx[a] - lvx[a];
        double y = gy[b] + lvy[b] - gy[a] - lvy[a];
        if (x == 0 || x != x || y == 0 || y != y) {
          stop = 1;
          continue;
        }
        double l = __builtin_sqrt(x * x + y * y);
        l = (l - r.distance) / l * alpha * r.strength;
        x *= l;
        y *= l;
        lvx[b] -= x * r.bias;
        lvy[b] -= y * r.bias;
        const double ob = 1 - r.bias;
        lvx[a] += x * ob;
        lvy[a] += y * ob;
      }
      if (__syncthreads_or(stop)) {
        draws = true;
        break;
      }
    }
  }
  if (draws && tid == 0) {
    uint32_t lcg = lcg_sh;
    for (uint32_t k = 0; k < n; ++k) {
      lvx[k] = vout_x[node0 + k];
      lvy[k] = vout_y[node0 + k];
    }
    for (double it = 0; it < link_iterations; it += 1) {
      const uint32_t e0 = job.link0, e1 = job.link0 + job.m;
      uint32_t a = src[e0] - node0, b = tgt[e0] - node0;
      double xa = gx[a], xb = gx[b], ya = gy[a], yb = gy[b];
      double dist = ldistance[e0], kstr = lstrength[e0], bias = lbias[e0];
      for (uint32_t e = e0; e < e1; ++e) {  // an edge's inputs no earlier edge writes are read one edge ahead
        uint32_t an = 0, bn = 0;
        double xan = 0, xbn = 0, yan = 0, ybn = 0, distn = 0, kn = 0, biasn = 0;
        if (e + 1 < e1) {
          an = src[e + 1] - node0;
          bn = tgt[e + 1] - node0;
          xan = gx[an];
          xbn = gx[bn];
          yan = gy[an];
          ybn = gy[bn];
          distn = ldistance[e + 1];
          kn = lstrength[e + 1];
          biasn = lbias[e + 1];
        }
        double x = xb + lvx[b] - xa - lvx[a];
        if (x == 0 || x != x) x = jiggle_of(lcg_next(lcg));
        double y = yb + lvy[b] - ya - lvy[a];
        if (y == 0 || y != y) y = jiggle_of(lcg_next(lcg));
        double l = __builtin_sqrt(x * x + y * y);
        l = (l - dist) / l * alpha * kstr;
        x *= l;
        y *= l;
        lvx[b] -= x * bias;
        lvy[b] -= y * bias;
        const double ob = 1 - bias;
        lvx[a] += x * ob;
        lvy[a] += y * ob;
        a = an;
        b = bn;
        xa = xan;
        xb = xbn;
        ya = yan;
        yb = ybn;
        dist = distn;
        kstr = kn;
        bias = biasn;
      }
    }
    lcg_sh = lcg;
  }
  __syncthreads();
  if (tid == 0) lcg_state[job.graph] = lcg_sh;
  for (uint32_t k = tid; k < n; k += SX_THREADS) {
    vout_x[node0 + k] = lvx[k];
    vout_y[node0 + k] = lvy[k];
  }
}

// The link schedule of one small graph, one workgroup a graph. On one thread,
// in link order: a link's level is the highest level of the earlier links at
// either end, plus one (from 0), counted per level; the links are then placed
// by level, in link order within a level; the records are packed in parallel.
// The serial passes hold one level per node in LDS.
__global__ __launch_bounds__(256) void k_small_schedule(SmallJob* __restrict__ jobs, const uint32_t* __restrict__ src,
                                                        const uint32_t* __restrict__ tgt,
                                                        const double* __restrict__ lstrength,
                                                        const double* __restrict__ ldistance,
                                                        const double* __restrict__ lbias, uint32_t* __restrict__ order,
                                                        uint32_t* __restrict__ levels, SmallLink* __restrict__ lrec) {
  __shared__ uint32_t last[SX_MAX];
  const SmallJob job = jobs[blockIdx.x];
  const uint32_t tid = threadIdx.x, m = job.m;
  uint32_t* lev = levels + job.lev0;  // m + 1 entries
  uint32_t* ord = order + job.rec0;   // m entries
  // a link's local ends; a range the host did not check reads node 0
  auto ends = [&](uint32_t e, uint32_t& a, uint32_t& b) {
    a = src[job.link0 + e] - job.node0;
    b = tgt[job.link0 + e] - job.node0;
    if (a >= job.n || a >= SX_MAX) a = 0;
    if (b >= job.n || b >= SX_MAX) b = 0;
  };
  for (uint32_t k = tid; k < SX_MAX; k += 256) last[k] = 0;
  for (uint32_t k = tid; k <= m; k += 256) lev[k] = 0;
  __syncthreads();
  if (tid == 0) {
    uint32_t nlev = 0;
    for (uint32_t e = 0; e < m; ++e) {
      uint32_t a, b;
      ends(e, a, b);
      const uint32_t level = last[a] > last[b] ? last[a] : last[b];
      last[a] = last[b] = level + 1;
      lev[level + 1] += 1;
      if (level + 1 > nlev) nlev = level + 1;
    }
    for (uint32_t k = 1; k <= nlev; ++k) lev[k] += lev[k - 1];  // lev[L]: where level L starts
    for (uint32_t k = 0; k < job.n && k < SX_MAX; ++k) last[k] = 0;
    last[0] = 0;  // also where out-of-range ends were counted
    for (uint32_t e = 0; e < m; ++e) {
      uint32_t a, b;
      ends(e, a, b);
      const uint32_t level = last[a] > last[b] ? last[a] : last[b];
      last[a] = last[b] = level + 1;
      ord[lev[level]++] = e;  // lev[L] ends up where level L + 1 starts
    }
    for (uint32_t k = nlev; k > 0; --k) lev[k] = lev[k - 1];
    lev[0] = 0;
    jobs[blockIdx.x].nlev = nlev;
  }
  __syncthreads();
  for (uint32_t q = tid; q < m; q += 256) {
    const uint32_t e = ord[q];
    uint32_t a, b;
    ends(e, a, b);
    SmallLink r;
    r.a = (uint16_t)a;
    r.b = (uint16_t)b;
    r.pad = 0;
    r.distance = ldistance[job.link0 + e];
    r.strength = lstrength[job.link0 + e];
    r.bias = lbias[job.link0 + e];
    lrec[job.rec0 + q] = r;
  }
}

__global__ __launch_bounds__(256) void k_small_scatter(const SmallJob* __restrict__ jobs,
                                                       const double* __restrict__ vout_x,
                                                       const double* __restrict__ vout_y, double* __restrict__ vx,
                                                       double* __restrict__ vy) {
  const SmallJob job = jobs[blockIdx.x];
  for (uint32_t k = threadIdx.x; k < job.n; k += 256) {
    vx[job.node0 + k] = vout_x[job.node0 + k];
    vy[job.node0 + k] = vout_y[job.node0 + k];
  }
}

}  // namespace

size_t small_cell_bytes() { return sizeof(SCell); }

hipError_t small_exact(const SmallJob* jobs, uint32_t count, const SmallExactArgs& a, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_small_exact, dim3(count), dim3(SX_THREADS), 0, stream, jobs, a.x, a.y, a.vx, a.vy, a.charge,
                     a.src, a.tgt, a.lstrength, a.ldistance, a.lbias, a.lrec, a.levels, a.alpha_dev, a.theta2,
                     a.dmin2, a.dmax2,
                     a.link_iterations, a.lcg, reinterpret_cast<SCell*>(a.cells), a.tcells, a.vout_x, a.vout_y,
                     a.error);
  return hipGetLastError();
}

hipError_t small_schedule(SmallJob* jobs, uint32_t count, const uint32_t* src, const uint32_t* tgt,
                          const double* lstrength, const double* ldistance, const double* lbias, uint32_t* order,
                          uint32_t* levels, SmallLink* lrec, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_small_schedule, dim3(count), dim3(256), 0, stream, jobs, src, tgt, lstrength, ldistance, lbias,
                     order, levels, lrec);
  return hipGetLastError();
}

hipError_t small_scatter(const SmallJob* jobs, uint32_t count, const double* vout_x, const double* vout_y, double* vx,
                         double* vy, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_small_scatter, dim3(count), dim3(256), 0, stream, jobs, vout_x, vout_y, vx, vy);
  return hipGetLastError();
}

}  // namespace d3f
