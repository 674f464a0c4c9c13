// forceManyBody, exact mode, on gfx950 (reference: src/core/many_body.zig,
// itself d3-force's manyBody.js over the quadtree).
//
// Accumulate (visitAfter): a leaf sums its chain from the head down
// (many_body.zig:125-132) and an internal cell combines children 0..3 in
// order (many_body.zig:112-121), so every sum has d3's operands in d3's
// order. Cells are numbered in preorder, so a subtree is a contiguous range
// whose size the tree build records; subtrees of at most ACC_SUB cells are
// summed by one workgroup each in LDS, depth by depth under workgroup
// barriers (k_acc_subtrees), and the few cells above them by one workgroup
// (k_acc_top). Kernel boundaries order the two, with no inter-workgroup
// atomics (whose cache maintenance spans all eight XCDs' L2s). Trees with
// cells below key resolution (whose serial split records no subtree sizes)
// go level by level, one launch per depth (k_accumulate_level).
//
// Traversal (visit): one lane per node, nodes in Morton order. Cells are
// numbered in d3's visit preorder (children 0..3), so a node's visit sequence
// needs no stack: from cell c the walk goes on to c + 1 when it opens c (its
// first child) and to c + skip (the cell after c's subtree) when it does not
// (an empty cell, an accepted cell, a leaf). Each lane walks its own node's
// sequence (many_body.zig:190-259) and accumulates its velocity in d3's
// order; c only grows, so every walk ends within the cell count. Lanes of a
// wave move through the cells in the same direction, so the records they read
// stay close in the caches; blocks are mapped to the eight XCDs in
// contiguous ranges of nodes, so each XCD's L2 serves one region of the tree.
//
// Jiggle (many_body.zig:65-83): with RECORD a lane that meets a zero delta
// stops without writing and is listed; the host then sorts the list, COUNT
// gives each listed node's number of draws (fixed by positions and the tree
// alone), the host draws them from the JS random source in ascending node
// order, and REPLAY recomputes those nodes consuming their draws in order.
#include "common.h"

namespace d3f {
namespace {

constexpr int ACC_THREADS = 256;
constexpr int TRAV_THREADS = 256;

__device__ inline bool truthy(double v) { return v != 0 && v == v; }

// A cell's traversal record. The 88-byte cell's x, y, value receive the same
// charges only when the tree is exported (k_publish_charges).
__device__ inline void put(TCell* __restrict__ tc, uint32_t c, const Cell& cell, double qx, double qy, double value,
                           double theta2, uint32_t skip, uint32_t head, uint32_t flags) {
  TCell t;
  t.qx = qx;
  t.qy = qy;
  t.value = value;
  const double w = cell.x1 - cell.x0;
  t.w2t = w * w / theta2;  // many_body.zig:202, identical operands per cell
  t.skip = skip;
  t.first = cell.first;
  t.count = cell.count;
  t.head = head;
  t.flags = flags;
  t.pad[0] = t.pad[1] = t.pad[2] = 0;
  tc[c] = t;
}

__global__ __launch_bounds__(ACC_THREADS) void k_publish_charges(Cell* __restrict__ cells,
                                                                 const TCell* __restrict__ tc, uint32_t len) {
  const uint32_t c = blockIdx.x * ACC_THREADS + threadIdx.x;
  if (c >= len) return;
  cells[c].x = tc[c].qx;
  cells[c].y = tc[c].qy;
  cells[c].value = tc[c].value;
}

// Cells grouped by depth (a counting sort): histogram, scan, scatter. The
// order inside a depth does not matter: cells of one depth never read each
// other.
constexpr int DEPTH_BINS = MAX_DEPTH + 64;

__global__ __launch_bounds__(ACC_THREADS) void k_depth_hist(const uint32_t* __restrict__ cdepth, uint32_t len,
                                                            uint32_t nbins, uint32_t* __restrict__ bins) {
  extern __shared__ uint32_t h[];
  for (uint32_t b = threadIdx.x; b < nbins; b += ACC_THREADS) h[b] = 0;
  __syncthreads();
  for (uint32_t c = blockIdx.x * ACC_THREADS + threadIdx.x; c < len; c += gridDim.x * ACC_THREADS)
    atomicAdd(&h[cdepth[c]], 1u);
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += ACC_THREADS)
    if (h[b]) atomicAdd(&bins[b], h[b]);
}

// Exclusive scan of the bins into start[0..nbins] and cursor (one thread:
// at most a few thousand bins).
__global__ void k_depth_scan(uint32_t* __restrict__ bins, uint32_t nbins, uint32_t* __restrict__ start) {
  if (threadIdx.x != 0) return;
  uint32_t run = 0;
  for (uint32_t b = 0; b < nbins; ++b) {
    const uint32_t v = bins[b];
    start[b] = run;
    bins[b] = run;  // becomes the scatter cursor
    run += v;
  }
  start[nbins] = run;
}

// A block places SCATTER_PER cells per thread with one global atomic per
// depth bin (the bins are few, so per-cell-block atomics would contend).
constexpr int SCATTER_PER = 16;

__global__ __launch_bounds__(ACC_THREADS) void k_depth_scatter(const uint32_t* __restrict__ cdepth, uint32_t len,
                                                               uint32_t nbins, uint32_t* __restrict__ cursor,
                                                               uint32_t* __restrict__ order) {
  extern __shared__ uint32_t h[];  // [nbins] local counts, then local bases
  const uint32_t base = blockIdx.x * ACC_THREADS * SCATTER_PER + threadIdx.x;
  for (uint32_t b = threadIdx.x; b < nbins; b += ACC_THREADS) h[b] = 0;
  __syncthreads();
  uint32_t d[SCATTER_PER], local[SCATTER_PER];
#pragma unroll
  for (int q = 0; q < SCATTER_PER; ++q) {
    const uint32_t c = base + q * ACC_THREADS;
    d[q] = 0;
    local[q] = 0;
    if (c < len) {
      d[q] = cdepth[c];
      local[q] = atomicAdd(&h[d[q]], 1u);
    }
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < nbins; b += ACC_THREADS)
    if (h[b]) h[b] = atomicAdd(&cursor[b], h[b]);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < SCATTER_PER; ++q) {
    const uint32_t c = base + q * ACC_THREADS;
    if (c < len) order[h[d[q]] + local[q]] = c;
  }
}

// One depth of visitAfter: leaves sum their chain from the head down
// (many_body.zig:125-132); internal cells combine children 0..3, whose depth
// was done by the previous launch (many_body.zig:112-121).
__global__ __launch_bounds__(ACC_THREADS) void k_accumulate_level(
    const Cell* __restrict__ cells, TCell* __restrict__ tc, const uint32_t* __restrict__ order,
    const uint32_t* __restrict__ start, uint32_t depth, const uint32_t* __restrict__ perm,
    const double* __restrict__ xs, const double* __restrict__ ys, const double* __restrict__ s, double theta2) {
  const uint32_t lo = start[depth], hi = start[depth + 1];
  for (uint32_t k = lo + blockIdx.x * ACC_THREADS + threadIdx.x; k < hi; k += gridDim.x * ACC_THREADS) {
    const uint32_t c = order[k];
    const Cell cell = cells[c];
    if (cell.child[0] == NONE && cell.child[1] == NONE && cell.child[2] == NONE && cell.child[3] == NONE) {
      const uint32_t head = perm[cell.first + cell.count - 1];
      double strength = 0;
      for (uint32_t j = cell.first + cell.count; j > cell.first;) {
        --j;
        strength += s[perm[j]];
      }
      put(tc, c, cell, xs[head], ys[head], strength, theta2, 1, head,
          TC_LEAF | (truthy(strength) ? TC_REACH : 0));
      continue;
    }
    double st = 0, weight = 0, x = 0, y = 0;
    bool reach = false;
    uint32_t skip = 1;
    for (int q = 0; q < 4; ++q) {
      const uint32_t ch = cell.child[q];
      if (ch == NONE) continue;
      skip += tc[ch].skip;
      const double v = tc[ch].value;
      const double w = fabs(v);
      if (w > 0) {  // `c = Math.abs(q.value)` is truthy
        st += v;
        weight += w;
        x += w * tc[ch].qx;
        y += w * tc[ch].qy;
      }
      reach = reach || (tc[ch].flags & TC_REACH);
    }
    put(tc, c, cell, x / weight, y / weight, st, theta2, skip, NONE, (reach && truthy(st)) ? TC_REACH : 0);
  }
}

// visitAfter over the subtrees in `cut` (at most ACC_SUB cells each, a
// contiguous preorder range [r, r + skip)), one subtree per workgroup at a
// time: leaves first (many_body.zig:125-132), then the internal cells one
// depth at a time from the deepest, children read from LDS (many_body.zig:
// 112-121). Each thread keeps its cells' children, depth and record fields
// in registers; the sums have d3's operands in d3's order.
constexpr int SUB_THREADS = 256;

// A subtree ends inside the tree (1 <= skip <= len - c by construction); the
// clamp keeps every walk moving forward whatever the arrays hold.
__device__ inline uint32_t clamp_skip(uint32_t skip, uint32_t c, uint32_t len) {
  return skip >= 1 && skip <= len - c ? skip : len - c;
}
constexpr int SUB_PER = (int)ACC_SUB / SUB_THREADS;

__global__ __launch_bounds__(SUB_THREADS) void k_acc_subtrees(
    const Cell* __restrict__ cells, TCell* __restrict__ tc, const uint2* __restrict__ cut,
    const uint32_t* __restrict__ cskip, const uint32_t* __restrict__ cdepth, const TreeMeta* __restrict__ meta,
    const uint32_t* __restrict__ perm, const double* __restrict__ xs, const double* __restrict__ ys,
    const double* __restrict__ s, double theta2) {
  __shared__ double sx[ACC_SUB], sy[ACC_SUB], sv[ACC_SUB];
  __shared__ uint8_t sreach[ACC_SUB];
  __shared__ uint32_t smaxd;
  const uint32_t tid = threadIdx.x;
  const uint32_t ncut = meta->ncut, len = meta->len;
  if (tid == 0) smaxd = 0;
  __syncthreads();
  for (uint32_t j = blockIdx.x; j < ncut; j += gridDim.x) {
    const uint2 e = cut[j];  // root, size | depth << CUT_SIZE_BITS (k_emit)
    const uint32_t r = e.x;
    const uint32_t size = min(min(e.y & CUT_SIZE_MASK, ACC_SUB), len - r);  // <= ACC_SUB by construction
    const uint32_t d0 = e.y >> CUT_SIZE_BITS;
    // Per cell: its children as 10-bit offsets in the subtree (0 = none: the
    // subtree's root, at offset 0, is no cell's child here), its depth, or
    // NONE for a leaf or a slot past the subtree. The loads are issued for
    // the four slots together, with no branch between them (a slot past the
    // subtree loads the root's again): records, then leaf heads, then
    // positions and strengths.
    uint4 ch4[SUB_PER];
    uint2 fc[SUB_PER];
    double x0[SUB_PER], x1[SUB_PER];
    uint32_t dep[SUB_PER], skp[SUB_PER];
#pragma unroll
    for (int q = 0; q < SUB_PER; ++q) {
      const uint32_t l = tid + (uint32_t)q * SUB_THREADS;
      const Cell* cp = cells + r + (l < size ? l : 0);
      ch4[q] = *reinterpret_cast<const uint4*>(cp->child);
      fc[q] = *reinterpret_cast<const uint2*>(&cp->first);
      x0[q] = cp->x0;
      x1[q] = cp->x1;
      dep[q] = cdepth[cp - cells];
      skp[q] = cskip[cp - cells];
    }
    uint64_t kids[SUB_PER];
    double w2t[SUB_PER];
    uint32_t head[SUB_PER];
    bool leaf[SUB_PER];
#pragma unroll
    for (int q = 0; q < SUB_PER; ++q) {
      const double w = x1[q] - x0[q];
      w2t[q] = w * w / theta2;  // many_body.zig:202, identical operands per cell
      const uint32_t k4[4] = {ch4[q].x, ch4[q].y, ch4[q].z, ch4[q].w};
      leaf[q] = k4[0] == NONE && k4[1] == NONE && k4[2] == NONE && k4[3] == NONE;
      kids[q] = 0;
      for (int k = 0; k < 4; ++k) {  // children lie inside the subtree's range
        const uint32_t o = k4[k] == NONE || k4[k] - r >= size ? 0u : k4[k] - r;
        kids[q] |= (uint64_t)o << (10 * k);
      }
      head[q] = perm[fc[q].x + fc[q].y - 1];
    }
    double hx[SUB_PER], hy[SUB_PER], hs[SUB_PER];
#pragma unroll
    for (int q = 0; q < SUB_PER; ++q) {
      hx[q] = xs[head[q]];
      hy[q] = ys[head[q]];
      hs[q] = s[head[q]];
    }
#pragma unroll
    for (int q = 0; q < SUB_PER; ++q) {
      const uint32_t l = tid + (uint32_t)q * SUB_THREADS;
      if (l >= size) {
        dep[q] = NONE;
        continue;
      }
      if (!leaf[q]) {
        atomicMax(&smaxd, dep[q]);
        continue;
      }
      // A leaf: its chain from the head down (many_body.zig:125-132); the
      // head's strength is loaded, a longer chain loads the rest here.
      double strength = 0;
      strength += hs[q];
      for (uint32_t p = fc[q].x + fc[q].y - 1; p > fc[q].x;) {
        --p;
        strength += s[perm[p]];
      }
      dep[q] = NONE;
      sx[l] = hx[q];
      sy[l] = hy[q];
      sv[l] = strength;
      sreach[l] = truthy(strength) ? 1 : 0;
      TCell t;
      t.qx = hx[q];
      t.qy = hy[q];
      t.value = strength;
      t.w2t = w2t[q];
      t.skip = 1;
      t.first = fc[q].x;
      t.count = fc[q].y;
      t.head = head[q];
      t.flags = TC_LEAF | (truthy(strength) ? TC_REACH : 0);
      t.pad[0] = t.pad[1] = t.pad[2] = 0;
      tc[r + l] = t;
    }
    __syncthreads();
    const uint32_t maxd = max(smaxd, d0);
    for (uint32_t d = maxd + 1; d-- > d0;) {
#pragma unroll
      for (int q = 0; q < SUB_PER; ++q) {
        if (dep[q] != d) continue;
        const uint32_t l = tid + (uint32_t)q * SUB_THREADS;
        // The four children's charges read together (slot 0 stands in for
        // an absent child), then summed in order.
        double cv[4], cx[4], cy[4];
        uint32_t ck[4];
        bool cr[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          ck[k] = (uint32_t)(kids[q] >> (10 * k)) & 1023u;
          cv[k] = sv[ck[k]];
          cx[k] = sx[ck[k]];
          cy[k] = sy[ck[k]];
          cr[k] = sreach[ck[k]] != 0;
        }
        double st = 0, weight = 0, x = 0, y = 0;
        bool reach = false;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (ck[k] == 0) continue;
          const double v = cv[k];
          const double w = fabs(v);
          if (w > 0) {  // `c = Math.abs(q.value)` is truthy
            st += v;
            weight += w;
            x += w * cx[k];
            y += w * cy[k];
          }
          reach = reach || cr[k];
        }
        const double qx = x / weight, qy = y / weight;
        const bool rr = reach && truthy(st);
        sx[l] = qx;
        sy[l] = qy;
        sv[l] = st;
        sreach[l] = rr ? 1 : 0;
        TCell t;
        t.qx = qx;
        t.qy = qy;
        t.value = st;
        t.w2t = w2t[q];
        t.skip = clamp_skip(skp[q], r + l, len);
        t.first = fc[q].x;
        t.count = fc[q].y;
        t.head = NONE;
        t.flags = rr ? TC_REACH : 0;
        t.pad[0] = t.pad[1] = t.pad[2] = 0;
        tc[r + l] = t;
      }
      __syncthreads();
    }
    if (tid == 0) smaxd = 0;  // every thread read it before the last barrier
    __syncthreads();
  }
}

// The cells above the subtrees (their subtrees exceed ACC_SUB cells), one
// workgroup, one depth at a time from the deepest; children come from the
// subtrees' records (the previous launch) or from a deeper pass of this one.
// Up to TOP_THREADS x TOP_PER cells keep their records in registers, so a
// depth costs one round of child loads.
constexpr int TOP_THREADS = 1024;
constexpr int TOP_PER = 4;

__device__ inline void acc_internal(TCell* __restrict__ tc, uint32_t c, const uint32_t* child, double w2t,
                                    uint32_t skip, uint32_t first, uint32_t count) {
  // The four children's charges are loaded together (an absent child loads
  // the cell's own record and is skipped), then summed in order.
  double cv[4], cx[4], cy[4];
  uint32_t cf[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const TCell& k = tc[child[q] == NONE ? c : child[q]];
    cv[q] = k.value;
    cx[q] = k.qx;
    cy[q] = k.qy;
    cf[q] = k.flags;
  }
  double st = 0, weight = 0, x = 0, y = 0;
  bool reach = false;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if (child[q] == NONE) continue;
    const double v = cv[q];
    const double w = fabs(v);
    if (w > 0) {  // `c = Math.abs(q.value)` is truthy (many_body.zig:112-121)
      st += v;
      weight += w;
      x += w * cx[q];
      y += w * cy[q];
    }
    reach = reach || (cf[q] & TC_REACH);
  }
  TCell t;
  t.qx = x / weight;
  t.qy = y / weight;
  t.value = st;
  t.w2t = w2t;
  t.skip = skip;
  t.first = first;
  t.count = count;
  t.head = NONE;
  t.flags = (reach && truthy(st)) ? TC_REACH : 0;
  t.pad[0] = t.pad[1] = t.pad[2] = 0;
  tc[c] = t;
}

__global__ __launch_bounds__(TOP_THREADS) void k_acc_top(const Cell* __restrict__ cells, TCell* __restrict__ tc,
                                                         const uint32_t* __restrict__ top,
                                                         const uint32_t* __restrict__ cskip,
                                                         const uint32_t* __restrict__ cdepth,
                                                         const TreeMeta* __restrict__ meta, double theta2) {
  const uint32_t ntop = meta->ntop, len = meta->len;
  if (ntop == 0) return;
  const uint32_t tid = threadIdx.x;
  const bool held = ntop <= (uint32_t)(TOP_THREADS * TOP_PER);
  uint32_t cc[TOP_PER], dep[TOP_PER], child[TOP_PER][4], skip[TOP_PER], first[TOP_PER], count[TOP_PER];
  double w2t[TOP_PER];
#pragma unroll
  for (int q = 0; q < TOP_PER; ++q) {
    const uint32_t k = tid + (uint32_t)q * TOP_THREADS;
    dep[q] = NONE;
    if (!held || k >= ntop) continue;
    const uint32_t c = top[k];
    const Cell cell = cells[c];
    cc[q] = c;
    dep[q] = cdepth[c];
    for (int j = 0; j < 4; ++j) child[q][j] = cell.child[j];
    const double w = cell.x1 - cell.x0;
    w2t[q] = w * w / theta2;
    skip[q] = clamp_skip(cskip[c], c, len);
    first[q] = cell.first;
    count[q] = cell.count;
  }
  for (uint32_t d = meta->topdepth + 1; d-- > 0;) {
    if (held) {
#pragma unroll
      for (int q = 0; q < TOP_PER; ++q)
        if (dep[q] == d) acc_internal(tc, cc[q], child[q], w2t[q], skip[q], first[q], count[q]);
    } else {
      for (uint32_t k = tid; k < ntop; k += TOP_THREADS) {
        const uint32_t c = top[k];
        if (cdepth[c] != d) continue;
        const Cell cell = cells[c];
        const double w = cell.x1 - cell.x0;
        acc_internal(tc, c, cell.child, w * w / theta2, clamp_skip(cskip[c], c, len), cell.first, cell.count);
      }
    }
    __syncthreads();
  }
}

__device__ inline double f64_of(uint32_t lo, uint32_t hi) {
  return __builtin_bit_cast(double, (uint64_t)hi << 32 | lo);
}

// d3's jiggle: (random() - 0.5) * 1e-6 (rng.zig:7-11).
__device__ inline double jiggle(double r) { return (r - 0.5) * 1e-6; }

constexpr uint32_t XCDS = 8;  // MI355X: workgroups go to the XCDs round-robin
constexpr int WALKS = 2;      // nodes per lane, walked side by side

// One node's walk: its position, its velocity being summed, the cell it is
// at, its draws (COUNT / REPLAY) and its work counts.
struct Walk {
  double nx, ny, vx, vy;
  uint32_t i, c, dptr, ndraw, n_visit, n_accept, n_pair;
  bool aborted;
};

// The jiggle and distanceMin rules of a visit (many_body.zig:203-207,
// 213-218): in RECORD mode a zero delta aborts the walk (false), else dx, dy
// take their draws and l its new value.
template <int MODE>
__device__ inline bool special(Walk& w, double& dx, double& dy, double& l, double dmin2,
                               const double* __restrict__ draws) {
  const bool jx = dx == 0, jy = dy == 0;
  if (MODE == TRAV_RECORD && (jx || jy)) {
    w.aborted = true;  // the first point where d3 would call random
    return false;
  }
  if (jx) {
    if (MODE == TRAV_COUNT) { ++w.ndraw; dx = 1; } else dx = jiggle(draws[w.dptr++]);
    l += dx * dx;
  }
  if (jy) {
    if (MODE == TRAV_COUNT) { ++w.ndraw; dy = 1; } else dy = jiggle(draws[w.dptr++]);
    l += dy * dy;
  }
  if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
  return true;
}

// One visit of walk w at its cell, whose first 48 bytes are g0 g1 g2; moves
// w.c on to the next cell of its sequence (len when it aborts).
template <int MODE>
__device__ inline void visit(Walk& w, const uint4& g0, const uint4& g1, const uint4& g2, uint32_t len,
                             const uint32_t* __restrict__ perm, const double* __restrict__ s, double dmin2,
                             double dmax2, double alpha, const double* __restrict__ draws) {
  const double value = f64_of(g1.x, g1.y);
  const uint32_t skip = g2.x;
  uint32_t next = w.c + (skip > 0 ? skip : 1);
  if (truthy(value)) {  // many_body.zig:191: `if (!quad.value) return true`
    ++w.n_visit;
    double dx = f64_of(g0.x, g0.y) - w.nx, dy = f64_of(g0.z, g0.w) - w.ny;
    double l = dx * dx + dy * dy;
    if (f64_of(g1.z, g1.w) < l) {  // w * w / theta2 < l: the Barnes-Hut approximation
      if (l < dmax2) {
        bool add = true;
        if (dx == 0 || dy == 0 || l < dmin2) add = special<MODE>(w, dx, dy, l, dmin2, draws);
        if (add) {
          w.vx += dx * value * alpha / l;
          w.vy += dy * value * alpha / l;
          ++w.n_accept;
        } else {
          next = len;
        }
      }
    } else if (skip > 1) {
      next = w.c + 1;  // opened: its first child
    } else if (!(l >= dmax2)) {
      // Leaf: points directly (many_body.zig:208-258).
      const uint32_t first = g2.y, cnt = g2.z, head = g2.w;
      bool go = true;
      if (head != w.i || cnt > 1) go = special<MODE>(w, dx, dy, l, dmin2, draws);
      if (go) {
        for (uint32_t k = first + cnt; k > first;) {
          --k;
          const uint32_t j = perm[k];
          if (j != w.i) {
            const double wj = s[j] * alpha / l;
            w.vx += dx * wj;
            w.vy += dy * wj;
            ++w.n_pair;
          }
        }
      } else {
        next = len;
      }
    }
  }
  w.c = next;
}

template <int MODE>
__global__ __launch_bounds__(TRAV_THREADS) void k_traverse(
    const TCell* __restrict__ tc, uint32_t len, const uint32_t* __restrict__ order, uint32_t count,
    const uint32_t* __restrict__ perm, const double* __restrict__ xs, const double* __restrict__ ys,
    const double* __restrict__ s, double* __restrict__ vxs, double* __restrict__ vys, double dmin2, double dmax2,
    double alpha, uint32_t* __restrict__ flagged, uint32_t* __restrict__ n_flagged,
    uint32_t* __restrict__ draws_count, const uint32_t* __restrict__ draw_off, const double* __restrict__ draws,
    uint32_t per_xcd, uint32_t blocks, unsigned long long* __restrict__ counters, const TreeMeta* __restrict__ meta,
    const double* __restrict__ alpha_dev) {
  // Block b runs on XCD b % 8; give each XCD a contiguous range of blocks.
  const uint32_t block = (blockIdx.x % XCDS) * per_xcd + blockIdx.x / XCDS;
  if (block >= blocks) return;
  if (meta) {  // a captured tick: the tree's sizes from the device
    if (meta->stop) return;
    count = meta->m;
    len = meta->len;
  }
  if (alpha_dev) alpha = *alpha_dev;
  const int lane = threadIdx.x & 63;
  // A wave takes WALKS * 64 consecutive slots (Morton-ordered nodes); lane L
  // walks slots base + L, base + 64 + L, ...
  const uint32_t base = (block * TRAV_THREADS + (threadIdx.x & ~63u)) * WALKS + lane;
  Walk w[WALKS];
#pragma unroll
  for (int k = 0; k < WALKS; ++k) {
    const uint32_t slot = base + 64 * k;
    const bool active = slot < count;
    Walk& a = w[k];
    a.i = active ? order[slot] : 0;
    a.nx = a.ny = a.vx = a.vy = 0;
    if (active) {
      a.nx = xs[a.i];
      a.ny = ys[a.i];
      a.vx = vxs[a.i];
      a.vy = vys[a.i];
    }
    a.c = active ? 0 : len;  // cell 0 is the root
    a.dptr = (MODE == TRAV_REPLAY && active) ? draw_off[slot] : 0;
    a.ndraw = a.n_visit = a.n_accept = a.n_pair = 0;
    a.aborted = false;
  }
  for (;;) {
    bool any = false;
#pragma unroll
    for (int k = 0; k < WALKS; ++k) any = any || w[k].c < len;
    if (!any) break;
    // Every walk's record first, so their loads are in flight together.
    uint4 g[WALKS][3];
#pragma unroll
    for (int k = 0; k < WALKS; ++k) {
      const uint4* r = reinterpret_cast<const uint4*>(tc + (w[k].c < len ? w[k].c : 0));
      g[k][0] = r[0];
      g[k][1] = r[1];
      g[k][2] = r[2];
    }
#pragma unroll
    for (int k = 0; k < WALKS; ++k)
      if (w[k].c < len) visit<MODE>(w[k], g[k][0], g[k][1], g[k][2], len, perm, s, dmin2, dmax2, alpha, draws);
  }
  if (MODE != TRAV_COUNT) {
    // Work of the walks whose results stand (an aborted walk is redone by REPLAY).
    uint32_t cv = 0, ca = 0, cp = 0;
#pragma unroll
    for (int k = 0; k < WALKS; ++k) {
      const bool keep = base + 64 * k < count && !(MODE == TRAV_RECORD && w[k].aborted);
      cv += keep ? w[k].n_visit : 0;
      ca += keep ? w[k].n_accept : 0;
      cp += keep ? w[k].n_pair : 0;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      cv += __shfl_xor(cv, d, 64);
      ca += __shfl_xor(ca, d, 64);
      cp += __shfl_xor(cp, d, 64);
    }
    if (lane == 0 && counters) {
      atomicAdd(&counters[0], (unsigned long long)cv);
      atomicAdd(&counters[1], (unsigned long long)ca);
      atomicAdd(&counters[2], (unsigned long long)cp);
    }
  }
#pragma unroll
  for (int k = 0; k < WALKS; ++k) {
    const uint32_t slot = base + 64 * k;
    if (slot >= count) continue;
    if (MODE == TRAV_RECORD && w[k].aborted) {
      const uint32_t q = atomicAdd(n_flagged, 1u);
      flagged[q] = w[k].i;
    } else if (MODE == TRAV_COUNT) {
      draws_count[slot] = w[k].ndraw;
    } else {
      vxs[w[k].i] = w[k].vx;
      vys[w[k].i] = w[k].vy;
    }
  }
}

__global__ void k_nan_nodes(const uint32_t* __restrict__ rest, uint32_t count, const TCell* __restrict__ tc,
                            double* __restrict__ vx, double* __restrict__ vy) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  if (!(tc[0].flags & TC_REACH)) return;
  const uint32_t i = rest[k];
  const double nan = __builtin_nan("");
  vx[i] = vx[i] + nan;
  vy[i] = vy[i] + nan;
}

// k_nan_nodes after a capturable build: the count n - m from the tree's meta.
__global__ void k_nan_nodes_dev(const uint32_t* __restrict__ rest, uint32_t n, const TreeMeta* __restrict__ meta,
                                const TCell* __restrict__ tc, double* __restrict__ vx, double* __restrict__ vy) {
  if (meta->stop || !meta->valid) return;
  const uint32_t m = meta->m;
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n || k >= n - m) return;
  if (!(tc[0].flags & TC_REACH)) return;
  const uint32_t i = rest[k];
  const double nan = __builtin_nan("");
  vx[i] = vx[i] + nan;
  vy[i] = vy[i] + nan;
}

}  // namespace

hipError_t tree_depth_order(TreeBuffers& t, hipStream_t stream) {
  const uint32_t len = t.meta_host->len;
  if (len == 0) return hipSuccess;
  const uint32_t nbins = t.meta_host->depth + 1;
  if (nbins + 1 > (uint32_t)DEPTH_BINS) HIPFAIL(hipErrorInvalidValue, "tree deeper than the depth bins");
  const size_t lds = (size_t)nbins * sizeof(uint32_t);
  HIPCHECK(hipMemsetAsync(t.dbins, 0, lds, stream));
  uint32_t hist_blocks = ceil_div(len, ACC_THREADS * 8);
  if (hist_blocks > 1024) hist_blocks = 1024;
  hipLaunchKernelGGL(k_depth_hist, dim3(hist_blocks), dim3(ACC_THREADS), lds, stream, t.cdepth, len, nbins, t.dbins);
  hipLaunchKernelGGL(k_depth_scan, dim3(1), dim3(64), 0, stream, t.dbins, nbins, t.dstart);
  hipLaunchKernelGGL(k_depth_scatter, dim3(ceil_div(len, ACC_THREADS * SCATTER_PER)), dim3(ACC_THREADS), lds, stream,
                     t.cdepth, len, nbins, t.dbins, t.dorder);
  return hipGetLastError();
}

hipError_t mb_accumulate(TreeBuffers& t, const double* x, const double* y, const double* strengths, double theta2,
                         hipStream_t stream, bool graph) {
  // A capturable build left no mixed group (else it stopped, and the lists
  // these kernels read are empty).
  const uint32_t len = graph ? t.cell_cap : t.meta_host->len;
  if (len == 0) return hipSuccess;
  if (graph || t.meta_host->mixed == 0) {
    // Subtree sizes and lists from the build (cells below key resolution,
    // written by the serial split, have none: those trees go level by level).
    // Enough blocks for one subtree each (subtrees average a few hundred
    // cells); blocks past the count return at once.
    uint32_t blocks = ceil_div(len, 128);
    if (blocks > 65535) blocks = 65535;
    hipLaunchKernelGGL(k_acc_subtrees, dim3(blocks), dim3(SUB_THREADS), 0, stream, t.cells, t.tcells, t.cut, t.cskip,
                       t.cdepth, t.meta, t.perm, x, y, strengths, theta2);
    hipLaunchKernelGGL(k_acc_top, dim3(1), dim3(TOP_THREADS), 0, stream, t.cells, t.tcells, t.top, t.cskip, t.cdepth,
                       t.meta, theta2);
    return hipGetLastError();
  }
  const uint32_t nbins = t.meta_host->depth + 1;
  HIPCHECK(tree_depth_order(t, stream));
  uint32_t level_blocks = ceil_div(len, ACC_THREADS);
  if (level_blocks > 2048) level_blocks = 2048;
  for (uint32_t d = nbins; d-- > 0;) {
    hipLaunchKernelGGL(k_accumulate_level, dim3(level_blocks), dim3(ACC_THREADS), 0, stream, t.cells, t.tcells,
                       t.dorder, t.dstart, d, t.perm, x, y, strengths, theta2);
  }
  return hipGetLastError();
}

hipError_t mb_publish_charges(TreeBuffers& t, hipStream_t stream) {
  const uint32_t len = t.meta_host->len;
  if (len == 0) return hipSuccess;
  hipLaunchKernelGGL(k_publish_charges, dim3(ceil_div(len, ACC_THREADS)), dim3(ACC_THREADS), 0, stream, t.cells,
                     t.tcells, len);
  return hipGetLastError();
}

hipError_t mb_traverse(const TravArgs& a, int mode, hipStream_t stream) {
  if (a.count == 0) return hipSuccess;
  const uint32_t blocks = ceil_div(a.count, TRAV_THREADS * WALKS);
  const uint32_t per_xcd = ceil_div(blocks, XCDS);
#define TRAV_LAUNCH(M)                                                                                            \
  hipLaunchKernelGGL((k_traverse<M>), dim3(per_xcd * XCDS), dim3(TRAV_THREADS), 0, stream, a.tc, a.len, a.order,  \
                     a.count, a.perm, a.x, a.y, a.s, a.vx, a.vy, a.dmin2, a.dmax2, a.alpha, a.flagged,           \
                     a.n_flagged, a.draws_count, a.draw_off, a.draws, per_xcd, blocks, a.counters, a.meta,        \
                     a.alpha_dev)
  if (mode == TRAV_RECORD) TRAV_LAUNCH(TRAV_RECORD);
  else if (mode == TRAV_COUNT) TRAV_LAUNCH(TRAV_COUNT);
  else TRAV_LAUNCH(TRAV_REPLAY);
#undef TRAV_LAUNCH
  return hipGetLastError();
}

hipError_t mb_nan_nodes_graph(const TreeBuffers& t, uint32_t n, double* vx, double* vy, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_nan_nodes_dev, dim3(ceil_div(n, 256)), dim3(256), 0, stream, t.rest, n, t.meta, t.tcells, vx,
                     vy);
  return hipGetLastError();
}

hipError_t mb_nan_nodes(const TreeBuffers& t, uint32_t n, double* vx, double* vy, hipStream_t stream) {
  const uint32_t m = t.meta_host->m;
  if (!t.meta_host->valid || n <= m) return hipSuccess;
  const uint32_t count = n - m;
  hipLaunchKernelGGL(k_nan_nodes, dim3(ceil_div(count, 256)), dim3(256), 0, stream, t.rest, count, t.tcells, vx, vy);
  return hipGetLastError();
}

}  // namespace d3f
