// forceManyBody, exact mode, on gfx950 (reference: src/core/many_body.zig,
// itself d3-force's manyBody.js over the quadtree).
//
// Accumulate (visitAfter): cells are bucketed by depth and processed from
// the deepest level up, one launch per level; a leaf sums its chain from the
// head down (many_body.zig:125-132) and an internal cell combines children
// 0..3 in order (many_body.zig:112-121), so every sum has d3's operands in
// d3's order. Kernel boundaries order the levels, with no inter-workgroup
// atomics (whose cache maintenance spans all eight XCDs' L2s).
//
// Traversal (visit): one lane per node, 64 nodes per wave in Morton order.
// The wave walks the union of its lanes' d3 traversals depth first with one
// shared LDS stack of (cell, lane mask) entries: a lane processes a cell only
// if every ancestor on the path was opened by that lane, so each lane sees
// exactly its own serial visit sequence (children 0..3, many_body.zig:190-259)
// and accumulates its velocity in d3's order. Cells are read with uniform
// (scalar) loads of one 64-byte record.
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
constexpr int TRAV_WAVES = TRAV_THREADS / 64;

__device__ inline bool truthy(double v) { return v != 0 && v == v; }

// A cell's traversal record. The 88-byte cell's x, y, value receive the same
// charges only when the tree is exported (k_publish_charges).
__device__ inline void put(TCell* __restrict__ tc, uint32_t c, const Cell& cell, double qx, double qy, double value,
                           double theta2, uint32_t head, uint32_t flags) {
  TCell t;
  t.qx = qx;
  t.qy = qy;
  t.value = value;
  const double w = cell.x1 - cell.x0;
  t.w2t = w * w / theta2;  // many_body.zig:202, identical operands per cell
  for (int k = 0; k < 4; ++k) t.child[k] = cell.child[k];
  t.first = cell.first;
  t.count = cell.count;
  t.head = head;
  t.flags = flags;
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
      put(tc, c, cell, xs[head], ys[head], strength, theta2, head,
          TC_LEAF | (truthy(strength) ? TC_REACH : 0));
      continue;
    }
    double st = 0, weight = 0, x = 0, y = 0;
    bool reach = false;
    for (int q = 0; q < 4; ++q) {
      const uint32_t ch = cell.child[q];
      if (ch == NONE) continue;
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
    put(tc, c, cell, x / weight, y / weight, st, theta2, NONE, (reach && truthy(st)) ? TC_REACH : 0);
  }
}

// A TCell as 16 dwords: qx 0-1, qy 2-3, value 4-5, w2t 6-7, child 8-11,
// first 12, count 13, head 14, flags 15.
typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
static_assert(sizeof(u32x16) == sizeof(TCell), "traversal record");

__device__ inline double rec_f64(const u32x16& r, int w) {
  return __builtin_bit_cast(double, (uint64_t)r[w + 1] << 32 | r[w]);
}

// d3's jiggle: (random() - 0.5) * 1e-6 (rng.zig:7-11).
__device__ inline double jiggle(double r) { return (r - 0.5) * 1e-6; }

template <int MODE, bool GLOBAL_STACK>
__global__ __launch_bounds__(TRAV_THREADS) void k_traverse(
    const TCell* __restrict__ tc, const uint32_t* __restrict__ order, uint32_t count,
    const uint32_t* __restrict__ perm, const double* __restrict__ xs, const double* __restrict__ ys,
    const double* __restrict__ s, double* __restrict__ vxs, double* __restrict__ vys, double dmin2, double dmax2,
    double alpha, uint32_t* __restrict__ flagged, uint32_t* __restrict__ n_flagged,
    uint32_t* __restrict__ draws_count, const uint32_t* __restrict__ draw_off, const double* __restrict__ draws,
    uint32_t stack_cap, uint32_t* __restrict__ gstack_cells, uint64_t* __restrict__ gstack_masks,
    unsigned long long* __restrict__ counters, uint32_t* __restrict__ err) {
  extern __shared__ uint64_t lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint64_t* st_mask;
  uint32_t* st_cell;
  if (GLOBAL_STACK) {
    const size_t w = (size_t)blockIdx.x * TRAV_WAVES + wave;
    st_mask = gstack_masks + w * stack_cap;
    st_cell = gstack_cells + w * stack_cap;
  } else {
    st_mask = lds + (size_t)wave * stack_cap;
    st_cell = (uint32_t*)(lds + (size_t)TRAV_WAVES * stack_cap) + (size_t)wave * stack_cap;
  }
  const uint32_t slot = blockIdx.x * TRAV_THREADS + threadIdx.x;
  const bool active = slot < count;
  const uint32_t i = active ? order[slot] : 0;
  double nx = 0, ny = 0, vx = 0, vy = 0;
  if (active) {
    nx = xs[i];
    ny = ys[i];
    vx = vxs[i];
    vy = vys[i];
  }
  uint64_t alive = __ballot(active);
  if (alive == 0) return;
  uint32_t ndraw = 0;
  uint32_t n_visit = 0, n_accept = 0, n_pair = 0;
  uint32_t dptr = (MODE == TRAV_REPLAY && active) ? draw_off[slot] : 0;
  bool aborted = false;
  if (lane == 0) {
    st_cell[0] = 0;
    st_mask[0] = alive;
  }
  int sp = 1;
  while (sp > 0) {
    --sp;
    __builtin_amdgcn_wave_barrier();
    const uint32_t c = __builtin_amdgcn_readfirstlane(st_cell[sp]);
    const uint64_t mraw = st_mask[sp];
    const uint32_t mlo = __builtin_amdgcn_readfirstlane((uint32_t)mraw);
    const uint32_t mhi = __builtin_amdgcn_readfirstlane((uint32_t)(mraw >> 32));
    const uint64_t M = ((uint64_t)mhi << 32 | mlo) & alive;
    if (M == 0) continue;
    // The whole 64-byte record in one scalar load (fields as TCell lays them
    // out), so a visit waits on one load rather than one per field group.
    const u32x16 rec = *reinterpret_cast<const u32x16*>(tc + c);
    const double value = rec_f64(rec, 4);
    if (!truthy(value)) continue;
    const bool in = (M >> lane) & 1;
    // Geometry on every lane (used where `in`); a lane accepts the cell when
    // the Barnes-Hut test passes and the cell lies within distanceMax.
    double dx = rec_f64(rec, 0) - nx, dy = rec_f64(rec, 2) - ny;
    double l = dx * dx + dy * dy;
    const bool far = rec_f64(rec, 6) < l;  // Barnes-Hut approximation
    const bool open = in && !far;
    const bool acc = in && far && l < dmax2;
    n_visit += in ? 1u : 0u;
    if (__ballot(acc) != 0) {
      if (__ballot(acc && (dx == 0 || dy == 0 || l < dmin2)) == 0) {
        // the common case, no lane needing a jiggle or distanceMin: d3's
        // terms added on the accepting lanes by select
        const double tx = dx * value * alpha / l, ty = dy * value * alpha / l;
        vx = acc ? vx + tx : vx;
        vy = acc ? vy + ty : vy;
        n_accept += acc ? 1u : 0u;
      } else if (acc) {
        const bool jx = dx == 0, jy = dy == 0;
        if (MODE == TRAV_RECORD && (jx || jy)) {
          aborted = true;
        } else {
          if (jx) {
            if (MODE == TRAV_COUNT) { ++ndraw; dx = 1; } else dx = jiggle(draws[dptr++]);
            l += dx * dx;
          }
          if (jy) {
            if (MODE == TRAV_COUNT) { ++ndraw; dy = 1; } else dy = jiggle(draws[dptr++]);
            l += dy * dy;
          }
          if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
          vx += dx * value * alpha / l;
          vy += dy * value * alpha / l;
          ++n_accept;
        }
      }
    }
    if (MODE == TRAV_RECORD) alive &= ~__ballot(aborted);
    const uint64_t O = __ballot(open);
    if (O == 0) continue;
    if (!(rec[15] & TC_LEAF)) {
      if (sp + 4 > (int)stack_cap) {  // cannot happen with cap = 3 depth + 4; if it does, say so
        if (lane == 0) atomicOr(err, 1u);
        break;
      }
      // Descend: push children 3..0 so 0 is visited first, in one store:
      // lane L < 4 writes child 3 - L, if present, above the present
      // children of higher index.
      const uint32_t h0 = rec[8], h1 = rec[9], h2 = rec[10], h3 = rec[11];
      const int e0 = h0 != NONE, e1 = h1 != NONE, e2 = h2 != NONE, e3 = h3 != NONE;
      if (lane < 4) {
        const uint32_t ch = lane == 0 ? h3 : lane == 1 ? h2 : lane == 2 ? h1 : h0;
        const int before = lane == 0 ? 0 : lane == 1 ? e3 : lane == 2 ? e3 + e2 : e3 + e2 + e1;
        if (ch != NONE) {
          st_cell[sp + before] = ch;
          st_mask[sp + before] = O;
        }
      }
      sp += e0 + e1 + e2 + e3;
      continue;
    }
    // Leaf: points directly (many_body.zig:208-258).
    if (open && !(l >= dmax2)) {
      const uint32_t first = rec[12], cnt = rec[13], head = rec[14];
      bool go = true;
      if (head != i || cnt > 1) {
        const bool jx = dx == 0, jy = dy == 0;
        if (MODE == TRAV_RECORD && (jx || jy)) {
          aborted = true;
          go = false;
        } else {
          if (jx) {
            if (MODE == TRAV_COUNT) { ++ndraw; dx = 1; } else dx = jiggle(draws[dptr++]);
            l += dx * dx;
          }
          if (jy) {
            if (MODE == TRAV_COUNT) { ++ndraw; dy = 1; } else dy = jiggle(draws[dptr++]);
            l += dy * dy;
          }
          if (l < dmin2) l = __builtin_sqrt(dmin2 * l);
        }
      }
      if (go) {
        for (uint32_t k = first + cnt; k > first;) {
          --k;
          const uint32_t j = perm[k];
          if (j != i) {
            const double wj = s[j] * alpha / l;
            vx += dx * wj;
            vy += dy * wj;
            ++n_pair;
          }
        }
      }
    }
    if (MODE == TRAV_RECORD) alive &= ~__ballot(aborted);
  }
  if (MODE != TRAV_COUNT) {
    // Work of the lanes whose results stand (an aborted lane is redone by REPLAY).
    const bool keep = active && !(MODE == TRAV_RECORD && aborted);
    uint32_t cv = keep ? n_visit : 0, ca = keep ? n_accept : 0, cp = keep ? n_pair : 0;
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
  if (!active) return;
  if (MODE == TRAV_RECORD && aborted) {
    const uint32_t k = atomicAdd(n_flagged, 1u);
    flagged[k] = i;
  } else if (MODE == TRAV_COUNT) {
    draws_count[slot] = ndraw;
  } else {
    vxs[i] = vx;
    vys[i] = vy;
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
                         hipStream_t stream) {
  const uint32_t len = t.meta_host->len;
  if (len == 0) return hipSuccess;
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
  const uint32_t blocks = ceil_div(a.count, TRAV_THREADS);
  const size_t lds = (size_t)TRAV_WAVES * a.stack_cap * (sizeof(uint64_t) + sizeof(uint32_t));
  const bool global = lds > 48 * 1024;
  TreeBuffers& st = *a.stacks;
  if (global) {
    const size_t need = (size_t)blocks * TRAV_WAVES * a.stack_cap;
    if (need > st.tstack_entries) {
      if (st.tstack_cells) HIPCHECK(hipFree(st.tstack_cells));
      if (st.tstack_masks) HIPCHECK(hipFree(st.tstack_masks));
      st.tstack_cells = nullptr;
      st.tstack_masks = nullptr;
      st.tstack_entries = 0;
      HIPCHECK(hipMalloc(&st.tstack_cells, need * sizeof(uint32_t)));
      HIPCHECK(hipMalloc(&st.tstack_masks, need * sizeof(uint64_t)));
      st.tstack_entries = need;
    }
  }
#define TRAV_LAUNCH(M, G)                                                                                        \
  hipLaunchKernelGGL((k_traverse<M, G>), dim3(blocks), dim3(TRAV_THREADS), G ? 0 : lds, stream, a.tc, a.order,   \
                     a.count, a.perm, a.x, a.y, a.s, a.vx, a.vy, a.dmin2, a.dmax2, a.alpha, a.flagged,          \
                     a.n_flagged, a.draws_count, a.draw_off, a.draws, a.stack_cap, st.tstack_cells,              \
                     st.tstack_masks, a.counters, a.error)
  if (global) {
    if (mode == TRAV_RECORD) TRAV_LAUNCH(TRAV_RECORD, true);
    else if (mode == TRAV_COUNT) TRAV_LAUNCH(TRAV_COUNT, true);
    else TRAV_LAUNCH(TRAV_REPLAY, true);
  } else {
    if (mode == TRAV_RECORD) TRAV_LAUNCH(TRAV_RECORD, false);
    else if (mode == TRAV_COUNT) TRAV_LAUNCH(TRAV_COUNT, false);
    else TRAV_LAUNCH(TRAV_REPLAY, false);
  }
#undef TRAV_LAUNCH
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
