// Exact forceCollide (collide.zig:17-130) on the GPU, bit-identical to the
// reference's serial Gauss-Seidel pass.
//
// One iteration of the reference: a quadtree on (x + vx, y + vy), each cell's
// radius bottom-up, then for i = 0 .. n-1 in order: xi = x_i + vx_i, a
// depth-first walk that prunes cells by their bounds against xi +- (ri + r),
// and for every visited leaf whose chain head j > i overlaps (l < r^2) a push
// that adds to vx_i and subtracts from vx_j. Node i's walk reads positions
// that earlier nodes' pushes changed, so the pass is a chain of dependent
// transactions; the GPU runs it as a dataflow over the conflicts.
//
// 1. k_collide_radius_level: the cell radii, one launch per depth, deepest
//    first (collide.zig:118-129).
// 2. k_collide_cand: one thread per node walks the tree from its tree
//    position p0 = (x + vx, y + vy) with every position allowed to move by
//    up to delta per axis during the pass (delta = scale |strength| R +
//    1e-9 S, R the largest radius, S the extent's magnitude; the host adapts
//    the scale). It descends every cell that
//    some position in that box would not prune, keeps the leaves j > i that
//    some positions in the boxes of i and j would make overlap (the
//    candidates, in the walk's preorder), and for each candidate its leaf
//    cell and its shallowest "uncertain" ancestor, the first cell on its
//    path whose prune test the box does not decide.
// 3. The (candidate j, node i) pairs are sorted by (j, i): a pair's rank is
//    the number of earlier nodes with j among their candidates, and indeg(j)
//    is the number of all such nodes.
// 4. k_collide_apply: a persistent dataflow kernel. Waves take nodes in
//    increasing index order; node i runs once done[i] == indeg(i) (every
//    earlier push onto i has landed) and done[j] == rank for each candidate
//    (every earlier toucher of j is done, no later one has started). It then
//    does exactly the reference's arithmetic for node i: xi from the current
//    vx_i, each candidate's visit decided by the prune tests at xi on its
//    path from the uncertain ancestor down, the overlap test and push at the current
//    vx_j, and finally increments done[] of every node it touched. Leaves that
//    are not candidates cannot overlap for any positions in the boxes, so
//    their reads need no order.
//    The boxes are checked as the pass runs: every pushed position must stay
//    within delta of p0 (per axis). A push outside the box, an exact zero
//    delta (the reference then draws a jiggle from the host's random, in
//    pass order) parks that node before it writes anything: the nodes that
//    do not depend on it finish, and k_collide_serial continues with the
//    unfinished ones in index order. A push outside the box, a full
//    candidate list or a wait that makes no progress for a second stops the
//    pass; the host then restores vx, vy and retries with wider boxes or runs
//    the whole iteration with k_collide_serial.
// 5. k_collide_serial: the reference's pass on one thread, stopping before a
//    pair that needs jiggles so that the host draws them in order.
//
// Memory ordering: vx and vy are read and written with agent-scope relaxed
// atomics (the waves of the dataflow run on all XCDs, whose L2s are not
// coherent for ordinary accesses); done[] increments follow an agent-scope
// release fence, and a node's reads follow an agent-scope acquire fence after
// its last wait.

#include "common.h"

namespace d3f {

namespace {

constexpr uint32_t COLLIDE_THREADS = 256;
constexpr uint64_t WAIT_LIMIT = 100000000ull;  // wall clock ticks (100 MHz): 1 s without progress
constexpr uint64_t PARKED_LIMIT = 200000ull;   // 2 ms, once a node has stopped for a jiggle

__device__ inline bool is_leaf(const Cell& q) {
  return q.child[0] == NONE && q.child[1] == NONE && q.child[2] == NONE && q.child[3] == NONE;
}

// collide.zig:89: the prune test of an internal cell for a node at (xi, yi)
// with radius ri (r = ri + quad.r computed as the reference does).
__device__ inline bool pruned(const CRec& q, double ri, double xi, double yi) {
  const double r = ri + q.r;
  return q.x0 > xi + r || q.x1 < xi - r || q.y0 > yi + r || q.y1 < yi - r;
}

// The box half-width: a push moves a node by at most |strength| times the
// pair's radius sum, so delta is a multiple (the host's scale) of
// |strength| R, R the largest radius, plus 1e-9 of the extent's magnitude,
// far above the rounding of the positions.
__device__ inline double collide_delta(const CollideArgs& a) {
  const double R = a.cells[0].r > 0 ? a.cells[0].r : 0.0;  // the root holds the largest radius (or 0)
  double S = 0;
  for (int k = 0; k < 4; ++k) S = fmax(S, fabs(a.meta->extent[k]));
  return a.scale * fabs(a.strength) * R + 1e-9 * S + 1e-300;
}

__device__ inline double ld(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void st(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint32_t ldu(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

// collide.zig:104-130 (Prep): a leaf takes its chain head's radius; an
// internal cell the largest child radius above 0 (NaN children never win).
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_radius_level(Cell* __restrict__ cells,
                                                                          const uint32_t* __restrict__ order,
                                                                          const uint32_t* __restrict__ start,
                                                                          uint32_t depth,
                                                                          const uint32_t* __restrict__ perm,
                                                                          const double* __restrict__ radii) {
  const uint32_t lo = start[depth], hi = start[depth + 1];
  for (uint32_t k = lo + blockIdx.x * COLLIDE_THREADS + threadIdx.x; k < hi; k += gridDim.x * COLLIDE_THREADS) {
    const uint32_t c = order[k];
    Cell& q = cells[c];
    double r;
    if (is_leaf(q)) {
      r = radii[perm[q.first + q.count - 1]];
    } else {
      r = 0;
      for (int ch = 0; ch < 4; ++ch) {
        const uint32_t cc = q.child[ch];
        if (cc != NONE && cells[cc].r > r) r = cells[cc].r;
      }
    }
    q.r = r;
  }
}

// Threads take the tree's points in tree (Morton) order, so the walks of a
// wave cover nearby cells; nodes outside the tree (NaN positions) keep the
// zero count the host wrote.
// The walks' 64-byte view of a cell: bounds, radius, children and the
// chain head of a leaf (NONE for an internal cell).
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_records(const Cell* __restrict__ cells, uint32_t len,
                                                                     const uint32_t* __restrict__ perm,
                                                                     CRec* __restrict__ rec) {
  const uint32_t c = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (c >= len) return;
  const Cell q = cells[c];
  CRec o;
  o.x0 = q.x0;
  o.y0 = q.y0;
  o.x1 = q.x1;
  o.y1 = q.y1;
  o.r = q.r;
  for (int k = 0; k < 4; ++k) o.child[k] = q.child[k];
  o.head = is_leaf(q) ? perm[q.first + q.count - 1] : NONE;
  o.pad = 0;
  rec[c] = o;
}

__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_cand(const CollideArgs a) {
  const uint32_t t = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (t >= a.m) return;
  const uint32_t i = a.perm[t];
  const double xi = a.px[i], yi = a.py[i];
  if (!__builtin_isfinite(xi) || !__builtin_isfinite(yi)) {
    // A non-finite position makes every overlap test false (l is inf or NaN)
    // and is never pushed: the node reads and writes nothing.
    a.ccount[i] = 0;
    return;
  }
  const double delta = collide_delta(a);
  if (!__builtin_isfinite(delta)) {
    atomicOr(a.status, COLLIDE_FULL);
    a.ccount[i] = 0;
    return;
  }
  const double D = 1.001 * delta;
  const double xlo = xi - D, xhi = xi + D, ylo = yi - D, yhi = yi + D;
  // Both ends of a pair move by at most delta per axis: the delta vector by
  // at most 2 sqrt(2) delta < 3 delta.
  const double pad = 3 * delta;
  const double ri = a.radii[i];
  uint32_t* cand = a.cand + (size_t)i * COLLIDE_K;
  uint32_t* cleaf = a.cleaf + (size_t)i * COLLIDE_K;
  uint32_t* ctop = a.ctop + (size_t)i * COLLIDE_K;
  // Stack entries carry the shallowest undecided ancestor of their cell
  // (NONE when the box decides every test above it): a candidate's visit is
  // then settled at run time by the tests from that ancestor down to its
  // leaf.
  uint32_t stk[COLLIDE_STACK], stop[COLLIDE_STACK];
  uint32_t sp = 1, cnt = 0, fail = 0;
  stk[0] = 0;
  stop[0] = NONE;
  while (sp > 0) {
    --sp;
    const uint32_t c = stk[sp], top = stop[sp];
    const CRec q = a.rec[c];
    const double r = ri + q.r;
    if (q.head != NONE) {
      const uint32_t j = q.head;
      if (j > i) {
        const double dx = xi - a.px[j], dy = yi - a.py[j];
        const double l = dx * dx + dy * dy;
        const double lim = fabs(r) + pad;
        if (l < lim * lim) {
          if (cnt == COLLIDE_K) {
            fail = COLLIDE_FULL;
            break;
          }
          cand[cnt] = j;
          cleaf[cnt] = c;
          ctop[cnt] = top;
          ++cnt;
        }
      }
      continue;
    }
    // Each half of the test is monotone in xi (or yi), so the box decides
    // it when its two ends agree.
    const bool always = q.x0 > xhi + r || q.x1 < xlo - r || q.y0 > yhi + r || q.y1 < ylo - r;
    if (always) continue;
    const bool never = !(q.x0 > xlo + r) && !(q.x1 < xhi - r) && !(q.y0 > ylo + r) && !(q.y1 < yhi - r);
    const uint32_t next = (never || top != NONE) ? top : c;
    if (sp + 4 > COLLIDE_STACK) {
      fail = COLLIDE_FULL_STACK;
      break;
    }
    for (int k = 3; k >= 0; --k) {
      if (q.child[k] != NONE) {
        stk[sp] = q.child[k];
        stop[sp] = next;
        ++sp;
      }
    }
  }
  if (fail) {
    atomicOr(a.status, fail);
    a.ccount[i] = 0;
    return;
  }
  a.ccount[i] = cnt;
}

// Pairs (j << key_shift | i) at node i's offset in the scan of the counts;
// the value is the pair's own position, where its rank goes.
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_pairs(const CollideArgs a) {
  const uint32_t i = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (i >= a.n) return;
  const uint32_t cnt = a.ccount[i], base = a.coff[i];
  for (uint32_t k = 0; k < cnt; ++k) {
    a.pkeys[base + k] = (uint64_t)a.cand[(size_t)i * COLLIDE_K + k] << a.key_shift | i;
    a.pvals[base + k] = base + k;
  }
}

// Sorted pairs (j, i): the first pair of each j records where j's run starts.
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_runs(const CollideArgs a, uint32_t pairs) {
  const uint32_t p = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (p >= pairs) return;
  const uint64_t j = a.pkeys[p] >> a.key_shift;
  if (p == 0 || (a.pkeys[p - 1] >> a.key_shift) != j) a.run_start[j] = p;
}

__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_ranks(const CollideArgs a, uint32_t pairs) {
  const uint32_t p = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (p >= pairs) return;
  const uint32_t j = (uint32_t)(a.pkeys[p] >> a.key_shift);
  const uint32_t r = p - a.run_start[j];
  a.rank[a.pvals[p]] = r;
  if (p + 1 == pairs || (uint32_t)(a.pkeys[p + 1] >> a.key_shift) != j) a.indeg[j] = r + 1;
}

__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_apply(const CollideArgs a) {
  const uint32_t lane = threadIdx.x & 63;
  const double delta = collide_delta(a);
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(a.ticket, 64u);
    base = __shfl(base, 0);
    if (base >= a.n) return;
    const uint32_t i = base + lane;
    // A node without candidates pushes nothing: its pass leaves every
    // velocity as it is, so it takes no part in the dataflow.
    uint32_t cnt = 0, off = 0;
    if (i < a.n) {
      cnt = a.ccount[i];
      off = a.coff[i];
      if (cnt == 0) a.fin[i] = 1;
    }
    bool pending = cnt != 0;
    uint64_t waiting = cnt ? (cnt == 64 ? ~0ull : (1ull << cnt) - 1) : 0;  // candidates not yet ready
    bool self_ready = false;
    uint64_t t0 = wall_clock64();
    for (;;) {
      if (pending) {
        if (!self_ready) self_ready = ldu(a.done + i) == a.indeg[i];
        for (uint64_t m = waiting; m; m &= m - 1) {
          const uint32_t k = __builtin_ctzll(m);
          const uint32_t j = a.cand[(size_t)i * COLLIDE_K + k];
          if (ldu(a.done + j) == a.rank[off + k]) waiting &= ~(1ull << k);
        }
      }
      const bool go = pending && self_ready && waiting == 0;
      if (__ballot(go)) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        t0 = wall_clock64();
      }
      uint32_t fail = 0;
      bool parked = false;
      if (go) {
        // collide.zig:46-50: xi from the current vx_i.
        const double ri = a.radii[i], ri2 = ri * ri;
        double vxi = ld(a.vx + i), vyi = ld(a.vy + i);
        const double xi = a.x[i] + vxi, yi = a.y[i] + vyi;
        if (!(fabs(xi - a.px[i]) <= delta) || !(fabs(yi - a.py[i]) <= delta)) fail = COLLIDE_MOVED;
        // First the visits and overlap tests (the node's pushes change no
        // value these read: each j appears once and xi is fixed), so that a
        // node needing a jiggle stops before writing anything.
        uint64_t hit = 0;
        for (uint32_t k = 0; k < cnt && !fail && !parked; ++k) {
          const uint32_t j = a.cand[(size_t)i * COLLIDE_K + k];
          // The reference visits the leaf when no cell on its root path is
          // pruned; the cells above ctop pass for any xi in the box. Cells
          // are numbered in preorder with children in ascending order, so
          // the path follows the last child whose index is not above the
          // leaf.
          bool visited = true;
          const uint32_t leaf = a.cleaf[(size_t)i * COLLIDE_K + k];
          uint32_t cc = a.ctop[(size_t)i * COLLIDE_K + k];
          for (uint32_t step = 0; cc != NONE && cc != leaf; ++step) {
            const CRec q = a.rec[cc];
            if (step == COLLIDE_STACK || q.head != NONE) {
              fail = COLLIDE_FULL_STACK;
              break;
            }
            if (pruned(q, ri, xi, yi)) {
              visited = false;
              break;
            }
            uint32_t nx = NONE;
            for (int ch = 0; ch < 4; ++ch)
              if (q.child[ch] != NONE && q.child[ch] <= leaf) nx = q.child[ch];
            if (nx == NONE) {
              fail = COLLIDE_FULL_STACK;
              break;
            }
            cc = nx;
          }
          if (fail || !visited) continue;
          const double x = xi - a.x[j] - ld(a.vx + j), y = yi - a.y[j] - ld(a.vy + j);
          const double l = x * x + y * y, r = ri + a.radii[j];
          if (l < r * r) {
            if (x == 0 || y == 0) parked = true;  // collide.zig:62-69 draws from the host's random
            hit |= 1ull << k;
          }
        }
        if (parked && !fail) {
          atomicMin(a.stop, i);
        } else if (!fail) {
          // collide.zig:58-82, in candidate (walk) order
          for (uint64_t m = hit; m; m &= m - 1) {
            const uint32_t k = __builtin_ctzll(m);
            const uint32_t j = a.cand[(size_t)i * COLLIDE_K + k];
            double vxj = ld(a.vx + j), vyj = ld(a.vy + j);
            double x = xi - a.x[j] - vxj, y = yi - a.y[j] - vyj;
            double l = x * x + y * y;
            double rj = a.radii[j], r = ri + rj;
            l = __builtin_sqrt(l);
            l = (r - l) / l * a.strength;
            x *= l;
            y *= l;
            rj *= rj;
            r = rj / (ri2 + rj);
            vxi += x * r;
            vyi += y * r;
            r = 1 - r;
            vxj -= x * r;
            vyj -= y * r;
            st(a.vx + j, vxj);
            st(a.vy + j, vyj);
            if (!(fabs(a.x[j] + vxj - a.px[j]) <= delta) || !(fabs(a.y[j] + vyj - a.py[j]) <= delta))
              fail = COLLIDE_MOVED;
          }
          st(a.vx + i, vxi);
          st(a.vy + i, vyi);
        }
      }
      if (__ballot(fail != 0)) {
        if (fail) atomicOr(a.status, fail);
        return;
      }
      if (__ballot(go)) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (go) {
          if (!parked) {
            for (uint32_t k = 0; k < cnt; ++k)
              __hip_atomic_fetch_add(a.done + a.cand[(size_t)i * COLLIDE_K + k], 1u, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
            a.fin[i] = 1;
          }
          pending = false;
        }
      }
      if (__ballot(pending) == 0) break;
      if (__ballot(ldu(a.status) != 0)) return;
      // Once a node has stopped for a jiggle, nodes that wait on it would
      // wait for ever: a wave that makes no progress for 2 ms leaves its
      // nodes to the serial continuation.
      const uint64_t limit = ldu(a.stop) != NONE ? PARKED_LIMIT : WAIT_LIMIT;
      if (wall_clock64() - t0 > limit) {
        if (limit == WAIT_LIMIT && lane == 0) atomicOr(a.status, COLLIDE_STALLED);
        return;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
}

// Nodes the parallel pass left unfinished, in index order.
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_unfinished_flags(const CollideArgs a,
                                                                              uint32_t* __restrict__ flag) {
  const uint32_t i = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (i < a.n) flag[i] = a.fin[i] == 0 ? 1u : 0u;
}
__global__ __launch_bounds__(COLLIDE_THREADS) void k_collide_unfinished_list(const CollideArgs a,
                                                                             const uint32_t* __restrict__ flag,
                                                                             const uint32_t* __restrict__ pos,
                                                                             uint32_t* __restrict__ list) {
  const uint32_t i = blockIdx.x * COLLIDE_THREADS + threadIdx.x;
  if (i < a.n && flag[i]) list[pos[i]] = i;
}

// collide.zig:45-101 on one thread, resumable: before a pair that needs a
// jiggle it stores its walk and returns; the host draws (x first, then y, as
// collide.zig:62-69) and launches again.
// With a list it continues a parallel pass: it runs only the listed nodes
// (those the pass left unfinished, in index order) and checks that every push
// keeps its node inside the pass's box, which the finished nodes' results
// assume; a push outside it stops the kernel with need = COLLIDE_NEED_MOVED.
__global__ void k_collide_serial(const CollideArgs a, uint32_t* __restrict__ stack, CollideSerialState* st,
                                 const uint32_t* __restrict__ list, uint32_t count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t p = st->node, sp = st->sp;
  bool inside = st->inside != 0;
  double xi = st->xi, yi = st->yi;
  const Cell* cells = a.cells;
  const double delta = list ? collide_delta(a) : 0.0;
  while (p < count) {
    const uint32_t i = list ? list[p] : p;
    const double ri = a.radii[i], ri2 = ri * ri;
    if (!inside) {
      xi = a.x[i] + a.vx[i];
      yi = a.y[i] + a.vy[i];
      stack[0] = 0;
      sp = 1;
      inside = true;
    }
    while (sp > 0) {
      const uint32_t c = stack[--sp];
      const Cell& q = cells[c];
      double rj = q.r, r = ri + rj;
      if (is_leaf(q)) {
        const uint32_t j = a.perm[q.first + q.count - 1];
        if (j > i) {
          double x = xi - a.x[j] - a.vx[j], y = yi - a.y[j] - a.vy[j];
          double l = x * x + y * y;
          if (l < r * r) {
            if ((x == 0 || y == 0) && !st->has_jig) {
              st->need = (x == 0 ? 1u : 0u) | (y == 0 ? 2u : 0u);
              stack[sp++] = c;
              st->node = p;
              st->sp = sp;
              st->inside = 1;
              st->xi = xi;
              st->yi = yi;
              return;
            }
            if (x == 0) {
              x = st->jig[0];
              l += x * x;
            }
            if (y == 0) {
              y = st->jig[1];
              l += y * y;
            }
            st->has_jig = 0;
            l = __builtin_sqrt(l);
            l = (r - l) / l * a.strength;
            x *= l;
            y *= l;
            rj *= rj;
            r = rj / (ri2 + rj);
            a.vx[i] += x * r;
            a.vy[i] += y * r;
            r = 1 - r;
            a.vx[j] -= x * r;
            a.vy[j] -= y * r;
            if (list && (!(fabs(a.x[j] + a.vx[j] - a.px[j]) <= delta) ||
                         !(fabs(a.y[j] + a.vy[j] - a.py[j]) <= delta))) {
              st->need = COLLIDE_NEED_MOVED;
              st->node = p;
              return;
            }
          }
        }
        continue;
      }
      if (q.x0 > xi + r || q.x1 < xi - r || q.y0 > yi + r || q.y1 < yi - r) continue;
      for (int k = 3; k >= 0; --k)
        if (q.child[k] != NONE) stack[sp++] = q.child[k];
    }
    inside = false;
    ++p;
  }
  st->node = count;
  st->need = 0;
}

hipError_t collide_radii(TreeBuffers& t, const double* radii, hipStream_t stream) {
  const uint32_t len = t.meta_host->len;
  if (len == 0) return hipSuccess;
  HIPCHECK(tree_depth_order(t, stream));
  uint32_t blocks = ceil_div(len, COLLIDE_THREADS);
  if (blocks > 2048) blocks = 2048;
  for (uint32_t d = t.meta_host->depth + 1; d-- > 0;)
    hipLaunchKernelGGL(k_collide_radius_level, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, t.cells, t.dorder,
                       t.dstart, d, t.perm, radii);
  return hipGetLastError();
}

hipError_t collide_records(const CollideArgs& a, uint32_t len, hipStream_t stream) {
  if (len == 0) return hipSuccess;
  hipLaunchKernelGGL(k_collide_records, dim3(ceil_div(len, COLLIDE_THREADS)), dim3(COLLIDE_THREADS), 0, stream,
                     a.cells, len, a.perm, a.rec);
  return hipGetLastError();
}

hipError_t collide_candidates(const CollideArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  HIPCHECK(hipMemsetAsync(a.ccount, 0, (size_t)a.n * sizeof(uint32_t), stream));
  if (a.m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_collide_cand, dim3(ceil_div(a.m, COLLIDE_THREADS)), dim3(COLLIDE_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t collide_pairs(const CollideArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_collide_pairs, dim3(ceil_div(a.n, COLLIDE_THREADS)), dim3(COLLIDE_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t collide_ranks(const CollideArgs& a, uint32_t pairs, hipStream_t stream) {
  if (pairs == 0) return hipSuccess;
  const uint32_t blocks = ceil_div(pairs, COLLIDE_THREADS);
  hipLaunchKernelGGL(k_collide_runs, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, a, pairs);
  hipLaunchKernelGGL(k_collide_ranks, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, a, pairs);
  return hipGetLastError();
}

hipError_t collide_apply(const CollideArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  uint32_t blocks = ceil_div(a.n, COLLIDE_THREADS);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_collide_apply, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, a);
  return hipGetLastError();
}

hipError_t collide_serial(const CollideArgs& a, uint32_t* stack, CollideSerialState* st, const uint32_t* list,
                          uint32_t count, hipStream_t stream) {
  hipLaunchKernelGGL(k_collide_serial, dim3(1), dim3(64), 0, stream, a, stack, st, list, count);
  return hipGetLastError();
}

hipError_t collide_unfinished(const CollideArgs& a, uint32_t* flag, uint32_t* pos, uint32_t* list,
                              const uint32_t* n_dev, uint32_t* partials, uint32_t* count_out, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  const uint32_t blocks = ceil_div(a.n, COLLIDE_THREADS);
  hipLaunchKernelGGL(k_collide_unfinished_flags, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, a, flag);
  HIPCHECK(scan_u32(flag, pos, a.n, n_dev, partials, count_out, stream));
  hipLaunchKernelGGL(k_collide_unfinished_list, dim3(blocks), dim3(COLLIDE_THREADS), 0, stream, a, flag, pos, list);
  return hipGetLastError();
}

}  // namespace d3f
