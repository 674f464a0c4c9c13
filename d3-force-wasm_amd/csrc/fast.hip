// Fast-mode link and collide on gfx950 (src/core/link_fast.zig,
// src/core/grid.zig), in FP64 and bit-identical to the reference: both are
// Jacobi sweeps in which every node sums its own terms in a fixed order and
// writes only its own velocity, so one thread per node reproduces them.
// Nothing here synchronises with the host: the grid's shape is computed on
// the device and read by the kernels after it.
#include "common.h"

namespace d3f {
namespace {

constexpr int EW = 256;
constexpr uint32_t TARGET_BIT = 0x80000000u;  // link_fast.zig:20
constexpr uint32_t LONG_ROW = 64;             // rows longer than this are summed by a wave

// ---------------------------------------------------------------- link sweep

// The serial CSR fill (link_fast.zig:105-130) walks the edges in order and
// fills each row from its end, so a row lists its entries by descending
// edge, and a self-link's target entry precedes its source entry. Position
// q of the sort input holds entry q of that order; a stable sort by node
// then yields the rows.
__global__ __launch_bounds__(EW) void k_csr_entries(const uint32_t* __restrict__ src,
                                                    const uint32_t* __restrict__ tgt, uint32_t m,
                                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t q = blockIdx.x * EW + threadIdx.x;
  if (q >= 2 * m) return;
  const uint32_t e = m - 1 - q / 2;
  if (q & 1) {
    keys[q] = src[e];
    vals[q] = e;
  } else {
    keys[q] = tgt[e];
    vals[q] = e | TARGET_BIT;
  }
}

// row[i] = first sorted position whose node is >= i; row[n] = 2m.
__global__ __launch_bounds__(EW) void k_csr_rows(const uint64_t* __restrict__ keys, uint32_t total, uint32_t n,
                                                 uint32_t* __restrict__ row) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i > n) return;
  uint32_t lo = 0, hi = total;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < i) lo = mid + 1; else hi = mid;
  }
  row[i] = i == n ? total : lo;
}

// One entry of node i's row (link_fast.zig:70-97) as the signed term its
// velocity adds: vx -= x * b on the target side is vx += -(x * b), which
// rounds identically, so a row's terms can be formed in parallel and added
// in row order. bf is b on the target side and 1 - b on the source side.
__device__ inline void link_term(const LinkEntry& le, const double2* __restrict__ pos, double2 pi, double alpha,
                                 uint64_t seed, double& tx, double& ty) {
  const uint32_t e = le.e & ~TARGET_BIT;
  const bool is_target = (le.e & TARGET_BIT) != 0;
  const double2 pj = pos[le.j];
  double x = is_target ? pi.x - pj.x : pj.x - pi.x;
  double y = is_target ? pi.y - pj.y : pj.y - pi.y;
  if (x == 0 || x != x) x = hash_jiggle(seed, e, 0, 0);
  if (y == 0 || y != y) y = hash_jiggle(seed, e, 0, 1);
  double l = __builtin_sqrt(x * x + y * y);
  l = (l - le.distance) / l * alpha * le.strength;
  x *= l;
  y *= l;
  if (is_target) {
    tx = -(x * le.bf);
    ty = -(y * le.bf);
  } else {
    tx = x * le.bf;
    ty = y * le.bf;
  }
}

// Each adjacency entry's parameters in row order (rebuilt with the rows).
__global__ __launch_bounds__(EW) void k_csr_pack(uint32_t total, const uint64_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ adj, const uint32_t* __restrict__ src,
                                                 const uint32_t* __restrict__ tgt,
                                                 const double* __restrict__ strength,
                                                 const double* __restrict__ distance,
                                                 const double* __restrict__ bias,
                                                 const uint32_t* __restrict__ node_graph,
                                                 const uint32_t* __restrict__ link_base, uint32_t* __restrict__ owner,
                                                 LinkEntry* __restrict__ entries) {
  const uint32_t a = blockIdx.x * EW + threadIdx.x;
  if (a >= total) return;
  const uint32_t entry = adj[a];
  const uint32_t e = entry & ~TARGET_BIT;
  const bool is_target = (entry & TARGET_BIT) != 0;
  const double b = bias[e];
  const uint32_t i = (uint32_t)keys[a];
  // The jiggle hashes the edge's index within its graph (the whole link
  // array for one simulation).
  const uint32_t local = node_graph ? e - link_base[node_graph[i]] : e;
  entries[a] = LinkEntry{is_target ? src[e] : tgt[e], local | (entry & TARGET_BIT), strength[e], distance[e],
                         is_target ? b : 1 - b};
  owner[a] = i;
}

// Phase 1 of a sweep (link_fast.zig:60-99 on the snapshot pos = x + vx,
// y + vy): every adjacency entry's term for its row's node, all in parallel.
__global__ __launch_bounds__(EW) void k_link_terms(uint32_t total, const uint32_t* __restrict__ owner,
                                                   const LinkEntry* __restrict__ entries,
                                                   const double2* __restrict__ pos,
                                                   double2* __restrict__ terms, double alpha, uint64_t seed,
                                                   const uint32_t* __restrict__ node_graph,
                                                   const uint64_t* __restrict__ seeds,
                                                   const double* __restrict__ alpha_dev,
                                                   const uint64_t* __restrict__ seed_dev) {
  const uint32_t a = blockIdx.x * EW + threadIdx.x;
  if (a >= total) return;
  if (alpha_dev) alpha = *alpha_dev;
  if (seed_dev) seed = *seed_dev;
  const uint32_t i = owner[a];
  if (node_graph) seed = seeds[node_graph[i]];
  double tx, ty;
  link_term(entries[a], pos, pos[i], alpha, seed, tx, ty);
  terms[a] = make_double2(tx, ty);
}

// Phase 2: each node adds its row's terms in row order, as the serial walk
// does. Rows longer than LONG_ROW are left to k_link_rows_long.
__global__ __launch_bounds__(EW) void k_link_rows(uint32_t n, const uint32_t* __restrict__ row,
                                                  const double2* __restrict__ terms, double* __restrict__ vxs,
                                                  double* __restrict__ vys) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const uint32_t a0 = row[i], a1 = row[i + 1];
  if (a0 == a1 || a1 - a0 > LONG_ROW) return;
  double vx = vxs[i], vy = vys[i];
  for (uint32_t a = a0; a < a1; ++a) {
    const double2 t = terms[a];
    vx += t.x;
    vy += t.y;
  }
  vxs[i] = vx;
  vys[i] = vy;
}

// A long row (a hub of the graph): the wave stages 64 terms at a time in
// LDS, x and y apart, the next chunk's loads in flight; lane 0 adds the x
// terms and lane 1 the y terms, each in row order (one dependent add per
// entry and lane, both lanes in one instruction).
__global__ __launch_bounds__(64) void k_link_rows_long(const uint32_t* __restrict__ rows_long,
                                                       const uint32_t* __restrict__ row,
                                                       const double2* __restrict__ terms, double* __restrict__ vxs,
                                                       double* __restrict__ vys) {
  __shared__ double buf[2][64];
  const uint32_t i = rows_long[blockIdx.x];
  const int lane = threadIdx.x;
  const uint32_t a0 = row[i], a1 = row[i + 1];
  const int comp = lane & 1;  // lane 0: x, lane 1: y
  double v = comp ? vys[i] : vxs[i];
  double2 next = a0 + lane < a1 ? terms[a0 + lane] : make_double2(0, 0);
  for (uint32_t base = a0; base < a1; base += 64) {
    buf[0][lane] = next.x;
    buf[1][lane] = next.y;
    __syncthreads();
    const uint32_t nb = base + 64 + lane;
    if (nb < a1) next = terms[nb];
    if (lane < 2) {
      const double* b = buf[comp];
      const uint32_t k = a1 - base < 64 ? a1 - base : 64;
      uint32_t q = 0;
      for (; q + 8 <= k; q += 8) {
        const double t0 = b[q], t1 = b[q + 1], t2 = b[q + 2], t3 = b[q + 3];
        const double t4 = b[q + 4], t5 = b[q + 5], t6 = b[q + 6], t7 = b[q + 7];
        v += t0;
        v += t1;
        v += t2;
        v += t3;
        v += t4;
        v += t5;
        v += t6;
        v += t7;
      }
      for (; q < k; ++q) v += b[q];
    }
    __syncthreads();
  }
  if (lane == 0) vxs[i] = v;
  if (lane == 1) vys[i] = v;
}

// Nodes whose row is longer than LONG_ROW (any order: each is one block).
__global__ __launch_bounds__(EW) void k_long_rows(const uint32_t* __restrict__ row, uint32_t n,
                                                  uint32_t* __restrict__ rows_long, uint32_t* __restrict__ count) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i < n && row[i + 1] - row[i] > LONG_ROW) rows_long[atomicAdd(count, 1u)] = i;
}

// ---------------------------------------------------------------- grid collide

struct GridPartial {
  double x0, y0, x1, y1, rmax;
  uint32_t ix0, iy0, ix1, iy1, m, pad;
};

// The serial scan keeps the first of equal values (it replaces on a strict
// < or >), so ties go to the lower node.
__device__ inline void gmin(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv < v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void gmax(double& v, uint32_t& i, double bv, uint32_t bi) {
  if (bv > v || (bv == v && bi < i)) { v = bv; i = bi; }
}
__device__ inline void gcombine(GridPartial& a, const GridPartial& b) {
  gmin(a.x0, a.ix0, b.x0, b.ix0);
  gmin(a.y0, a.iy0, b.y0, b.iy0);
  gmax(a.x1, a.ix1, b.x1, b.ix1);
  gmax(a.y1, a.iy1, b.y1, b.iy1);
  if (b.rmax > a.rmax) a.rmax = b.rmax;
  a.m += b.m;
}
__device__ inline GridPartial gempty() {
  const double inf = __builtin_huge_val();
  return GridPartial{inf, inf, -inf, -inf, 0, NONE, NONE, NONE, NONE, 0, 0};
}

__device__ inline GridPartial block_combine(GridPartial p) {
  __shared__ GridPartial sh[EW];
  sh[threadIdx.x] = p;
  __syncthreads();
  for (int s = EW / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) gcombine(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  return sh[0];
}

// grid.zig:66-91: shifted positions, their bounds and the largest radius
// over the points without a NaN coordinate.
__global__ __launch_bounds__(EW) void k_grid_bounds(const double* __restrict__ xs, const double* __restrict__ ys,
                                                    const double* __restrict__ vxs, const double* __restrict__ vys,
                                                    const double* __restrict__ radii, uint32_t n,
                                                    double* __restrict__ px, double* __restrict__ py,
                                                    GridPartial* __restrict__ partials) {
  GridPartial p = gempty();
  for (uint32_t i = blockIdx.x * EW + threadIdx.x; i < n; i += gridDim.x * EW) {
    const double x = xs[i] + vxs[i], y = ys[i] + vys[i];
    px[i] = x;
    py[i] = y;
    if (x != x || y != y) continue;
    gmin(p.x0, p.ix0, x, i);
    gmin(p.y0, p.iy0, y, i);
    gmax(p.x1, p.ix1, x, i);
    gmax(p.y1, p.iy1, y, i);
    const double rad = radii[i];
    if (rad > p.rmax) p.rmax = rad;
    p.m++;
  }
  p = block_combine(p);
  if (threadIdx.x == 0) partials[blockIdx.x] = p;
}

// grid.zig:224-253: the grid's shape. Cells are 2 r_max wide, doubled until
// there are at most 4n + 16 of them; `active` is 0 when the iteration does
// nothing (fewer than two points or no positive radius).
__global__ __launch_bounds__(EW) void k_grid_setup(const GridPartial* __restrict__ partials, uint32_t blocks,
                                                   uint32_t n, GridParams* __restrict__ gp) {
  GridPartial q = gempty();
  for (uint32_t b = threadIdx.x; b < blocks; b += EW) gcombine(q, partials[b]);
  q = block_combine(q);
  if (threadIdx.x != 0) return;
  GridParams r{};
  r.m = q.m;
  r.active = q.m >= 2 && q.rmax > 0;
  if (r.active) {
    double h = 2 * q.rmax;
    uint32_t gx = (uint32_t)(floor((q.x1 - q.x0) / h) + 1);
    uint32_t gy = (uint32_t)(floor((q.y1 - q.y0) / h) + 1);
    while ((uint64_t)gx * gy > 4 * (uint64_t)n + 16) {
      h *= 2;
      gx = (uint32_t)(floor((q.x1 - q.x0) / h) + 1);
      gy = (uint32_t)(floor((q.y1 - q.y0) / h) + 1);
    }
    r.x0 = q.x0;
    r.y0 = q.y0;
    r.inv_h = 1 / h;
    r.gx = gx;
    r.gy = gy;
    r.cells = gx * gy;
  }
  r.sort_n = r.active ? n : 0;
  *gp = r;
}

// grid.zig:95-111: sort input in descending node order keyed by cell
// (`cells`, sorting last, for a NaN position).
__global__ __launch_bounds__(EW) void k_grid_bin(const double* __restrict__ px, const double* __restrict__ py,
                                                 uint32_t n, const GridParams* __restrict__ gp,
                                                 uint64_t* __restrict__ keys, uint32_t* __restrict__ items) {
  const uint32_t k = blockIdx.x * EW + threadIdx.x;
  if (k >= n || !gp->active) return;
  const uint32_t i = n - 1 - k;
  const uint32_t cells = gp->cells;
  uint32_t cell = cells;
  if (px[i] == px[i] && py[i] == py[i]) {
    const uint32_t cx = (uint32_t)floor((px[i] - gp->x0) * gp->inv_h);
    const uint32_t cy = (uint32_t)floor((py[i] - gp->y0) * gp->inv_h);
    cell = cy * gp->gx + cx;
  }
  keys[k] = cell;
  items[k] = i;
}

// grid.zig:115-127 and 255-258: start[c] = first sorted position whose cell
// is >= c (m for the cells after the last binned node, and for `cells`).
__global__ __launch_bounds__(EW) void k_grid_starts(const uint64_t* __restrict__ keys,
                                                    const GridParams* __restrict__ gp, uint32_t* __restrict__ start) {
  const uint32_t c = blockIdx.x * EW + threadIdx.x;
  if (!gp->active) return;
  const uint32_t cells = gp->cells;
  if (c > cells) return;
  if (c == cells) {
    start[c] = gp->m;
    return;
  }
  uint32_t lo = 0, hi = gp->m;  // keys[m..n) hold `cells`
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (keys[mid] < c) lo = mid + 1; else hi = mid;
  }
  start[c] = lo;
}

// The binned nodes' positions and radii in sorted (cell) order, so a cell's
// candidates are contiguous.
__global__ __launch_bounds__(EW) void k_grid_gather(const GridParams* __restrict__ gp,
                                                    const uint32_t* __restrict__ items, const double* __restrict__ px,
                                                    const double* __restrict__ py, const double* __restrict__ radii,
                                                    double4* __restrict__ cand) {
  const uint32_t k = blockIdx.x * EW + threadIdx.x;
  if (!gp->active || k >= gp->m) return;
  const uint32_t j = items[k];
  cand[k] = make_double4(px[j], py[j], radii[j], __longlong_as_double((long long)j));
}

// grid.zig:129-195: node i scans its 3 x 3 cells (rows, then columns, then
// each cell's nodes high index first) and adds its share of each push.
// Threads run in cell order (thread k takes the k-th binned node), so a
// wave's nodes share their candidate cells; each node's own sum is the
// reference's, term for term.
__global__ __launch_bounds__(EW) void k_grid_sweep(const GridParams* __restrict__ gp,
                                                   const uint64_t* __restrict__ keys,
                                                   const uint32_t* __restrict__ start,
                                                   const double4* __restrict__ cand, double* __restrict__ vxs,
                                                   double* __restrict__ vys, double strength, uint64_t seed) {
  const uint32_t k = blockIdx.x * EW + threadIdx.x;
  if (!gp->active || k >= gp->m) return;
  const uint32_t ci = (uint32_t)keys[k];
  const double4 me = cand[k];
  const uint32_t i = (uint32_t)__double_as_longlong(me.w);
  const uint32_t gx = gp->gx, gy = gp->gy;
  const double ri = me.z;
  const double ri2 = ri * ri;
  const double xi = me.x, yi = me.y;
  double vx = vxs[i], vy = vys[i];
  const uint32_t cxi = ci % gx, cyi = ci / gx;
  const uint32_t xa = cxi > 0 ? cxi - 1 : 0, xb = cxi + 1 < gx - 1 ? cxi + 1 : gx - 1;
  const uint32_t ya = cyi > 0 ? cyi - 1 : 0, yb = cyi + 1 < gy - 1 ? cyi + 1 : gy - 1;
  for (uint32_t cy = ya; cy <= yb; ++cy) {
    // The cells xa..xb of a row are consecutive ids, so their nodes are one
    // run of the sorted items, in the reference's cell-then-node order.
    const uint32_t a1 = start[cy * gx + xb + 1];
    for (uint32_t a = start[cy * gx + xa]; a < a1; ++a) {
      const double4 c = cand[a];
      const uint32_t j = (uint32_t)__double_as_longlong(c.w);
      if (j == i) continue;
      const double rj = c.z;
      double rr = ri + rj;
      double x = xi - c.x, y = yi - c.y;
      double l = x * x + y * y;
      if (l < rr * rr) {
        if (x == 0 || y == 0) {
          const uint32_t lo = i < j ? i : j, hi = i < j ? j : i;
          const double sign = i == lo ? 1.0 : -1.0;
          if (x == 0) {
            x = sign * hash_jiggle(seed, lo, hi, 0);
            l += x * x;
          }
          if (y == 0) {
            y = sign * hash_jiggle(seed, lo, hi, 1);
            l += y * y;
          }
        }
        l = __builtin_sqrt(l);
        l = (rr - l) / l * strength;
        x *= l;
        y *= l;
        rr = rj * rj / (ri2 + rj * rj);
        vx += x * rr;
        vy += y * rr;
      }
    }
  }
  vxs[i] = vx;
  vys[i] = vy;
}

constexpr uint32_t GRID_BLOCKS = 512;

}  // namespace

hipError_t link_csr(const LinkCsrArgs& a, const SortBuffers& sb, uint32_t* long_count, hipStream_t stream) {
  const uint32_t total = 2 * a.m, n = a.n;
  if (total) {
    hipLaunchKernelGGL(k_csr_entries, dim3(ceil_div(total, EW)), dim3(EW), 0, stream, a.src, a.tgt, a.m, sb.k0,
                       a.adj);
    HIPCHECK(hipMemcpyAsync(sb.count, &total, sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    int digits = 0;
    for (uint64_t v = n; v > 0; v >>= 8) ++digits;
    HIPCHECK(sort_digits(sb.hist, sb.glob, sb.k0, a.adj, sb.k1, sb.v1, total, sb.count, 0, digits, stream));
    hipLaunchKernelGGL(k_csr_pack, dim3(ceil_div(total, EW)), dim3(EW), 0, stream, total, sb.k0, a.adj, a.src, a.tgt,
                       a.strength, a.distance, a.bias, a.node_graph, a.link_base, a.owner, a.entries);
  }
  hipLaunchKernelGGL(k_csr_rows, dim3(ceil_div(n + 1, EW)), dim3(EW), 0, stream, sb.k0, total, n, a.row);
  HIPCHECK(hipMemsetAsync(sb.count, 0, sizeof(uint32_t), stream));
  if (n) hipLaunchKernelGGL(k_long_rows, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, a.row, n, a.rows_long, sb.count);
  HIPCHECK(hipMemcpyAsync(long_count, sb.count, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipGetLastError());
  return hipStreamSynchronize(stream);  // `total` and *long_count live on the host
}

hipError_t link_sweep(const LinkCsrArgs& a, uint32_t long_count, const double2* pos, double* vx, double* vy,
                      double alpha, uint64_t seed, hipStream_t stream, const double* alpha_dev,
                      const uint64_t* seed_dev) {
  const uint32_t total = 2 * a.m;
  if (a.n == 0 || total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_link_terms, dim3(ceil_div(total, EW)), dim3(EW), 0, stream, total, a.owner, a.entries, pos,
                     (double2*)a.terms, alpha, seed, a.node_graph, a.seeds, alpha_dev, seed_dev);
  hipLaunchKernelGGL(k_link_rows, dim3(ceil_div(a.n, EW)), dim3(EW), 0, stream, a.n, a.row, (const double2*)a.terms,
                     vx, vy);
  if (long_count)
    hipLaunchKernelGGL(k_link_rows_long, dim3(long_count), dim3(64), 0, stream, a.rows_long, a.row,
                       (const double2*)a.terms, vx, vy);
  return hipGetLastError();
}

size_t grid_partials_bytes() { return GRID_BLOCKS * sizeof(GridPartial); }

hipError_t grid_collide_iteration(const double* x, const double* y, double* vx, double* vy, const double* radii,
                                  uint32_t n, double strength, uint64_t seed, const GridBuffers& g,
                                  const SortBuffers& sb, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  uint32_t blocks = ceil_div(n, EW * 4);
  if (blocks > GRID_BLOCKS) blocks = GRID_BLOCKS;
  const uint32_t nb = ceil_div(n, EW);
  hipLaunchKernelGGL(k_grid_bounds, dim3(blocks), dim3(EW), 0, stream, x, y, vx, vy, radii, n, g.px, g.py,
                     (GridPartial*)g.partials);
  hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(EW), 0, stream, (const GridPartial*)g.partials, blocks, n,
                     g.params);
  hipLaunchKernelGGL(k_grid_bin, dim3(nb), dim3(EW), 0, stream, g.px, g.py, n, g.params, sb.k0, sb.v0);
  // Cell ids are at most 4n + 16 (grid.zig:243-250); the count is 0 when
  // the iteration is inactive.
  int digits = 0;
  for (uint64_t v = 4 * (uint64_t)n + 16; v > 0; v >>= 8) ++digits;
  HIPCHECK(sort_digits(sb.hist, sb.glob, sb.k0, sb.v0, sb.k1, sb.v1, n, &g.params->sort_n, 0, digits, stream));
  hipLaunchKernelGGL(k_grid_starts, dim3(ceil_div(4 * n + 17, EW)), dim3(EW), 0, stream, sb.k0, g.params, g.start);
  hipLaunchKernelGGL(k_grid_gather, dim3(nb), dim3(EW), 0, stream, g.params, sb.v0, g.px, g.py, radii, g.cand);
  hipLaunchKernelGGL(k_grid_sweep, dim3(nb), dim3(EW), 0, stream, g.params, sb.k0, g.start, g.cand, vx, vy, strength,
                     seed);
  return hipGetLastError();
}

}  // namespace d3f
