// forceLink, exact mode, on gfx950 (reference: src/core/link.zig, d3's
// Gauss-Seidel pass over edges in user order).
//
// Dependency-level schedule: level(e) = 1 + max(level of the previous edge at
// its source, level of the previous edge at its target). Edges of one level
// share no node, so they commute bit for bit, and running the levels in
// order gives every node its edges in user order with the same operands as
// the serial pass. The schedule depends only on source/target and is built
// when the links change (runtime.hip); edge arrays are stored level-major so
// each level is one coalesced launch. The levels run on a packed copy of the
// node state (x, y, vx, vy in one 32-byte record per node): an edge gathers
// two records and writes two velocity pairs instead of touching twelve
// separate words.
//
// A zero or NaN delta makes d3 draw from random in edge order
// (link.zig:23-27). The level pass flags it and the packed velocities are not
// copied back; the host then runs k_link_serial on the untouched velocities,
// which stops before each such edge so the host can draw in order and resume.
#include "common.h"

namespace d3f {
namespace {

constexpr int LINK_THREADS = 256;

__global__ __launch_bounds__(LINK_THREADS) void k_link_pack(const double* __restrict__ xs,
                                                             const double* __restrict__ ys,
                                                             const double* __restrict__ vxs,
                                                             const double* __restrict__ vys, uint32_t n,
                                                             LinkNode* __restrict__ nodes) {
  const uint32_t i = blockIdx.x * LINK_THREADS + threadIdx.x;
  if (i >= n) return;
  LinkNode r;
  r.x = xs[i];
  r.y = ys[i];
  r.vx = vxs[i];
  r.vy = vys[i];
  nodes[i] = r;
}

// One dependency level (link.zig:14-38 per edge): the edges of a level share
// no node. A self-link's delta, (x + vx) - x - vx, is a rounding residue that
// need not be zero; its source update then applies to the updated target
// velocity, as in the serial pass.
__device__ __forceinline__ void link_edge(const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt,
                                          const double* __restrict__ strength, const double* __restrict__ distance,
                                          const double* __restrict__ bias, uint32_t e, LinkNode* __restrict__ nodes,
                                          double alpha, uint32_t* __restrict__ zero_flag) {
  const uint32_t s = src[e], t = tgt[e];
  const LinkNode a = nodes[s], c = nodes[t];
  double x = c.x + c.vx - a.x - a.vx;
  double y = c.y + c.vy - a.y - a.vy;
  if (x == 0 || x != x || y == 0 || y != y) {
    atomicOr(zero_flag, 1u);
    return;
  }
  double l = __builtin_sqrt(x * x + y * y);
  l = (l - distance[e]) / l * alpha * strength[e];
  x *= l;
  y *= l;
  double b = bias[e];
  const double2 vt = make_double2(c.vx - x * b, c.vy - y * b);
  b = 1 - b;
  const double2 v0 = s == t ? vt : make_double2(a.vx, a.vy);
  const double2 vs = make_double2(v0.x + x * b, v0.y + y * b);
  if (s != t) *reinterpret_cast<double2*>(&nodes[t].vx) = vt;
  *reinterpret_cast<double2*>(&nodes[s].vx) = vs;
}

__global__ __launch_bounds__(LINK_THREADS) void k_link_level(
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt, const double* __restrict__ strength,
    const double* __restrict__ distance, const double* __restrict__ bias, uint32_t lo, uint32_t hi,
    LinkNode* __restrict__ nodes, double alpha, uint32_t* __restrict__ zero_flag, const double* __restrict__ alpha_dev) {
  const uint32_t e = lo + blockIdx.x * LINK_THREADS + threadIdx.x;
  if (e >= hi) return;
  if (alpha_dev) alpha = *alpha_dev;
  link_edge(src, tgt, strength, distance, bias, e, nodes, alpha, zero_flag);
}

// A run of consecutive small levels in one workgroup, a barrier between
// levels (the workgroup's waves share one CU and its cache, so a level's
// writes are seen by the next after the barrier): one launch instead of one
// per level for the deep levels, which hold few links each.
constexpr int RUN_THREADS = 1024;
constexpr uint32_t RUN_MAX_LEVELS = 62;
constexpr uint32_t RUN_MAX_EDGES = 2 * RUN_THREADS;  // levels of at most this many links join a run
struct LevelRun {
  uint32_t count;
  uint32_t off[RUN_MAX_LEVELS + 1];
};

__global__ __launch_bounds__(RUN_THREADS) void k_link_run(
    const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt, const double* __restrict__ strength,
    const double* __restrict__ distance, const double* __restrict__ bias, LevelRun run,
    LinkNode* __restrict__ nodes, double alpha, uint32_t* __restrict__ zero_flag, const double* __restrict__ alpha_dev) {
  if (alpha_dev) alpha = *alpha_dev;
  for (uint32_t l = 0; l < run.count; ++l) {
    for (uint32_t e = run.off[l] + threadIdx.x; e < run.off[l + 1]; e += RUN_THREADS)
      link_edge(src, tgt, strength, distance, bias, e, nodes, alpha, zero_flag);
    __syncthreads();
  }
}

__global__ __launch_bounds__(LINK_THREADS) void k_link_unpack(const LinkNode* __restrict__ nodes, uint32_t n,
                                                               const uint32_t* __restrict__ zero_flag,
                                                               double* __restrict__ vxs, double* __restrict__ vys) {
  const uint32_t i = blockIdx.x * LINK_THREADS + threadIdx.x;
  if (i >= n || *zero_flag) return;
  const double2 v = *reinterpret_cast<const double2*>(&nodes[i].vx);
  vxs[i] = v.x;
  vys[i] = v.y;
}

// The serial pass from state->edge, stopping before an edge that needs a
// draw unless state carries its jiggle values (one thread).
__global__ void k_link_serial(const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt,
                              const double* __restrict__ strength, const double* __restrict__ distance,
                              const double* __restrict__ bias, uint32_t m, const double* __restrict__ xs,
                              const double* __restrict__ ys, double* __restrict__ vxs, double* __restrict__ vys,
                              double alpha, LinkSerialState* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  uint32_t e = st->edge;
  bool has_jig = st->has_jig != 0;
  for (; e < m; ++e) {
    const uint32_t s = src[e], t = tgt[e];
    double x = xs[t] + vxs[t] - xs[s] - vxs[s];
    double y = ys[t] + vys[t] - ys[s] - vys[s];
    const bool nx = x == 0 || x != x, ny = y == 0 || y != y;
    if (nx || ny) {
      if (!has_jig) {
        st->edge = e;
        st->need = (nx ? 1u : 0u) | (ny ? 2u : 0u);
        st->has_jig = 0;
        return;
      }
      if (nx) x = st->jig[0];
      if (ny) y = st->jig[1];
    }
    has_jig = false;
    double l = __builtin_sqrt(x * x + y * y);
    l = (l - distance[e]) / l * alpha * strength[e];
    x *= l;
    y *= l;
    double b = bias[e];
    vxs[t] -= x * b;
    vys[t] -= y * b;
    b = 1 - b;
    vxs[s] += x * b;
    vys[s] += y * b;
  }
  st->edge = m;
  st->need = 0;
  st->has_jig = 0;
}

}  // namespace

hipError_t link_pack(const LinkLevelArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_link_pack, dim3(ceil_div(a.n, LINK_THREADS)), dim3(LINK_THREADS), 0, stream, a.x, a.y, a.vx,
                     a.vy, a.n, a.nodes);
  return hipGetLastError();
}

hipError_t link_level(const LinkLevelArgs& a, uint32_t lo, uint32_t hi, hipStream_t stream) {
  if (hi <= lo) return hipSuccess;
  hipLaunchKernelGGL(k_link_level, dim3(ceil_div(hi - lo, LINK_THREADS)), dim3(LINK_THREADS), 0, stream, a.src,
                     a.tgt, a.strength, a.distance, a.bias, lo, hi, a.nodes, a.alpha, a.zero_flag, a.alpha_dev);
  return hipGetLastError();
}

hipError_t link_levels(const LinkLevelArgs& a, const uint32_t* off, uint32_t levels, hipStream_t stream) {
  uint32_t l = 1;
  while (l <= levels) {
    if (off[l + 1] - off[l] > RUN_MAX_EDGES) {
      HIPCHECK(link_level(a, off[l], off[l + 1], stream));
      ++l;
      continue;
    }
    LevelRun run;
    run.count = 0;
    run.off[0] = off[l];
    while (l <= levels && run.count < RUN_MAX_LEVELS && off[l + 1] - off[l] <= RUN_MAX_EDGES) {
      run.off[++run.count] = off[l + 1];
      ++l;
    }
    if (run.count == 1) {
      HIPCHECK(link_level(a, run.off[0], run.off[1], stream));
      continue;
    }
    hipLaunchKernelGGL(k_link_run, dim3(1), dim3(RUN_THREADS), 0, stream, a.src, a.tgt, a.strength, a.distance,
                       a.bias, run, a.nodes, a.alpha, a.zero_flag, a.alpha_dev);
    HIPCHECK(hipGetLastError());
  }
  return hipSuccess;
}

hipError_t link_unpack(const LinkLevelArgs& a, hipStream_t stream) {
  if (a.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_link_unpack, dim3(ceil_div(a.n, LINK_THREADS)), dim3(LINK_THREADS), 0, stream, a.nodes, a.n,
                     a.zero_flag, a.vx, a.vy);
  return hipGetLastError();
}

hipError_t link_serial(const uint32_t* src, const uint32_t* tgt, const double* strength, const double* distance,
                       const double* bias, uint32_t m, const double* x, const double* y, double* vx, double* vy,
                       double alpha, LinkSerialState* state, hipStream_t stream) {
  hipLaunchKernelGGL(k_link_serial, dim3(1), dim3(64), 0, stream, src, tgt, strength, distance, bias, m, x, y, vx,
                     vy, alpha, state);
  return hipGetLastError();
}

}  // namespace d3f
