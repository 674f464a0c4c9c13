// center, x, y, radial, integrate and the shifted positions on gfx950
// (reference: src/core/simple.zig, src/core/sim.zig:129-146). All but center
// are per node. forceCenter's mean is d3's serial sum over nodes 0..n-1
// (simple.zig:18-22), which no reassociation may change: one workgroup stages
// x and y through LDS and one lane per axis adds them in index order.
#include "common.h"

namespace d3f {
namespace {

constexpr int EW = 256;

// d3's center sums: sx = x0 + x1 + ... + x(n-1) in index order (and sy),
// simple.zig:18-22. The chain of dependent FP64 adds is the whole cost, so
// each axis gets its own SIMD: of the 8 waves (wave w issues on SIMD w % 4)
// wave 0 adds x and wave 1 adds y, one lane each, from a double-buffered LDS
// chunk that waves 2, 3, 6 and 7 (SIMDs 2 and 3) fill with the next chunk
// meanwhile; waves 4 and 5 only keep the barrier. Each lane loads 32 values
// at once, then adds them in order.
constexpr int CS_THREADS = 512;
constexpr int CS_CHUNK = 2048;
constexpr int CS_PRODUCERS = 256;  // threads of waves 2, 3, 6, 7
constexpr int CS_BLOCK = 32;

__device__ inline bool cs_producer(int wave) { return (wave & 3) >= 2; }
__device__ inline int cs_slot(int wave, int lane) { return ((wave >> 2) * 2 + ((wave & 3) - 2)) * 64 + lane; }

__device__ inline void cs_load(const double* __restrict__ xs, const double* __restrict__ ys, uint32_t n,
                               uint32_t base, double (*buf)[CS_CHUNK], int tid) {
  const uint32_t cnt = n - base < (uint32_t)CS_CHUNK ? n - base : (uint32_t)CS_CHUNK;
  constexpr int U = CS_CHUNK / CS_PRODUCERS;
  double vx[U], vy[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t k = tid + u * CS_PRODUCERS;
    vx[u] = k < cnt ? xs[base + k] : 0;
    vy[u] = k < cnt ? ys[base + k] : 0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t k = tid + u * CS_PRODUCERS;
    if (k < cnt) {
      buf[0][k] = vx[u];
      buf[1][k] = vy[u];
    }
  }
}

__device__ inline double cs_add(const double* p, uint32_t cnt, double s) {
  if (cnt == (uint32_t)CS_CHUNK) {
    for (uint32_t k = 0; k < (uint32_t)CS_CHUNK; k += CS_BLOCK) {
      double a[CS_BLOCK];
#pragma unroll
      for (int j = 0; j < CS_BLOCK; ++j) a[j] = p[k + j];
#pragma unroll
      for (int j = 0; j < CS_BLOCK; ++j) s += a[j];
    }
  } else {
    for (uint32_t k = 0; k < cnt; ++k) s += p[k];
  }
  return s;
}

__global__ __launch_bounds__(CS_THREADS) void k_center_raw_sum(const double* __restrict__ xs,
                                                               const double* __restrict__ ys, uint32_t n,
                                                               double* __restrict__ out) {
  __shared__ double buf[2][2][CS_CHUNK];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const uint32_t chunks = ceil_div(n, CS_CHUNK);
  double s = 0;
  if (chunks && cs_producer(wave)) cs_load(xs, ys, n, 0, buf[0], cs_slot(wave, lane));
  __syncthreads();
  for (uint32_t c = 0; c < chunks; ++c) {
    const uint32_t base = c * CS_CHUNK;
    const uint32_t cnt = n - base < (uint32_t)CS_CHUNK ? n - base : (uint32_t)CS_CHUNK;
    if (cs_producer(wave)) {
      if (c + 1 < chunks) cs_load(xs, ys, n, base + CS_CHUNK, buf[(c + 1) & 1], cs_slot(wave, lane));
    } else if (wave < 2 && lane == 0) {
      s = cs_add(buf[c & 1][wave], cnt, s);
    }
    __syncthreads();
  }
  if (wave < 2 && lane == 0) out[wave] = s;
}

// A batch of graphs: one wave per graph; lane 0 adds x and lane 1 adds y in
// node order (simple.zig:18-22) from LDS, the next 64 values already in
// flight.
__global__ __launch_bounds__(64) void k_batch_center_sums(const double* __restrict__ xs,
                                                          const double* __restrict__ ys,
                                                          const uint32_t* __restrict__ node_off,
                                                          double* __restrict__ sums) {
  __shared__ double buf[2][64];
  const uint32_t g = blockIdx.x;
  const int lane = threadIdx.x, axis = lane & 1;
  const uint32_t lo = node_off[g], hi = node_off[g + 1];
  double s = 0;
  double nx = lo + lane < hi ? xs[lo + lane] : 0, ny = lo + lane < hi ? ys[lo + lane] : 0;
  for (uint32_t base = lo; base < hi; base += 64) {
    buf[0][lane] = nx;
    buf[1][lane] = ny;
    __syncthreads();
    const uint32_t k = base + 64 + lane;
    nx = k < hi ? xs[k] : 0;
    ny = k < hi ? ys[k] : 0;
    if (lane < 2) {
      const uint32_t cnt = hi - base < 64 ? hi - base : 64;
      for (uint32_t q = 0; q < cnt; ++q) s += buf[axis][q];
    }
    __syncthreads();
  }
  if (lane < 2) sums[2 * g + axis] = s;
}

__global__ __launch_bounds__(EW) void k_batch_center_apply(double* __restrict__ xs, double* __restrict__ ys,
                                                           uint32_t n, const uint32_t* __restrict__ node_off,
                                                           const uint32_t* __restrict__ node_graph,
                                                           const double* __restrict__ sums, double cx, double cy,
                                                           double strength) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const uint32_t g = node_graph[i];
  const double nf = (double)(node_off[g + 1] - node_off[g]);
  xs[i] -= (sums[2 * g] / nf - cx) * strength;
  ys[i] -= (sums[2 * g + 1] / nf - cy) * strength;
}

// x -= (sx / n - cx) * strength, y likewise (simple.zig:23-30); every thread
// forms the same shift from the same operands.
__global__ __launch_bounds__(EW) void k_center_apply(double* __restrict__ xs, double* __restrict__ ys, uint32_t n,
                                                     const double* __restrict__ sums, double cx, double cy,
                                                     double strength) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const double nf = (double)n;
  const double sx = (sums[0] / nf - cx) * strength;
  const double sy = (sums[1] / nf - cy) * strength;
  xs[i] -= sx;
  ys[i] -= sy;
}

// simple.zig:90-123: `x += vx *= decay`, or `x = fx, vx = 0` on a fixed axis.
__global__ __launch_bounds__(EW) void k_integrate(double* __restrict__ xs, double* __restrict__ ys,
                                                  double* __restrict__ vxs, double* __restrict__ vys,
                                                  const double* __restrict__ fxs, const double* __restrict__ fys,
                                                  const uint8_t* __restrict__ fixed, uint32_t n, double decay) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  const uint8_t f = fixed[i];
  if ((f & 1) == 0) {
    const double v = vxs[i] * decay;
    vxs[i] = v;
    xs[i] = xs[i] + v;
  } else {
    xs[i] = fxs[i];
    vxs[i] = 0;
  }
  if ((f & 2) == 0) {
    const double v = vys[i] * decay;
    vys[i] = v;
    ys[i] = ys[i] + v;
  } else {
    ys[i] = fys[i];
    vys[i] = 0;
  }
}

// simple.zig:51-62: v += (target - p) * strength * alpha.
__global__ __launch_bounds__(EW) void k_axis(const double* __restrict__ pos, double* __restrict__ v,
                                             const double* __restrict__ strength,
                                             const double* __restrict__ target, uint32_t n, double alpha) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  v[i] += (target[i] - pos[i]) * strength[i] * alpha;
}

// simple.zig:64-86 (`a || 1e-6`, never jiggle).
__global__ __launch_bounds__(EW) void k_radial(const double* __restrict__ xs, const double* __restrict__ ys,
                                               double* __restrict__ vxs, double* __restrict__ vys,
                                               const double* __restrict__ strength,
                                               const double* __restrict__ radius, uint32_t n, double cx, double cy,
                                               double alpha) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  double dx = xs[i] - cx;
  if (dx == 0 || dx != dx) dx = 1e-6;
  double dy = ys[i] - cy;
  if (dy == 0 || dy != dy) dy = 1e-6;
  const double l = __builtin_sqrt(dx * dx + dy * dy);
  const double k = (radius[i] - l) * strength[i] * alpha / l;
  vxs[i] += dx * k;
  vys[i] += dy * k;
}

__global__ __launch_bounds__(EW) void k_shift(const double* __restrict__ xs, const double* __restrict__ ys,
                                              const double* __restrict__ vxs, const double* __restrict__ vys,
                                              double* __restrict__ px, double* __restrict__ py, uint32_t n) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  px[i] = xs[i] + vxs[i];
  py[i] = ys[i] + vys[i];
}

}  // namespace

hipError_t batch_center(double* x, double* y, const uint32_t* node_off, const uint32_t* node_graph, uint32_t graphs,
                        uint32_t n, double* sums, double cx, double cy, double strength, hipStream_t stream) {
  if (graphs == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_center_sums, dim3(graphs), dim3(64), 0, stream, x, y, node_off, sums);
  hipLaunchKernelGGL(k_batch_center_apply, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, n, node_off, node_graph,
                     sums, cx, cy, strength);
  return hipGetLastError();
}

hipError_t center_raw_sum(const double* x, const double* y, uint32_t n, double* sums, hipStream_t stream) {
  hipLaunchKernelGGL(k_center_raw_sum, dim3(1), dim3(CS_THREADS), 0, stream, x, y, n, sums);
  return hipGetLastError();
}

hipError_t center_apply(double* x, double* y, uint32_t n, const double* sums, double cx, double cy, double strength,
                        hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_center_apply, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, n, sums, cx, cy, strength);
  return hipGetLastError();
}

hipError_t integrate(double* x, double* y, double* vx, double* vy, const double* fx, const double* fy,
                     const uint8_t* fixed, uint32_t n, double decay, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_integrate, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, fx, fy, fixed, n, decay);
  return hipGetLastError();
}

hipError_t axis_force(const double* pos, double* v, const double* strength, const double* target, uint32_t n,
                      double alpha, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axis, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, pos, v, strength, target, n, alpha);
  return hipGetLastError();
}

hipError_t radial_force(const double* x, const double* y, double* vx, double* vy, const double* strength,
                        const double* radius, uint32_t n, double cx, double cy, double alpha, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_radial, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, strength, radius, n, cx,
                     cy, alpha);
  return hipGetLastError();
}

hipError_t shift_positions(const double* x, const double* y, const double* vx, const double* vy, double* px,
                           double* py, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, px, py, n);
  return hipGetLastError();
}

}  // namespace d3f
