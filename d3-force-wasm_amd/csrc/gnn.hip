// Learned warm start: a message-passing network that predicts a layout, with
// hand-written forward and backward kernels on gfx950's BF16 matrix cores
// (no reference counterpart; the reference always starts from d3's
// phyllotaxis).
//
//   H1    = relu(X W0 + b0)                          X: n x 8 node features
//   H_l+1 = relu(H_l Ws_l + mean_nbr(H_l) Wn_l + b_l)  l = 1..4, width 64
//   P     = H5 Wo + bo                               n x 2 positions
//
// Activations and weights are BF16, every product accumulates in FP32
// (v_mfma_f32_16x16x16_bf16, one wave per 16-row tile); biases, the loss and
// the Adam state are FP32. The loss is the stress of the predicted layout over
// given node pairs (i, j, d): sum ((|p_i - p_j| - d) / d)^2 / pairs, which
// does not depend on the layout's rotation or translation.
//
// Operand layout of the 16x16x16 BF16 MFMA (lane l of 64): A supplies row
// l % 16, k = 4 (l / 16) + 0..3; B supplies column l % 16, k = 4 (l / 16) +
// 0..3; the FP32 result holds rows 4 (l / 16) + 0..3 of column l % 16.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/d3fgpu.h"
#include "buffers.h"
#include "common.h"

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace d3f {
namespace {

constexpr int F_IN = 8;     // input features
constexpr int K_IN = 16;    // input features padded to one MFMA k-step
constexpr int HID = 64;     // width
constexpr int LAYERS = 4;   // message-passing layers
constexpr int OUT = 2;      // positions
constexpr int OUT_PAD = 16;

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float f) {  // round to nearest even
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Parameter layout (FP32 master copy, also the gradient and Adam layout):
// W0[K_IN][HID] (rows >= F_IN zero), b0[HID], per layer Ws[HID][HID],
// Wn[HID][HID], b[HID], then Wo[HID][OUT_PAD] (columns >= OUT zero), bo[OUT_PAD].
constexpr size_t OFF_W0 = 0;
constexpr size_t OFF_B0 = OFF_W0 + K_IN * HID;
constexpr size_t OFF_L = OFF_B0 + HID;
constexpr size_t L_STRIDE = 2 * HID * HID + HID;
constexpr size_t OFF_WO = OFF_L + LAYERS * L_STRIDE;
constexpr size_t OFF_BO = OFF_WO + HID * OUT_PAD;
constexpr size_t N_PARAMS = OFF_BO + OUT_PAD;

// ---------------------------------------------------------------- GEMM

// Y[n x N] = A[n x K] B[K x N] (+ A2 B2) + bias (+ add), for K in {16, 64},
// N in {16, 64}: A, A2 row-major BF16 with K columns; B given transposed
// (BT[N][K], BF16); add a BF16 [n x N] addend. Writes the FP32
// pre-activation (if z) and relu or identity as BF16 (if h) and FP32 (if
// yf), rows < `rows` only. One wave per 16 rows, rows padded to 16.
template <int K, int N>
__global__ __launch_bounds__(64) void k_gemm(const uint16_t* __restrict__ a, const uint16_t* __restrict__ bt,
                                             const uint16_t* __restrict__ a2, const uint16_t* __restrict__ bt2,
                                             const float* __restrict__ bias, const uint16_t* __restrict__ add,
                                             uint32_t rows, int relu, float* __restrict__ z,
                                             uint16_t* __restrict__ h, float* __restrict__ yf) {
  const uint32_t r0 = blockIdx.x * 16;
  const int lane = threadIdx.x, lr = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[N / 16];
#pragma unroll
  for (int t = 0; t < N / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int pass = 0; pass < (a2 ? 2 : 1); ++pass) {
    const uint16_t* A = pass ? a2 : a;
    const uint16_t* BT = pass ? bt2 : bt;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += 16) {
      const bf16x4 av = *(const bf16x4*)(A + (size_t)(r0 + lr) * K + k0 + lk);
#pragma unroll
      for (int t = 0; t < N / 16; ++t) {
        const bf16x4 bv = *(const bf16x4*)(BT + (size_t)(t * 16 + lr) * K + k0 + lk);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < N / 16; ++t) {
    const int col = t * 16 + lr;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t row = r0 + lk + j;
      if (row >= rows) continue;
      float v = acc[t][j] + (bias ? bias[col] : 0.f);
      if (add) v += from_bf16(add[(size_t)row * N + col]);
      if (z) z[(size_t)row * N + col] = v;
      const float y = relu && v < 0.f ? 0.f : v;
      if (h) h[(size_t)row * N + col] = to_bf16(y);
      if (yf) yf[(size_t)row * N + col] = y;
    }
  }
}

// Weight gradient G[K x N] = A^T[K x rows] D[rows x N] for K, N in {16, 64}:
// each block takes 64 rows, stages A and D tiles in LDS and leaves its
// partial product in part[block]; k_reduce adds the partials in block order.
template <int K, int N>
__global__ __launch_bounds__(64) void k_gemm_tn(const uint16_t* __restrict__ a, const uint16_t* __restrict__ d,
                                                uint32_t rows, float* __restrict__ part) {
  __shared__ uint16_t sa[64][K + 4];
  __shared__ uint16_t sd[64][N + 4];
  const uint32_t r0 = blockIdx.x * 64;
  const int lane = threadIdx.x, lr = lane & 15, lk = (lane >> 4) * 4;
  for (int q = lane; q < 64 * K; q += 64) {
    const int r = q / K, c = q % K;
    sa[r][c] = r0 + r < rows ? a[(size_t)(r0 + r) * K + c] : 0;
  }
  for (int q = lane; q < 64 * N; q += 64) {
    const int r = q / N, c = q % N;
    sd[r][c] = r0 + r < rows ? d[(size_t)(r0 + r) * N + c] : 0;
  }
  __syncthreads();
  float* out = part + (size_t)blockIdx.x * K * N;
#pragma unroll
  for (int mt = 0; mt < K / 16; ++mt) {
#pragma unroll
    for (int nt = 0; nt < N / 16; ++nt) {
      f32x4 acc = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k0 = 0; k0 < 64; k0 += 16) {  // k runs over the block's rows
        bf16x4 av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          av[j] = (short)sa[k0 + lk + j][mt * 16 + lr];
          bv[j] = (short)sd[k0 + lk + j][nt * 16 + lr];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) out[(size_t)(mt * 16 + lk + j) * N + nt * 16 + lr] = acc[j];
    }
  }
}

// g[k] = sum over blocks of part[block][k] in block order.
__global__ __launch_bounds__(256) void k_reduce(const float* __restrict__ part, uint32_t blocks, uint32_t size,
                                                float* __restrict__ g) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= size) return;
  float s = 0.f;
  for (uint32_t b = 0; b < blocks; ++b) s += part[(size_t)b * size + k];
  g[k] = s;
}

// Bias gradient: db[c] = sum over rows of dz[row][c] (FP32, row order).
// Column sums of dz (the bias gradients) in a fixed order: block y sums
// rows [y COLSUM_ROWS, (y + 1) COLSUM_ROWS) of each column in row order,
// then k_colsum_final adds the block sums in block order.
constexpr uint32_t COLSUM_ROWS = 1024;

__global__ __launch_bounds__(64) void k_colsum_part(const float* __restrict__ dz, uint32_t rows, int cols,
                                                    float* __restrict__ part) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= cols) return;
  const uint32_t r0 = blockIdx.y * COLSUM_ROWS;
  const uint32_t r1 = rows - r0 < COLSUM_ROWS ? rows : r0 + COLSUM_ROWS;
  float s = 0.f;
  for (uint32_t r = r0; r < r1; ++r) s += dz[(size_t)r * cols + c];
  part[(size_t)blockIdx.y * cols + c] = s;
}

__global__ __launch_bounds__(64) void k_colsum_final(const float* __restrict__ part, uint32_t parts, int cols,
                                                     float* __restrict__ db) {
  const int c = blockIdx.x * 64 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (uint32_t b = 0; b < parts; ++b) s += part[(size_t)b * cols + c];
  db[c] = s;
}

// ---------------------------------------------------------------- graph ops

// out[i] = mean of h over i's neighbours (CSR row i), BF16 in and out, FP32
// sum in row order; transpose = sum over neighbours j of h[j] / deg(j) (the
// adjoint of the mean on a symmetric graph). One wave per node, one lane per
// feature.
__global__ __launch_bounds__(64) void k_mean(const uint16_t* __restrict__ h, const uint32_t* __restrict__ row,
                                             const uint32_t* __restrict__ col, uint32_t n, int transpose,
                                             uint16_t* __restrict__ out) {
  const uint32_t i = blockIdx.x;
  const int f = threadIdx.x;
  const uint32_t a0 = row[i], a1 = row[i + 1];
  float s = 0.f;
  for (uint32_t a = a0; a < a1; ++a) {
    const uint32_t j = col[a];
    const float v = from_bf16(h[(size_t)j * HID + f]);
    s += transpose ? v / (float)(row[j + 1] - row[j]) : v;
  }
  if (!transpose && a1 > a0) s /= (float)(a1 - a0);
  out[(size_t)i * HID + f] = to_bf16(s);
}

// dz = dh * relu'(z) as BF16 (and FP32 for the bias gradient).
__global__ __launch_bounds__(256) void k_relu_back(const float* __restrict__ dh, const float* __restrict__ z,
                                                   uint32_t count, uint16_t* __restrict__ dz,
                                                   float* __restrict__ dzf) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count) return;
  const float v = z[k] > 0.f ? dh[k] : 0.f;
  dz[k] = to_bf16(v);
  dzf[k] = v;
}

// Stress over pairs: loss += ((r - d) / d)^2 / pairs; dP accumulates
// the gradient with atomics.
__global__ __launch_bounds__(256) void k_stress(const float* __restrict__ p, const uint32_t* __restrict__ pi,
                                                const uint32_t* __restrict__ pj, const float* __restrict__ pd,
                                                uint32_t pairs, float* __restrict__ dp, float* __restrict__ loss) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (k < pairs) {
    const uint32_t i = pi[k], j = pj[k];
    const float dx = p[i * OUT_PAD] - p[j * OUT_PAD], dy = p[i * OUT_PAD + 1] - p[j * OUT_PAD + 1];
    const float r = sqrtf(dx * dx + dy * dy + 1e-12f), d = pd[k];
    const float e = (r - d) / d;
    l = e * e / (float)pairs;
    const float g = 2.f * e / d / r / (float)pairs;
    atomicAdd(&dp[i * OUT_PAD], g * dx);
    atomicAdd(&dp[i * OUT_PAD + 1], g * dy);
    atomicAdd(&dp[j * OUT_PAD], -g * dx);
    atomicAdd(&dp[j * OUT_PAD + 1], -g * dy);
  }
  for (int d = 32; d > 0; d >>= 1) l += __shfl_xor(l, d, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ in, uint32_t count,
                                                 uint16_t* __restrict__ out) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < count) out[k] = to_bf16(in[k]);
}

// BF16 transposed copies of the weights (BT[N][K] for each W[K][N]) from the
// FP32 master parameters; also W itself (for dX = dZ W^T, whose B^T is W).
__global__ __launch_bounds__(256) void k_pack_weights(const float* __restrict__ w, uint16_t* __restrict__ wt,
                                                      uint16_t* __restrict__ wn) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= N_PARAMS) return;
  // find the matrix containing k
  size_t off = 0;
  int kk = 0, nn = 0;
  if (k < OFF_B0) { off = OFF_W0; kk = K_IN; nn = HID; }
  else if (k >= OFF_L && k < OFF_WO) {
    const size_t l = (k - OFF_L) / L_STRIDE, r = (k - OFF_L) % L_STRIDE;
    if (r >= 2 * HID * HID) return;
    off = OFF_L + l * L_STRIDE + (r >= HID * HID ? HID * HID : 0);
    kk = HID;
    nn = HID;
  } else if (k >= OFF_WO && k < OFF_BO) { off = OFF_WO; kk = HID; nn = OUT_PAD; }
  else return;
  const size_t e = k - off;
  const int r = (int)(e / nn), c = (int)(e % nn);
  wt[off + (size_t)c * kk + r] = to_bf16(w[k]);
  wn[k] = to_bf16(w[k]);
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, uint32_t count, float lr,
                                              float b1, float b2, float c1, float c2, float eps) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= count) return;
  const float gk = g[k];
  const float mk = b1 * m[k] + (1 - b1) * gk, vk = b2 * v[k] + (1 - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  w[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// Keeps the padding of the parameters at zero (rows >= F_IN of W0, columns
// >= OUT of Wo and bo).
__global__ __launch_bounds__(256) void k_zero_padding(float* __restrict__ w) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k < K_IN * HID && k / HID >= F_IN) w[OFF_W0 + k] = 0.f;
  if (k < HID * OUT_PAD && k % OUT_PAD >= OUT) w[OFF_WO + k] = 0.f;
  if (k < OUT_PAD && k >= OUT) w[OFF_BO + k] = 0.f;
}

// ---------------------------------------------------------------- host

// ---------------------------------------------------------------- inputs from the links
//
// The network's inputs from the links alone, on the GPU:
//  - the symmetric adjacency rows without self-links or repeats: keys
//    a << B | b of both directions of every link radix-sorted, the first of
//    each key kept, rows from the counts;
//  - 8 features per node: log(1 + degree), 1 / sqrt(1 + degree) and the 6
//    smoothest non-constant modes of the lazy random walk (I + D^-1 A) / 2.
//    They come from subspace iteration on the similar symmetric operator
//    L = (I + D^-1/2 A D^-1/2) / 2 from seeded Gaussian vectors (7 columns),
//    L's top vector sqrt(degree) projected out and the columns orthonormalised
//    (Cholesky QR, twice) every 5 steps, then Rayleigh-Ritz; a mode of the walk
//    is D^-1/2 times one of L's. Each feature column has mean 0, unit variance
//    and a nonnegative third moment. FP64 throughout; the 8 x 8 Cholesky and
//    the 7 x 7 symmetric eigenproblem run on the host.
constexpr int IN_MODES = 6;
constexpr int IN_COLS = IN_MODES + 1;  // iterated columns; column 7 of V holds sqrt(degree)
constexpr uint32_t IN_BLOCKS = 256;    // blocks of the reductions

struct Mat7 {
  double m[IN_COLS][IN_COLS];
  double c[IN_COLS];
};

__global__ __launch_bounds__(256) void k_in_keys(const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt,
                                                 uint32_t m, uint32_t n, int B, uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m) return;
  const uint64_t a = src[e], b = tgt[e];
  const uint64_t none = (uint64_t)n << B;  // a self-link: after every real row, dropped
  keys[2 * e] = a == b ? none : (a << B | b);
  keys[2 * e + 1] = a == b ? none : (b << B | a);
  vals[2 * e] = 2 * e;
  vals[2 * e + 1] = 2 * e + 1;
}

__global__ __launch_bounds__(256) void k_in_unique(const uint64_t* __restrict__ keys, uint32_t total, uint32_t n,
                                                   int B, uint32_t* __restrict__ flag, uint32_t* __restrict__ deg) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const uint64_t k = keys[i];
  const bool first = (i == 0 || keys[i - 1] != k) && (k >> B) < n;
  flag[i] = first ? 1u : 0u;
  if (first) atomicAdd(&deg[k >> B], 1u);
}

__global__ __launch_bounds__(256) void k_in_cols(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ flag,
                                                 const uint32_t* __restrict__ pos, uint32_t total, int B,
                                                 uint32_t* __restrict__ col) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= total || !flag[i]) return;
  col[pos[i]] = (uint32_t)(keys[i] & ((1ull << B) - 1));
}

// Link endpoint counts with multiplicity (d3's link strength 1 / min count).
__global__ __launch_bounds__(256) void k_in_count(const uint32_t* __restrict__ src, const uint32_t* __restrict__ tgt,
                                                  uint32_t m, uint32_t* __restrict__ cnt) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m) return;
  atomicAdd(&cnt[src[e]], 1u);
  atomicAdd(&cnt[tgt[e]], 1u);
}

__device__ inline uint64_t in_hash(uint64_t x) {  // splitmix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double in_uniform(uint64_t seed, uint64_t k) {  // (0, 1)
  return ((double)(in_hash(seed * 0x632BE59BD9B4E019ull + k) >> 11) + 0.5) * 0x1.0p-53;
}

// Degree features, D^-1/2, seeded Gaussian columns 0..6 (Box-Muller) and
// column 7 = sqrt(degree).
__global__ __launch_bounds__(256) void k_in_init(const uint32_t* __restrict__ row, uint32_t n, uint64_t seed,
                                                 double* __restrict__ v, double* __restrict__ dis,
                                                 float* __restrict__ f) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double deg = (double)(row[i + 1] - row[i]);
  f[8 * (size_t)i] = (float)log1p(deg);
  f[8 * (size_t)i + 1] = (float)(1.0 / sqrt(1.0 + deg));
  dis[i] = deg > 0 ? 1.0 / sqrt(deg) : 0.0;
  for (int k = 0; k < IN_COLS; ++k) {
    const uint64_t c = ((uint64_t)i * IN_COLS + k) * 2;
    const double u1 = in_uniform(seed, c), u2 = in_uniform(seed, c + 1);
    v[8 * (size_t)i + k] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
  }
  v[8 * (size_t)i + 7] = sqrt(deg);
}

// y = L v on columns 0..6 (column 7 copied: L keeps sqrt(degree)). A row
// is summed by a group of IN_GROUP lanes (neighbour a by lane a mod
// IN_GROUP, then a fixed-order shuffle reduction), so hub rows of
// thousands of neighbours do not hold one lane for the whole pass.
constexpr int IN_GROUP = 16;
__global__ __launch_bounds__(256) void k_in_walk(const uint32_t* __restrict__ row, const uint32_t* __restrict__ col,
                                                 const double* __restrict__ dis, const double* __restrict__ v,
                                                 double* __restrict__ y, uint32_t n) {
  const uint32_t i = (blockIdx.x * 256 + threadIdx.x) / IN_GROUP;
  const int g = threadIdx.x % IN_GROUP;
  if (i >= n) return;  // whole groups leave together
  double s[IN_COLS];
#pragma unroll
  for (int k = 0; k < IN_COLS; ++k) s[k] = 0;
  for (uint32_t a = row[i] + g; a < row[i + 1]; a += IN_GROUP) {
    const uint32_t j = col[a];
    const double w = dis[j];
    const double* vj = v + 8 * (size_t)j;
#pragma unroll
    for (int k = 0; k < IN_COLS; ++k) s[k] += w * vj[k];
  }
#pragma unroll
  for (int k = 0; k < IN_COLS; ++k)
    for (int d = IN_GROUP / 2; d > 0; d >>= 1) s[k] += __shfl_xor(s[k], d, IN_GROUP);
  if (g == 0) {
    const double di = dis[i];
#pragma unroll
    for (int k = 0; k < IN_COLS; ++k) y[8 * (size_t)i + k] = 0.5 * (v[8 * (size_t)i + k] + di * s[k]);
    y[8 * (size_t)i + 7] = v[8 * (size_t)i + 7];
  }
}

// Per-block partial sums of V^T W (8 x 8; W = V when null) into part[block][64].
__global__ __launch_bounds__(256) void k_in_gram(const double* __restrict__ v, const double* __restrict__ w, uint32_t n,
                                                 double* __restrict__ part) {
  __shared__ double red[4][64];
  double acc[64];
#pragma unroll
  for (int k = 0; k < 64; ++k) acc[k] = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double a[8], b[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) a[k] = v[8 * (size_t)i + k];
#pragma unroll
    for (int k = 0; k < 8; ++k) b[k] = w ? w[8 * (size_t)i + k] : a[k];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[r * 8 + c] += a[r] * b[c];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 64; ++k) {
    double x = acc[k];
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    if (lane == 0) red[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < 64)
    part[(size_t)blockIdx.x * 64 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// out[k] = sum over blocks of part[block][k]: one block per k, the
// partials added by a fixed tree in LDS.
__global__ __launch_bounds__(256) void k_in_sum(const double* __restrict__ part, uint32_t blocks, uint32_t width,
                                                double* __restrict__ out) {
  __shared__ double red[256];
  const uint32_t k = blockIdx.x;
  double s = 0;
  for (uint32_t b = threadIdx.x; b < blocks; b += 256) s += part[(size_t)b * width + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int d = 128; d > 0; d >>= 1) {
    if ((int)threadIdx.x < d) red[threadIdx.x] += red[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

// V[0..6] <- (V[0..6] - V[7] c^T) M, column 7 kept.
__global__ __launch_bounds__(256) void k_in_transform(double* __restrict__ v, uint32_t n, Mat7 t) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double* vi = v + 8 * (size_t)i;
  double a[IN_COLS];
#pragma unroll
  for (int k = 0; k < IN_COLS; ++k) a[k] = vi[k] - vi[7] * t.c[k];
#pragma unroll
  for (int k = 0; k < IN_COLS; ++k) {
    double s = 0;
#pragma unroll
    for (int r = 0; r < IN_COLS; ++r) s += a[r] * t.m[r][k];
    vi[k] = s;
  }
}

// u = D^-1/2 V Q (the 6 leading Ritz vectors, Q in t.m columns 0..5) and
// per-block sums of u, u^2, u^3 per column into part[block][18].
__global__ __launch_bounds__(256) void k_in_ritz(const double* __restrict__ v, const double* __restrict__ dis,
                                                 uint32_t n, Mat7 t, double* __restrict__ u, double* __restrict__ part) {
  __shared__ double red[4][3 * IN_MODES];
  double acc[3 * IN_MODES];
#pragma unroll
  for (int k = 0; k < 3 * IN_MODES; ++k) acc[k] = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const double* vi = v + 8 * (size_t)i;
#pragma unroll
    for (int k = 0; k < IN_MODES; ++k) {
      double s = 0;
#pragma unroll
      for (int r = 0; r < IN_COLS; ++r) s += vi[r] * t.m[r][k];
      s *= dis[i];
      u[8 * (size_t)i + k] = s;
      acc[k] += s;
      acc[IN_MODES + k] += s * s;
      acc[2 * IN_MODES + k] += s * s * s;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 3 * IN_MODES; ++k) {
    double x = acc[k];
    for (int d = 32; d > 0; d >>= 1) x += __shfl_xor(x, d, 64);
    if (lane == 0) red[wave][k] = x;
  }
  __syncthreads();
  if (threadIdx.x < 3 * IN_MODES)
    part[(size_t)blockIdx.x * 3 * IN_MODES + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Feature columns 2..7 = sign (u - mean) / sd (FP32), then all 8 as BF16
// into the network's input rows (K_IN columns, the rest zero).
__global__ __launch_bounds__(256) void k_in_write(const double* __restrict__ u, uint32_t n, Mat7 t, int modes_on,
                                                  float* __restrict__ f, uint16_t* __restrict__ x) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float* fi = f + 8 * (size_t)i;
  for (int k = 0; k < IN_MODES; ++k)  // t.m[0]: mean, t.m[1]: scale (sign / sd), per mode
    fi[2 + k] = modes_on ? (float)((u[8 * (size_t)i + k] - t.m[0][k]) * t.m[1][k]) : 0.f;
  for (int k = 0; k < F_IN; ++k) x[(size_t)i * K_IN + k] = to_bf16(fi[k]);
}

// Prediction: per-block sums of the positions (x, y).
__global__ __launch_bounds__(256) void k_in_psum(const float* __restrict__ p, uint32_t n, double* __restrict__ part) {
  __shared__ double red[4][2];
  double sx = 0, sy = 0;
  for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    sx += p[(size_t)i * OUT_PAD];
    sy += p[(size_t)i * OUT_PAD + 1];
  }
  for (int d = 32; d > 0; d >>= 1) {
    sx += __shfl_xor(sx, d, 64);
    sy += __shfl_xor(sy, d, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = sx;
    red[wave][1] = sy;
  }
  __syncthreads();
  if (threadIdx.x < 2)
    part[(size_t)blockIdx.x * 2 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// Per-block sums over links of k r^2, k r and r (r the predicted link
// length, k = 1 / min(count[source], count[target])).
__global__ __launch_bounds__(256) void k_in_link_sums(const uint32_t* __restrict__ src,
                                                      const uint32_t* __restrict__ tgt, uint32_t m,
                                                      const uint32_t* __restrict__ cnt, const float* __restrict__ p,
                                                      double* __restrict__ part) {
  __shared__ double red[4][3];
  double a = 0, b = 0, c = 0;
  for (uint32_t e = blockIdx.x * 256 + threadIdx.x; e < m; e += gridDim.x * 256) {
    const uint32_t s_ = src[e], t_ = tgt[e];
    const double dx = (double)p[(size_t)s_ * OUT_PAD] - (double)p[(size_t)t_ * OUT_PAD];
    const double dy = (double)p[(size_t)s_ * OUT_PAD + 1] - (double)p[(size_t)t_ * OUT_PAD + 1];
    const double r = sqrt(dx * dx + dy * dy);
    const uint32_t cs = cnt[s_], ct = cnt[t_];
    const double k = 1.0 / (double)(cs < ct ? cs : ct);
    a += k * r * r;
    b += k * r;
    c += r;
  }
  for (int d = 32; d > 0; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[wave][0] = a;
    red[wave][1] = b;
    red[wave][2] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    part[(size_t)blockIdx.x * 3 + threadIdx.x] =
        red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
}

// positions (FP64, n x 2) = (p - mean) s + uniform(-spread, spread) from the seed.
__global__ __launch_bounds__(256) void k_in_place(const float* __restrict__ p, uint32_t n, double mx, double my,
                                                  double s, double spread, uint64_t seed, double* __restrict__ out) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const double jx = spread > 0 ? (2 * in_uniform(seed, 2 * (uint64_t)i) - 1) * spread : 0.0;
  const double jy = spread > 0 ? (2 * in_uniform(seed, 2 * (uint64_t)i + 1) - 1) * spread : 0.0;
  out[2 * (size_t)i] = ((double)p[(size_t)i * OUT_PAD] - mx) * s + jx;
  out[2 * (size_t)i + 1] = ((double)p[(size_t)i * OUT_PAD + 1] - my) * s + jy;
}

// Cyclic Jacobi eigen-decomposition of a symmetric 7 x 7 matrix: values in
// w, vectors in the columns of q.
void jacobi7(double a[IN_COLS][IN_COLS], double w[IN_COLS], double q[IN_COLS][IN_COLS]) {
  for (int r = 0; r < IN_COLS; ++r)
    for (int c = 0; c < IN_COLS; ++c) q[r][c] = r == c ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0, diag = 0;
    for (int r = 0; r < IN_COLS; ++r)
      for (int c = 0; c < IN_COLS; ++c) (r == c ? diag : off) += a[r][c] * a[r][c];
    if (off <= 1e-30 * diag) break;
    for (int p_ = 0; p_ < IN_COLS; ++p_)
      for (int q_ = p_ + 1; q_ < IN_COLS; ++q_) {
        if (a[p_][q_] == 0) continue;
        const double theta = (a[q_][q_] - a[p_][p_]) / (2 * a[p_][q_]);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
        const double c = 1 / std::sqrt(t * t + 1), s_ = t * c;
        for (int k = 0; k < IN_COLS; ++k) {  // A <- J^T A J, Q <- Q J
          const double akp = a[k][p_], akq = a[k][q_];
          a[k][p_] = c * akp - s_ * akq;
          a[k][q_] = s_ * akp + c * akq;
        }
        for (int k = 0; k < IN_COLS; ++k) {
          const double apk = a[p_][k], aqk = a[q_][k];
          a[p_][k] = c * apk - s_ * aqk;
          a[q_][k] = s_ * apk + c * aqk;
        }
        for (int k = 0; k < IN_COLS; ++k) {
          const double qkp = q[k][p_], qkq = q[k][q_];
          q[k][p_] = c * qkp - s_ * qkq;
          q[k][q_] = s_ * qkp + c * qkq;
        }
      }
  }
  for (int k = 0; k < IN_COLS; ++k) w[k] = a[k][k];
}

struct Net {
  uint32_t n = 0, rows = 0, nnz = 0;  // rows: n padded to 64
  Pinned<float> hparams, hgrad;
  Device<float> w, g, m, v, part, cpart, z[LAYERS + 2], dh, dzf, p, dp, loss;
  Device<uint16_t> wt, wn, x, hh[LAYERS + 2], agg[LAYERS], dz, dhb, tmp;
  Device<uint32_t> row, col, pi, pj;
  Device<float> pd;
  // inputs from the links (set_links) and the prediction (predict)
  uint32_t links = 0;
  Device<uint32_t> lsrc, ltgt, lcnt, kvals, kvals_t, flag, pos, deg, hist, glob, partials, counts;
  Device<uint64_t> keys, keys_t;
  Device<double> fv, fy, fu, dis, fpart, fred, place;
  Device<float> ff;
  uint32_t pairs = 0;
  int step = 0;
  hipStream_t stream = nullptr;

  ~Net() {
    if (stream) (void)hipStreamSynchronize(stream);
    hparams.release();
    hgrad.release();
    for (auto* a : {&w, &g, &m, &v, &part, &dh, &dzf, &p, &dp, &loss, &pd}) a->release();
    for (auto& a : z) a.release();
    for (auto* a : {&wt, &wn, &x, &dz, &dhb, &tmp}) a->release();
    for (auto& a : hh) a.release();
    for (auto& a : agg) a.release();
    for (auto* a : {&row, &col, &pi, &pj}) a->release();
    for (auto* a : {&lsrc, &ltgt, &lcnt, &kvals, &kvals_t, &flag, &pos, &deg, &hist, &glob, &partials, &counts})
      a->release();
    for (auto* a : {&keys, &keys_t}) a->release();
    for (auto* a : {&fv, &fy, &fu, &dis, &fpart, &fred, &place}) a->release();
    ff.release();
    if (stream) (void)hipStreamDestroy(stream);
  }

  hipError_t init(uint64_t seed) {
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hparams.reserve(N_PARAMS, 0));
    HIPCHECK(hgrad.reserve(N_PARAMS, 0));
    for (auto* a : {&w, &g, &m, &v}) HIPCHECK(a->reserve(N_PARAMS));
    HIPCHECK(wt.reserve(N_PARAMS));
    HIPCHECK(wn.reserve(N_PARAMS));
    HIPCHECK(loss.reserve(1));
    // Glorot-uniform weights from a splitmix64 stream, zero biases.
    std::memset(hparams.p, 0, N_PARAMS * sizeof(float));
    uint64_t s = seed;
    auto uni = [&]() {
      uint64_t h = (s += 0x9E3779B97F4A7C15ull);
      h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
      h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
      h ^= h >> 31;
      return (double)(h >> 11) * 0x1.0p-53;
    };
    auto fill = [&](size_t off, int k, int nn, int kreal, int nreal) {
      const double a = std::sqrt(6.0 / (kreal + nreal));
      for (int r = 0; r < k; ++r)
        for (int c = 0; c < nn; ++c)
          hparams.p[off + (size_t)r * nn + c] = (r < kreal && c < nreal) ? (float)((2 * uni() - 1) * a) : 0.f;
    };
    fill(OFF_W0, K_IN, HID, F_IN, HID);
    for (int l = 0; l < LAYERS; ++l) {
      fill(OFF_L + l * L_STRIDE, HID, HID, HID, HID);
      fill(OFF_L + l * L_STRIDE + HID * HID, HID, HID, HID, HID);
    }
    fill(OFF_WO, HID, OUT_PAD, HID, OUT);
    return upload_params();
  }

  hipError_t upload_params() {
    HIPCHECK(hipMemcpyAsync(w.p, hparams.p, N_PARAMS * sizeof(float), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemsetAsync(m.p, 0, N_PARAMS * sizeof(float), stream));
    HIPCHECK(hipMemsetAsync(v.p, 0, N_PARAMS * sizeof(float), stream));
    step = 0;
    return pack();
  }

  hipError_t pack() {
    HIPCHECK(hipMemsetAsync(wt.p, 0, N_PARAMS * sizeof(uint16_t), stream));
    hipLaunchKernelGGL(k_pack_weights, dim3(ceil_div((uint32_t)N_PARAMS, 256)), dim3(256), 0, stream, w.p, wt.p,
                       wn.p);
    return hipGetLastError();
  }

  hipError_t set_graph(uint32_t nn, const uint32_t* hrow, const uint32_t* hcol, const float* feat) {
    HIPCHECK(hipStreamSynchronize(stream));
    n = nn;
    rows = ceil_div(n > 0 ? n : 1, 64) * 64;
    nnz = hrow[n];
    HIPCHECK(row.reserve(n + 1));
    HIPCHECK(col.reserve(nnz > 0 ? nnz : 1));
    HIPCHECK(hipMemcpy(row.p, hrow, (n + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (nnz) HIPCHECK(hipMemcpy(col.p, hcol, nnz * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHECK(alloc_rows());
    // features as BF16, padded with zeros to K_IN columns and to `rows` rows
    std::vector<uint16_t> xb((size_t)rows * K_IN, 0);
    for (uint32_t i = 0; i < n; ++i)
      for (int f = 0; f < F_IN; ++f) {
        uint32_t u;
        std::memcpy(&u, &feat[(size_t)i * F_IN + f], 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) u |= 0x400000u;
        else u += 0x7fffu + ((u >> 16) & 1u);
        xb[(size_t)i * K_IN + f] = (uint16_t)(u >> 16);
      }
    HIPCHECK(hipMemcpy(x.p, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    return hipSuccess;
  }

  // The per-node buffers for n rows (padded to `rows`), cleared; x zeroed.
  hipError_t alloc_rows() {
    const size_t r = rows;
    HIPCHECK(x.reserve(r * K_IN));
    HIPCHECK(hipMemsetAsync(x.p, 0, r * K_IN * sizeof(uint16_t), stream));
    for (auto& a : hh) HIPCHECK(a.reserve(r * HID));
    for (auto& a : agg) HIPCHECK(a.reserve(r * HID));
    for (auto& a : z) HIPCHECK(a.reserve(r * HID));
    for (auto* a : {&dz, &dhb, &tmp}) HIPCHECK(a->reserve(r * HID));
    HIPCHECK(dh.reserve(r * HID));
    HIPCHECK(dzf.reserve(r * HID));
    HIPCHECK(p.reserve(r * OUT_PAD));
    HIPCHECK(dp.reserve(r * OUT_PAD));
    HIPCHECK(part.reserve((size_t)(r / 64) * HID * HID));
    HIPCHECK(cpart.reserve((size_t)ceil_div(r, COLSUM_ROWS) * HID));
    for (auto& a : hh) HIPCHECK(hipMemsetAsync(a.p, 0, r * HID * sizeof(uint16_t), stream));
    for (auto& a : agg) HIPCHECK(hipMemsetAsync(a.p, 0, r * HID * sizeof(uint16_t), stream));
    for (auto& a : z) HIPCHECK(hipMemsetAsync(a.p, 0, r * HID * sizeof(float), stream));
    for (auto* a : {&dz, &dhb, &tmp}) HIPCHECK(hipMemsetAsync(a->p, 0, r * HID * sizeof(uint16_t), stream));
    HIPCHECK(hipMemsetAsync(dh.p, 0, r * HID * sizeof(float), stream));
    HIPCHECK(hipMemsetAsync(dzf.p, 0, r * HID * sizeof(float), stream));
    HIPCHECK(hipMemsetAsync(p.p, 0, r * OUT_PAD * sizeof(float), stream));
    HIPCHECK(hipMemsetAsync(dp.p, 0, r * OUT_PAD * sizeof(float), stream));
    return hipStreamSynchronize(stream);
  }

  // ---- inputs from the links (see the kernels above)

  // Sum of `width` values over the IN_BLOCKS partials in fpart, to the host.
  hipError_t reduce_host(uint32_t width, double* out) {
    hipLaunchKernelGGL(k_in_sum, dim3(width), dim3(256), 0, stream, fpart.p, IN_BLOCKS, width, fred.p);
    HIPCHECK(hipMemcpyAsync(out, fred.p, width * sizeof(double), hipMemcpyDeviceToHost, stream));
    return hipStreamSynchronize(stream);
  }

  // G = V^T W (8 x 8) on the host; W = V when null.
  hipError_t gram(const double* w, double g[8][8]) {
    hipLaunchKernelGGL(k_in_gram, dim3(IN_BLOCKS), dim3(256), 0, stream, fv.p, w, n, fpart.p);
    return reduce_host(64, &g[0][0]);
  }

  // Columns 0..6 of V: sqrt(degree) (column 7) projected out, then
  // orthonormal (Cholesky QR). False when the columns are (numerically)
  // dependent.
  hipError_t orthonormalise(bool* ok) {
    double g[8][8];
    HIPCHECK(gram(nullptr, g));
    Mat7 t{};
    const double g77 = g[7][7];
    for (int k = 0; k < IN_COLS; ++k) t.c[k] = g77 > 0 ? g[k][7] / g77 : 0.0;
    double a[IN_COLS][IN_COLS];  // Gram matrix of the projected columns
    for (int r = 0; r < IN_COLS; ++r)
      for (int c = 0; c < IN_COLS; ++c) a[r][c] = g[r][c] - (g77 > 0 ? g[r][7] * g[c][7] / g77 : 0.0);
    double l[IN_COLS][IN_COLS] = {};  // a = l l^T
    for (int j = 0; j < IN_COLS; ++j) {
      double d = a[j][j];
      for (int k = 0; k < j; ++k) d -= l[j][k] * l[j][k];
      if (!(d > 1e-300) || !std::isfinite(d)) {
        *ok = false;
        return hipSuccess;
      }
      l[j][j] = std::sqrt(d);
      for (int r = j + 1; r < IN_COLS; ++r) {
        double v = a[r][j];
        for (int k = 0; k < j; ++k) v -= l[r][k] * l[j][k];
        l[r][j] = v / l[j][j];
      }
    }
    // M = (l^T)^-1: upper triangular, V M has orthonormal columns
    for (int c = 0; c < IN_COLS; ++c)
      for (int r = IN_COLS - 1; r >= 0; --r) {
        double v = r == c ? 1.0 : 0.0;
        for (int k = r + 1; k < IN_COLS; ++k) v -= l[k][r] * t.m[k][c];
        t.m[r][c] = v / l[r][r];
      }
    hipLaunchKernelGGL(k_in_transform, dim3(ceil_div(n, 256u)), dim3(256), 0, stream, fv.p, n, t);
    HIPCHECK(hipGetLastError());
    *ok = true;
    return hipSuccess;
  }

  // The graph from its links (endpoints < nn, checked by the caller): rows,
  // features into x (and to feats, n x 8 FP32, when given), link counts.
  hipError_t set_links(uint32_t nn, uint32_t mm, const uint32_t* hsrc, const uint32_t* htgt, uint64_t seed,
                       float* feats) {
    HIPCHECK(hipStreamSynchronize(stream));
    n = nn;
    links = mm;
    rows = ceil_div(n > 0 ? n : 1, 64) * 64;
    const size_t cm = links > 0 ? links : 1, c2 = 2 * cm, cn = n > 0 ? n : 1;
    HIPCHECK(lsrc.reserve(cm));
    HIPCHECK(ltgt.reserve(cm));
    if (links) {
      HIPCHECK(hipMemcpyAsync(lsrc.p, hsrc, links * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipMemcpyAsync(ltgt.p, htgt, links * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    }
    // rows: sort the keys of both directions, keep the first of each
    int bits = 1;
    while (bits < 32 && ((uint64_t)1 << bits) <= n) ++bits;  // n itself fits (the self-link row)
    const int B = 8 * ((bits + 7) / 8);
    const int digits = (B + bits + 7) / 8;
    for (auto* a : {&kvals, &kvals_t, &flag, &pos}) HIPCHECK(a->reserve(c2));
    for (auto* a : {&keys, &keys_t}) HIPCHECK(a->reserve(c2));
    HIPCHECK(hist.reserve(sort_hist_words((uint32_t)c2)));
    HIPCHECK(glob.reserve(8 * 256));
    HIPCHECK(partials.reserve(ceil_div((uint32_t)std::max(c2, cn), 4096u) + 2));
    HIPCHECK(counts.reserve(2));
    HIPCHECK(deg.reserve(cn));
    HIPCHECK(lcnt.reserve(cn));
    HIPCHECK(row.reserve(n + 1));
    HIPCHECK(col.reserve(c2));
    const uint32_t total = 2 * links, hc[2] = {total, n};
    HIPCHECK(hipMemcpyAsync(counts.p, hc, sizeof(hc), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemsetAsync(deg.p, 0, cn * sizeof(uint32_t), stream));
    HIPCHECK(hipMemsetAsync(lcnt.p, 0, cn * sizeof(uint32_t), stream));
    if (links) {
      hipLaunchKernelGGL(k_in_keys, dim3(ceil_div(links, 256u)), dim3(256), 0, stream, lsrc.p, ltgt.p, links, n, B, keys.p,
                         kvals.p);
      HIPCHECK(sort_digits(hist.p, glob.p, keys.p, kvals.p, keys_t.p, kvals_t.p, total, counts.p, 0, digits, stream));
      hipLaunchKernelGGL(k_in_unique, dim3(ceil_div(total, 256u)), dim3(256), 0, stream, keys.p, total, n, B, flag.p,
                         deg.p);
      HIPCHECK(scan_u32(flag.p, pos.p, total, counts.p, partials.p, nullptr, stream));
      hipLaunchKernelGGL(k_in_cols, dim3(ceil_div(total, 256u)), dim3(256), 0, stream, keys.p, flag.p, pos.p, total,
                         B, col.p);
      hipLaunchKernelGGL(k_in_count, dim3(ceil_div(links, 256u)), dim3(256), 0, stream, lsrc.p, ltgt.p, links, lcnt.p);
    }
    if (n) HIPCHECK(scan_u32(deg.p, row.p, n, counts.p + 1, partials.p, row.p + n, stream));
    else HIPCHECK(hipMemsetAsync(row.p, 0, sizeof(uint32_t), stream));
    HIPCHECK(hipMemcpyAsync(&nnz, row.p + n, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(alloc_rows());
    // features
    const size_t r8 = (size_t)cn * 8;
    for (auto* a : {&fv, &fy, &fu}) HIPCHECK(a->reserve(r8));
    HIPCHECK(dis.reserve(cn));
    HIPCHECK(ff.reserve(r8));
    HIPCHECK(fpart.reserve((size_t)IN_BLOCKS * 64));
    HIPCHECK(fred.reserve(64));
    const uint32_t nb = ceil_div(n > 0 ? n : 1, 256u);
    const uint32_t wb = (uint32_t)(((uint64_t)(n > 0 ? n : 1) * IN_GROUP + 255) / 256);
    hipLaunchKernelGGL(k_in_init, dim3(nb), dim3(256), 0, stream, row.p, n, seed, fv.p, dis.p, ff.p);
    bool ok = n > (uint32_t)IN_COLS;  // warmstart.py's features: modes only above 7 nodes
    Mat7 t{};
    constexpr int ITERATIONS = 40;
    for (int it = 0; it < ITERATIONS && ok; ++it) {
      hipLaunchKernelGGL(k_in_walk, dim3(wb), dim3(256), 0, stream, row.p, col.p, dis.p, fv.p, fy.p, n);
      std::swap(fv.p, fy.p);
      if (it % 5 == 4 || it == ITERATIONS - 1)
        for (int pass = 0; pass < 2 && ok; ++pass) HIPCHECK(orthonormalise(&ok));
    }
    if (ok) {  // Rayleigh-Ritz: H = V^T L V, smoothest modes first
      hipLaunchKernelGGL(k_in_walk, dim3(wb), dim3(256), 0, stream, row.p, col.p, dis.p, fv.p, fy.p, n);
      double g[8][8], h[IN_COLS][IN_COLS], w[IN_COLS], q[IN_COLS][IN_COLS];
      HIPCHECK(gram(fy.p, g));
      for (int r = 0; r < IN_COLS; ++r)
        for (int c = 0; c < IN_COLS; ++c) h[r][c] = 0.5 * (g[r][c] + g[c][r]);
      jacobi7(h, w, q);
      int order[IN_COLS];
      for (int k = 0; k < IN_COLS; ++k) order[k] = k;
      std::sort(order, order + IN_COLS, [&](int a, int b) { return w[a] > w[b]; });
      for (int r = 0; r < IN_COLS; ++r)
        for (int k = 0; k < IN_COLS; ++k) t.m[r][k] = k < IN_MODES ? q[r][order[k]] : 0.0;
      hipLaunchKernelGGL(k_in_ritz, dim3(IN_BLOCKS), dim3(256), 0, stream, fv.p, dis.p, n, t, fu.p, fpart.p);
      double mom[3 * IN_MODES];
      HIPCHECK(reduce_host(3 * IN_MODES, mom));
      t = Mat7{};
      for (int k = 0; k < IN_MODES; ++k) {
        const double mean = mom[k] / n, e2 = mom[IN_MODES + k] / n, e3 = mom[2 * IN_MODES + k] / n;
        const double var = e2 - mean * mean;
        const double sd = var > 0 ? std::sqrt(var) : 0.0;
        const double third = e3 - 3 * mean * e2 + 2 * mean * mean * mean;  // n E[(u - mean)^3]
        t.m[0][k] = mean;
        t.m[1][k] = (sd > 0 ? 1 / sd : 1.0) * (third < 0 ? -1.0 : 1.0);
      }
    }
    hipLaunchKernelGGL(k_in_write, dim3(nb), dim3(256), 0, stream, fu.p, n, t, ok ? 1 : 0, ff.p, x.p);
    HIPCHECK(hipGetLastError());
    if (feats && n)
      HIPCHECK(hipMemcpyAsync(feats, ff.p, (size_t)n * 8 * sizeof(float), hipMemcpyDeviceToHost, stream));
    return hipStreamSynchronize(stream);
  }

  // Predicted layout (n x 2, FP64) of the graph set by set_links: the
  // network's output centred, scaled to the size that minimises d3's energy
  // over uniform scalings (links k (s r - 30)^2 / 2 against 30 log(1 / s r)
  // over all pairs: s = (b + sqrt(b^2 + 4 a c)) / 2 a with a = sum k r^2,
  // b = 30 sum k r, c = 30 n (n - 1) / 2), plus a seeded uniform jitter of
  // +-jitter times the mean scaled link length.
  hipError_t predict(double jitter, uint64_t seed, double* out) {
    HIPCHECK(forward());
    double ps[2] = {0, 0}, ls[3] = {0, 0, 0};
    hipLaunchKernelGGL(k_in_psum, dim3(IN_BLOCKS), dim3(256), 0, stream, p.p, n, fpart.p);
    HIPCHECK(reduce_host(2, ps));
    if (links) {
      hipLaunchKernelGGL(k_in_link_sums, dim3(IN_BLOCKS), dim3(256), 0, stream, lsrc.p, ltgt.p, links, lcnt.p, p.p,
                         fpart.p);
      HIPCHECK(reduce_host(3, ls));
    }
    const double mx = n ? ps[0] / n : 0.0, my = n ? ps[1] / n : 0.0;
    const double a = ls[0], b = 30.0 * ls[1], c = 30.0 * (double)n * ((double)n - 1) / 2;
    const double s_ = a > 0 ? (b + std::sqrt(b * b + 4 * a * c)) / (2 * a) : 1.0;
    const double spread = links ? jitter * s_ * ls[2] / links : jitter;
    HIPCHECK(place.reserve((size_t)(n > 0 ? n : 1) * 2));
    hipLaunchKernelGGL(k_in_place, dim3(ceil_div(n > 0 ? n : 1, 256u)), dim3(256), 0, stream, p.p, n, mx, my, s_,
                       spread, seed, place.p);
    if (n) HIPCHECK(hipMemcpyAsync(out, place.p, (size_t)n * 2 * sizeof(double), hipMemcpyDeviceToHost, stream));
    return hipStreamSynchronize(stream);
  }

  // Forward; hh[l] holds H_l (l = 1..5, BF16), z[l] the pre-activations,
  // p the positions (FP32, OUT_PAD columns).
  hipError_t forward() {
    const uint32_t tiles = rows / 16;
    hipLaunchKernelGGL((k_gemm<K_IN, HID>), dim3(tiles), dim3(64), 0, stream, x.p, wt.p + OFF_W0, nullptr, nullptr,
                       w.p + OFF_B0, nullptr, n, 1, z[0].p, hh[1].p, nullptr);
    for (int l = 0; l < LAYERS; ++l) {
      const size_t o = OFF_L + l * L_STRIDE;
      hipLaunchKernelGGL(k_mean, dim3(n), dim3(64), 0, stream, hh[l + 1].p, row.p, col.p, n, 0, agg[l].p);
      hipLaunchKernelGGL((k_gemm<HID, HID>), dim3(tiles), dim3(64), 0, stream, hh[l + 1].p, wt.p + o, agg[l].p,
                         wt.p + o + HID * HID, w.p + o + 2 * HID * HID, nullptr, n, 1, z[l + 1].p, hh[l + 2].p,
                         nullptr);
    }
    hipLaunchKernelGGL((k_gemm<HID, OUT_PAD>), dim3(tiles), dim3(64), 0, stream, hh[LAYERS + 1].p, wt.p + OFF_WO,
                       nullptr, nullptr, w.p + OFF_BO, nullptr, n, 0, nullptr, nullptr, p.p);
    return hipGetLastError();
  }

  // Weight gradient of one matrix: g[off] = A^T D.
  template <int K, int N>
  hipError_t wgrad(const uint16_t* a, const uint16_t* d, size_t off) {
    const uint32_t blocks = rows / 64;
    hipLaunchKernelGGL((k_gemm_tn<K, N>), dim3(blocks), dim3(64), 0, stream, a, d, n, part.p);
    hipLaunchKernelGGL(k_reduce, dim3(ceil_div(K * N, 256)), dim3(256), 0, stream, part.p, blocks, (uint32_t)(K * N),
                       g.p + off);
    return hipGetLastError();
  }

  // db[c] = sum over the n node rows of m[r][c] (m has `cols` <= HID columns).
  void colsum(const float* m, int cols, float* db) {
    const uint32_t parts = ceil_div(n > 0 ? n : 1, COLSUM_ROWS);
    hipLaunchKernelGGL(k_colsum_part, dim3(ceil_div(cols, 64), parts), dim3(64), 0, stream, m, n, cols, cpart.p);
    hipLaunchKernelGGL(k_colsum_final, dim3(ceil_div(cols, 64)), dim3(64), 0, stream, cpart.p, parts, cols, db);
  }

  // Backward from dp (FP32, OUT_PAD columns) into g.
  hipError_t backward() {
    const uint32_t tiles = rows / 16, cnt = rows * HID;
    HIPCHECK(hipMemsetAsync(g.p, 0, N_PARAMS * sizeof(float), stream));
    // output head: dWo = H5^T dP, dbo, dH5 = dP Wo^T (B^T of that product is Wo)
    hipLaunchKernelGGL(k_to_bf16, dim3(ceil_div(rows * OUT_PAD, 256)), dim3(256), 0, stream, dp.p, rows * OUT_PAD,
                       tmp.p);
    HIPCHECK((wgrad<HID, OUT_PAD>(hh[LAYERS + 1].p, tmp.p, OFF_WO)));
    colsum(dp.p, OUT_PAD, g.p + OFF_BO);
    hipLaunchKernelGGL((k_gemm<OUT_PAD, HID>), dim3(tiles), dim3(64), 0, stream, tmp.p, wn.p + OFF_WO, nullptr, nullptr,
                       nullptr, nullptr, n, 0, nullptr, nullptr, dh.p);
    for (int l = LAYERS - 1; l >= 0; --l) {
      const size_t o = OFF_L + l * L_STRIDE;
      hipLaunchKernelGGL(k_relu_back, dim3(ceil_div(cnt, 256)), dim3(256), 0, stream, dh.p, z[l + 1].p, cnt, dz.p,
                         dzf.p);
      HIPCHECK((wgrad<HID, HID>(hh[l + 1].p, dz.p, o)));
      HIPCHECK((wgrad<HID, HID>(agg[l].p, dz.p, o + HID * HID)));
      colsum(dzf.p, HID, g.p + o + 2 * HID * HID);
      // dH_l = dZ Ws^T + mean^T(dZ Wn^T)
      hipLaunchKernelGGL((k_gemm<HID, HID>), dim3(tiles), dim3(64), 0, stream, dz.p, wn.p + o + HID * HID, nullptr,
                         nullptr, nullptr, nullptr, n, 0, nullptr, dhb.p, nullptr);
      hipLaunchKernelGGL(k_mean, dim3(n), dim3(64), 0, stream, dhb.p, row.p, col.p, n, 1, tmp.p);
      hipLaunchKernelGGL((k_gemm<HID, HID>), dim3(tiles), dim3(64), 0, stream, dz.p, wn.p + o, nullptr, nullptr, nullptr,
                         tmp.p, n, 0, nullptr, nullptr, dh.p);
    }
    // input layer
    hipLaunchKernelGGL(k_relu_back, dim3(ceil_div(cnt, 256)), dim3(256), 0, stream, dh.p, z[0].p, cnt, dz.p, dzf.p);
    HIPCHECK((wgrad<K_IN, HID>(x.p, dz.p, OFF_W0)));
    colsum(dzf.p, HID, g.p + OFF_B0);
    hipLaunchKernelGGL(k_zero_padding, dim3(ceil_div(K_IN * HID, 256)), dim3(256), 0, stream, g.p);
    return hipGetLastError();
  }
};

std::vector<Net*> g_nets;
Net* nget(uint32_t h) { return h == 0 || h > g_nets.size() ? nullptr : g_nets[h - 1]; }

int32_t nstatus(hipError_t e) {
  publish_failure(e);
  if (e == hipSuccess) return 0;
  return e == hipErrorOutOfMemory ? -2 : -4;
}

}  // namespace
}  // namespace d3f

using namespace d3f;

EXPORT uint32_t gnn_create(uint64_t seed) {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return 0;
  Net* net = new (std::nothrow) Net();
  if (!net) return 0;
  if (net->init(seed) != hipSuccess) {
    delete net;
    return 0;
  }
  g_nets.push_back(net);
  return (uint32_t)g_nets.size();
}

EXPORT int32_t gnn_destroy(uint32_t h) {
  Net* net = nget(h);
  if (!net) return -1;
  delete net;
  g_nets[h - 1] = nullptr;
  return 0;
}

EXPORT uint32_t gnn_param_count(void) { return (uint32_t)N_PARAMS; }

EXPORT float* gnn_params(uint32_t h) {
  Net* net = nget(h);
  return net ? net->hparams.p : nullptr;
}

EXPORT int32_t gnn_sync_params(uint32_t h, uint32_t direction) {
  Net* net = nget(h);
  if (!net) return -1;
  if (direction == D3F_SYNC_H2D) return nstatus(net->upload_params());
  hipError_t e = hipMemcpyAsync(net->hparams.p, net->w.p, N_PARAMS * sizeof(float), hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  return nstatus(e);
}

EXPORT int32_t gnn_set_graph(uint32_t h, uint32_t n, const uint32_t* row, const uint32_t* col, const float* features) {
  Net* net = nget(h);
  if (!net) return -1;
  if (!row || (row[n] && !col) || (n && !features)) return -3;
  for (uint32_t i = 0; i < n; ++i)
    if (row[i + 1] < row[i]) return -3;
  for (uint32_t a = 0; a < row[n]; ++a)
    if (col[a] >= n) return -3;
  return nstatus(net->set_graph(n, row, col, features));
}

EXPORT int32_t gnn_set_links(uint32_t h, uint32_t n, uint32_t m, const uint32_t* source, const uint32_t* target,
                             uint64_t seed, float* features) {
  Net* net = nget(h);
  if (!net) return -1;
  if (m && (!source || !target)) return -3;
  for (uint32_t e = 0; e < m; ++e)
    if (source[e] >= n || target[e] >= n) return -3;
  return nstatus(net->set_links(n, m, source, target, seed, features));
}

EXPORT int32_t gnn_predict(uint32_t h, double jitter, uint64_t seed, double* positions) {
  Net* net = nget(h);
  if (!net) return -1;
  if (net->n && !positions) return -3;
  return nstatus(net->predict(jitter, seed, positions));
}

EXPORT int32_t gnn_forward(uint32_t h, float* positions) {
  Net* net = nget(h);
  if (!net) return -1;
  hipError_t e = net->forward();
  std::vector<float> out((size_t)net->rows * OUT_PAD);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data(), net->p.p, out.size() * sizeof(float), hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  if (e == hipSuccess)
    for (uint32_t i = 0; i < net->n; ++i) {
      positions[2 * i] = out[(size_t)i * OUT_PAD];
      positions[2 * i + 1] = out[(size_t)i * OUT_PAD + 1];
    }
  return nstatus(e);
}

EXPORT int32_t gnn_set_pairs(uint32_t h, uint32_t pairs, const uint32_t* pi, const uint32_t* pj, const float* d) {
  Net* net = nget(h);
  if (!net) return -1;
  for (uint32_t k = 0; k < pairs; ++k)
    if (pi[k] >= net->n || pj[k] >= net->n || !(d[k] > 0)) return -3;
  hipError_t e = net->pi.reserve(pairs > 0 ? pairs : 1);
  if (e == hipSuccess) e = net->pj.reserve(pairs > 0 ? pairs : 1);
  if (e == hipSuccess) e = net->pd.reserve(pairs > 0 ? pairs : 1);
  if (e == hipSuccess && pairs) e = hipMemcpy(net->pi.p, pi, pairs * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && pairs) e = hipMemcpy(net->pj.p, pj, pairs * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess && pairs) e = hipMemcpy(net->pd.p, d, pairs * 4, hipMemcpyHostToDevice);
  net->pairs = pairs;
  return nstatus(e);
}

// Gradient of sum(dp . P) for a given dP (n x 2, host) into grad (host,
// gnn_param_count() floats), for checking the backward pass.
EXPORT int32_t gnn_gradient(uint32_t h, const float* dp, float* grad) {
  Net* net = nget(h);
  if (!net) return -1;
  std::vector<float> d((size_t)net->rows * OUT_PAD, 0.f);
  for (uint32_t i = 0; i < net->n; ++i) {
    d[(size_t)i * OUT_PAD] = dp[2 * i];
    d[(size_t)i * OUT_PAD + 1] = dp[2 * i + 1];
  }
  hipError_t e = net->forward();
  if (e == hipSuccess) e = hipMemcpyAsync(net->dp.p, d.data(), d.size() * 4, hipMemcpyHostToDevice, net->stream);
  if (e == hipSuccess) e = net->backward();
  if (e == hipSuccess) e = hipMemcpyAsync(grad, net->g.p, N_PARAMS * 4, hipMemcpyDeviceToHost, net->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
  return nstatus(e);
}

// One Adam step on the stress of the current pairs; *loss receives the
// stress before the step.
EXPORT int32_t gnn_train_step(uint32_t h, float lr, float* loss) {
  Net* net = nget(h);
  if (!net) return -1;
  hipStream_t s = net->stream;
  hipError_t e = net->forward();
  if (e == hipSuccess) e = hipMemsetAsync(net->dp.p, 0, (size_t)net->rows * OUT_PAD * 4, s);
  if (e == hipSuccess) e = hipMemsetAsync(net->loss.p, 0, 4, s);
  if (e == hipSuccess && net->pairs)
    hipLaunchKernelGGL(k_stress, dim3(ceil_div(net->pairs, 256)), dim3(256), 0, s, net->p.p, net->pi.p, net->pj.p,
                       net->pd.p, net->pairs, net->dp.p, net->loss.p);
  if (e == hipSuccess) e = net->backward();
  net->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1 - std::pow(b1, (float)net->step), c2 = 1 - std::pow(b2, (float)net->step);
  if (e == hipSuccess)
    hipLaunchKernelGGL(k_adam, dim3(ceil_div((uint32_t)N_PARAMS, 256)), dim3(256), 0, s, net->w.p, net->g.p, net->m.p,
                       net->v.p, (uint32_t)N_PARAMS, lr, b1, b2, c1, c2, 1e-8f);
  if (e == hipSuccess) hipLaunchKernelGGL(k_zero_padding, dim3(ceil_div(K_IN * HID, 256)), dim3(256), 0, s, net->w.p);
  if (e == hipSuccess) e = net->pack();
  if (e == hipSuccess && loss) e = hipMemcpyAsync(loss, net->loss.p, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return nstatus(e);
}
