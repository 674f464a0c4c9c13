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

struct Net {
  uint32_t n = 0, rows = 0, nnz = 0;  // rows: n padded to 64
  Pinned<float> hparams, hgrad;
  Device<float> w, g, m, v, part, cpart, z[LAYERS + 2], dh, dzf, p, dp, loss;
  Device<uint16_t> wt, wn, x, hh[LAYERS + 2], agg[LAYERS], dz, dhb, tmp;
  Device<uint32_t> row, col, pi, pj;
  Device<float> pd;
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
    const size_t r = rows;
    HIPCHECK(x.reserve(r * K_IN));
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
    // features as BF16, padded with zeros to K_IN columns and to `rows` rows
    std::vector<uint16_t> xb(r * K_IN, 0);
    for (uint32_t i = 0; i < n; ++i)
      for (int f = 0; f < F_IN; ++f) {
        uint32_t u;
        std::memcpy(&u, &feat[(size_t)i * F_IN + f], 4);
        if ((u & 0x7fffffffu) > 0x7f800000u) u |= 0x400000u;
        else u += 0x7fffu + ((u >> 16) & 1u);
        xb[(size_t)i * K_IN + f] = (uint16_t)(u >> 16);
      }
    HIPCHECK(hipMemcpy(x.p, xb.data(), xb.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
    for (auto& a : hh) HIPCHECK(hipMemset(a.p, 0, r * HID * sizeof(uint16_t)));
    for (auto& a : agg) HIPCHECK(hipMemset(a.p, 0, r * HID * sizeof(uint16_t)));
    for (auto& a : z) HIPCHECK(hipMemset(a.p, 0, r * HID * sizeof(float)));
    for (auto* a : {&dz, &dhb, &tmp}) HIPCHECK(hipMemset(a->p, 0, r * HID * sizeof(uint16_t)));
    HIPCHECK(hipMemset(dh.p, 0, r * HID * sizeof(float)));
    HIPCHECK(hipMemset(dzf.p, 0, r * HID * sizeof(float)));
    HIPCHECK(hipMemset(p.p, 0, r * OUT_PAD * sizeof(float)));
    HIPCHECK(hipMemset(dp.p, 0, r * OUT_PAD * sizeof(float)));
    return hipSuccess;
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
