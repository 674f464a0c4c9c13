// center, x, y, radial, integrate and the shifted positions on gfx950
// (reference: src/core/simple.zig, src/core/sim.zig:129-146). All but center
// are per node. forceCenter's mean is d3's serial sum over nodes 0..n-1
// (simple.zig:18-22), which no reassociation may change: one workgroup stages
// x and y through LDS and one lane per axis adds them in index order.
#include <mutex>
#include <vector>

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

// ---- the same sums, by binade chunks and segments (n >= CHUNKED_MIN)
//
// Let s be a partial sum in the binade [2^e, 2^(e+1)) (or its negative) with
// ulp u = 2^(e-52): s = m u for an integer m, and while s + x lands strictly
// inside the same binade, fl(s + x) = (m + r) u with r = x / u rounded to the
// nearest integer, ties to the r that makes m + r even. So over a chunk of 32
// values whose partial sums all stay at least one ulp inside the binade, the
// chain is exactly m + (r_1 + ... + r_32); the r depend on m only through its
// parity at the ties. For each start parity p a chunk is a transform "start
// m in [L_p, H_p] -> m + D_p" (ulp units, int64, exact), and consecutive
// chunks in one binade compose into one transform of the same form: for
// T then B, L = max(L_T, L_B' - D_T), H = min(H_T, H_B' - D_T), D = D_T +
// D_B' with B' the variant for parity (p + D_T) & 1. The binade of each
// chunk is guessed from an approximate prefix sum (chunk sums scanned in
// another order); chunks whose guessed start is not safely inside their
// interval, or that have no transform, are "needing" chunks added one by
// one (their values are staged in order); runs of the others, within one
// binade and one 256-chunk block, become segments. The serial kernel walks
// the records: s += D when lo <= s <= hi (lo = L u, hi = H u, exact
// doubles), else the values one by one, so a wrong guess costs time only.
constexpr uint32_t CHUNKED_MIN = 65536;
constexpr uint32_t CC = 32;        // values per chunk
constexpr uint32_t SEG_CAP = 256;  // segments do not cross multiples of this many chunks
constexpr uint32_t CD = 8;         // doubles per record: lo0 hi0 D0 lo1 hi1 D1 (fidx, need) (chunk, len)
constexpr uint32_t CR = 128;       // records per staging round (two per lane); every needing one is staged
static_assert(CR == 128, "the chain reads two records per lane");
constexpr int CK_THREADS = 256;    // wave 0 adds, waves 1-3 stage
constexpr size_t CS_CHAIN_LDS = 65536;  // LDS a chain block takes in all (see center_raw_sum)
constexpr uint32_t CS_RTAB = 4096; // rounds whose staging indices are kept in LDS
constexpr uint32_t META_VALID = 1u << 13, META_NEG = 1u << 12;

__global__ __launch_bounds__(EW) void k_cs_chunk_sums(const double* __restrict__ xs, const double* __restrict__ ys,
                                                      uint32_t n, uint32_t nch, double* __restrict__ csum) {
  const uint32_t t = blockIdx.x * EW + threadIdx.x;
  if (t >= 2 * nch) return;
  const uint32_t axis = t >= nch, j = t - axis * nch;
  const double* p = axis ? ys : xs;
  const uint32_t b = j * CC, e = n - b < CC ? n : b + CC;
  double acc = 0;
  for (uint32_t k = b; k < e; ++k) acc += p[k];
  csum[t] = acc;
}

// Exclusive scan of one axis's chunk sums (block = axis).
__global__ __launch_bounds__(1024) void k_cs_start(const double* __restrict__ csum, uint32_t nch,
                                                   double* __restrict__ start) {
  __shared__ double part[1024];
  const uint32_t axis = blockIdx.x, tid = threadIdx.x;
  const uint32_t per = ceil_div(nch, 1024u);
  const uint32_t b = tid * per, e = b + per < nch ? b + per : nch;
  const double* c = csum + (size_t)axis * nch;
  double acc = 0;
  for (uint32_t k = b; k < e; ++k) acc += c[k];
  part[tid] = acc;
  __syncthreads();
  for (uint32_t d = 1; d < 1024; d <<= 1) {
    const double v = tid >= d ? part[tid - d] : 0.0;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  double run = tid ? part[tid - 1] : 0.0;
  for (uint32_t k = b; k < e; ++k) {
    start[(size_t)axis * nch + k] = run;
    run += c[k];
  }
}

__device__ inline double cs_pack(uint32_t lo, uint32_t hi) { return __hiloint2double((int)hi, (int)lo); }

// One chunk: its binade (from the guessed start), its two transforms and
// whether the guess is safely inside both.
__global__ __launch_bounds__(EW) void k_cs_desc(const double* __restrict__ xs, const double* __restrict__ ys,
                                                uint32_t n, uint32_t nch, const double* __restrict__ start,
                                                long long* __restrict__ ct, uint32_t* __restrict__ meta,
                                                uint32_t* __restrict__ need, uint32_t* __restrict__ count) {
  const uint32_t t = blockIdx.x * EW + threadIdx.x;
  if (t == 0) *count = 2 * nch;
  if (t >= 2 * nch) return;
  const uint32_t axis = t >= nch, j = t - axis * nch;
  const double* p = axis ? ys : xs;
  const double g = start[t];
  const uint32_t b = j * CC;
  uint32_t mt = 0;
  bool nd = true;
  long long L[2] = {1, 1}, H[2] = {0, 0}, D[2] = {0, 0};
  if (b + CC <= n && g != 0 && g - g == 0) {
    int E;
    (void)__builtin_frexp(g, &E);
    const int e = E - 1;  // |g| in [2^e, 2^(e+1))
    if (e > -960 && e < 1000) {
      const long long two52 = 4503599627370496LL, two53 = 2 * two52;
      // start and partial-sum ranges in ulp units
      const long long in_lo = g > 0 ? two52 : -(two53 - 1), in_hi = g > 0 ? two53 - 1 : -two52;
      const long long out_lo = g > 0 ? two52 + 1 : -(two53 - 1), out_hi = g > 0 ? two53 - 1 : -(two52 + 1);
      double v[CC];
#pragma unroll
      for (uint32_t k = 0; k < CC; ++k) v[k] = p[b + k];
      const double lim = 1152921504606846976.0;  // 2^60
      bool all_ok = true;
      for (int par0 = 0; par0 < 2; ++par0) {
        long long P = 0, mn = 0, mx = 0;
        int par = par0;
        bool ok = true;
#pragma unroll
        for (uint32_t k = 0; k < CC; ++k) {
          double q = __builtin_ldexp(v[k], 52 - e);  // x / u, exact unless |q| < 2^-1022
          if (!(__builtin_fabs(q) <= lim)) {
            ok = false;
            q = 0;
          }
          const double f = __builtin_floor(q), fr = q - f;
          long long r = (long long)f;
          if (fr > 0.5 || (fr == 0.5 && ((par + r) & 1))) r += 1;
          P += r;
          par = (int)((par + r) & 1);
          if (k == 0) {
            mn = P;
            mx = P;
          } else {
            mn = P < mn ? P : mn;
            mx = P > mx ? P : mx;
          }
        }
        if (!ok || mn < -two53 || mx > two53) {
          all_ok = false;
          continue;
        }
        L[par0] = in_lo > out_lo - mn ? in_lo : out_lo - mn;
        H[par0] = in_hi < out_hi - mx ? in_hi : out_hi - mx;
        D[par0] = P;
      }
      if (all_ok) {
        mt = META_VALID | (g < 0 ? META_NEG : 0u) | (uint32_t)(e + 1100);
        // the guess is inexact: keep it 2^-36 |g| inside both intervals
        const double u = __builtin_ldexp(1.0, e - 52);
        const double gm = g / u, mg = __builtin_fabs(gm) * 1.4551915228366852e-11;
        nd = !((double)L[0] + mg <= gm && gm <= (double)H[0] - mg && (double)L[1] + mg <= gm &&
               gm <= (double)H[1] - mg);
      }
    }
  }
  long long* c = ct + (size_t)t * 6;
  c[0] = L[0];
  c[1] = H[0];
  c[2] = D[0];
  c[3] = L[1];
  c[4] = H[1];
  c[5] = D[1];
  meta[t] = mt;
  need[t] = nd ? 1u : 0u;
}

// Segment starts, and each segment's composed record at its first chunk.
__global__ __launch_bounds__(EW) void k_cs_segments(uint32_t nch, const long long* __restrict__ ct,
                                                    const uint32_t* __restrict__ meta,
                                                    const uint32_t* __restrict__ need, uint32_t* __restrict__ segflag,
                                                    double* __restrict__ segtmp) {
  const uint32_t t = blockIdx.x * EW + threadIdx.x;
  if (t >= 2 * nch) return;
  const uint32_t j = t - (t >= nch) * nch;
  const bool st = need[t] || j % SEG_CAP == 0 || need[t - 1] || meta[t] != meta[t - 1];
  segflag[t] = st;
  if (!st) return;
  const double inf = __builtin_huge_val();
  double* rec = segtmp + (size_t)t * CD;
  uint32_t len = 1;
  if (need[t]) {
    rec[0] = inf;
    rec[1] = -inf;
    rec[2] = 0;
    rec[3] = inf;
    rec[4] = -inf;
    rec[5] = 0;
  } else {
    const long long* c = ct + (size_t)t * 6;
    long long L[2] = {c[0], c[3]}, H[2] = {c[1], c[4]}, D[2] = {c[2], c[5]};
    for (uint32_t k = t + 1; k < t - j + nch && (k - t + j) % SEG_CAP != 0 && !need[k] && meta[k] == meta[t]; ++k) {
      const long long* cb = ct + (size_t)k * 6;
#pragma unroll
      for (int pp = 0; pp < 2; ++pp) {
        const int q = (int)((pp + D[pp]) & 1);
        const long long lb = cb[3 * q] - D[pp], hb = cb[3 * q + 1] - D[pp];
        L[pp] = L[pp] > lb ? L[pp] : lb;
        H[pp] = H[pp] < hb ? H[pp] : hb;
        D[pp] += cb[3 * q + 2];
      }
      ++len;
    }
    const int e = (int)(meta[t] & 0xfff) - 1100;
    const double u = __builtin_ldexp(1.0, e - 52);
#pragma unroll
    for (int pp = 0; pp < 2; ++pp) {
      const bool okp = L[pp] <= H[pp];
      rec[3 * pp] = okp ? (double)L[pp] * u : inf;  // |L|, |H| <= 2^53: exact
      rec[3 * pp + 1] = okp ? (double)H[pp] * u : -inf;
      rec[3 * pp + 2] = okp ? (double)D[pp] * u : 0.0;  // |D| <= 2^53 when the interval is not empty
    }
  }
  rec[7] = cs_pack(j, len);
}

// Segment records in order, each with its staging index; per-axis record
// ranges; the staging index at each round of CR records.
__global__ __launch_bounds__(EW) void k_cs_seglist(uint32_t nch, const uint32_t* __restrict__ segflag,
                                                   const uint32_t* __restrict__ segidx,
                                                   const uint32_t* __restrict__ need,
                                                   const uint32_t* __restrict__ fidx,
                                                   const double* __restrict__ segtmp, double* __restrict__ seglist,
                                                   uint32_t* __restrict__ round_f, uint32_t* __restrict__ segrange,
                                                   uint32_t rstride) {
  const uint32_t t = blockIdx.x * EW + threadIdx.x;
  if (t >= 2 * nch) return;
  const uint32_t axis = t >= nch, j = t - axis * nch;
  const uint32_t base = segidx[axis * nch];
  if (j == 0) segrange[2 * axis] = base;
  if (j == nch - 1) {
    const uint32_t end = segidx[t] + segflag[t];
    segrange[2 * axis + 1] = end;
    round_f[axis * rstride + ceil_div(end - base, CR)] = fidx[t] + need[t];
  }
  if (!segflag[t]) return;
  const uint32_t k = segidx[t];
  double* r = seglist + (size_t)k * CD;
  const double* tmp = segtmp + (size_t)t * CD;
#pragma unroll
  for (int f = 0; f < 6; ++f) r[f] = tmp[f];
  r[6] = cs_pack(fidx[t], need[t]);
  r[7] = tmp[7];
  if ((k - base) % CR == 0) round_f[axis * rstride + (k - base) / CR] = fidx[t];
}

// Copies the values of needing chunks (the last, partial chunk padded with
// -0, which leaves any sum unchanged) into their staging slots.
__global__ __launch_bounds__(EW) void k_cs_stage(const double* __restrict__ xs, const double* __restrict__ ys,
                                                 uint32_t n, uint32_t nch, const uint32_t* __restrict__ need,
                                                 const uint32_t* __restrict__ fidx, double* __restrict__ staged) {
  const uint32_t t = blockIdx.x * EW + threadIdx.x;
  if (t >= 2 * nch || !need[t]) return;
  const uint32_t axis = t >= nch, j = t - axis * nch;
  const double* p = axis ? ys : xs;
  const uint32_t b = j * CC;
  const uint32_t f = fidx[t];
  for (uint32_t k = 0; k < CC; ++k) staged[(size_t)f * CC + k] = b + k < n ? p[b + k] : -0.0;
}

// The serial chain of one axis (block = axis): wave 0 walks the records;
// waves 1-3 stage the next round's records and staged values in LDS. Lane l
// of wave 0 reads records l and l + 64 of the round once; record k's fields
// are taken with readlane into scalar registers, and the interval test is a
// ballot, so every branch is wave-uniform (every lane holds the same s).
__global__ __launch_bounds__(CK_THREADS) void k_cs_chain(const double* __restrict__ xs, const double* __restrict__ ys,
                                                         uint32_t n, const double* __restrict__ seglist,
                                                         const double* __restrict__ staged,
                                                         const uint32_t* __restrict__ round_f,
                                                         const uint32_t* __restrict__ segrange, uint32_t rstride,
                                                         double* __restrict__ out) {
  __shared__ double dbuf[2][CR * CD];
  __shared__ double sbuf[2][CR * CC];
  __shared__ uint32_t f0buf[2];
  __shared__ uint32_t rtab[CS_RTAB];
  const uint32_t axis = blockIdx.x;
  const double* p = axis ? ys : xs;
  const uint32_t base = segrange[2 * axis], S = segrange[2 * axis + 1] - base;
  const double* rec = seglist + (size_t)base * CD;
  const uint32_t rounds = ceil_div(S, CR);
  const uint32_t* rf = round_f + axis * rstride;
  const bool tab = rounds + 1 <= CS_RTAB;
  if (tab)
    for (uint32_t q = threadIdx.x; q <= rounds; q += CK_THREADS) rtab[q] = rf[q];
  __syncthreads();
  const int wave = threadIdx.x >> 6;
  const uint32_t ptid = threadIdx.x - 64;  // producer thread 0..191
  // All loads of a round are issued before any LDS store: one memory round
  // trip per round.
  auto stage = [&](uint32_t r, int b) {
    constexpr uint32_t PT = CK_THREADS - 64;
    constexpr uint32_t ND = (CR * CD + PT - 1) / PT, NS = (CR * CC + PT - 1) / PT;
    const uint32_t c0 = r * CR, cnt = S - c0 < CR ? S - c0 : CR;
    const uint32_t f0 = tab ? rtab[r] : rf[r], m = (tab ? rtab[r + 1] : rf[r + 1]) - f0;  // m <= cnt
    double vd[ND], vs[NS];
#pragma unroll
    for (uint32_t u = 0; u < ND; ++u) {
      const uint32_t q = ptid + u * PT;
      vd[u] = q < cnt * CD ? rec[(size_t)c0 * CD + q] : 0.0;
    }
#pragma unroll
    for (uint32_t u = 0; u < NS; ++u) {
      const uint32_t q = ptid + u * PT;
      vs[u] = q < m * CC ? staged[(size_t)f0 * CC + q] : 0.0;
    }
#pragma unroll
    for (uint32_t u = 0; u < ND; ++u) {
      const uint32_t q = ptid + u * PT;
      if (q < cnt * CD) dbuf[b][q] = vd[u];
    }
#pragma unroll
    for (uint32_t u = 0; u < NS; ++u) {
      const uint32_t q = ptid + u * PT;
      if (q < m * CC) sbuf[b][q] = vs[u];
    }
    if (ptid == 0) f0buf[b] = f0;
  };
  double s = 0;
  if (wave > 0 && rounds) stage(0, 0);
  __syncthreads();
  for (uint32_t r = 0; r < rounds; ++r) {
    const int b = r & 1;
    if (wave > 0) {
      if (r + 1 < rounds) stage(r + 1, b ^ 1);
    } else {
      const uint32_t c0 = r * CR, cnt = S - c0 < CR ? S - c0 : CR;
      const uint32_t f0 = __builtin_amdgcn_readfirstlane(f0buf[b]);
      const int lane = threadIdx.x & 63;
      const double* d = dbuf[b];
      double da[CD], db[CD];
#pragma unroll
      for (uint32_t f = 0; f < CD; ++f) {
        da[f] = d[lane * CD + f];
        db[f] = d[(lane + 64) * CD + f];
      }
      auto rl = [](double v, uint32_t l) {
        return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                                __builtin_amdgcn_readlane(__double2loint(v), l));
      };
      auto record = [&](const double (&v)[CD], uint32_t l) {
        const double lo0 = rl(v[0], l), hi0 = rl(v[1], l), d0 = rl(v[2], l);
        const double lo1 = rl(v[3], l), hi1 = rl(v[4], l), d1 = rl(v[5], l);
        const bool odd = __builtin_amdgcn_readfirstlane(__double2loint(s)) & 1;
        const double lo = odd ? lo1 : lo0, hi = odd ? hi1 : hi0, dd = odd ? d1 : d0;
        if (__builtin_amdgcn_ballot_w64(lo <= s && s <= hi)) {
          s += dd;
          return;
        }
        const uint32_t fi = __builtin_amdgcn_readlane(__double2loint(v[6]), l);
        const uint32_t nd = __builtin_amdgcn_readlane(__double2hiint(v[6]), l);
        if (nd) {
          const double* x = sbuf[b] + (fi - f0) * CC;
          double a[CC];
#pragma unroll
          for (uint32_t e = 0; e < CC; ++e) a[e] = x[e];
#pragma unroll
          for (uint32_t e = 0; e < CC; ++e) s += a[e];
        } else {  // a segment whose guess failed: its values from global memory
          const uint32_t j = __builtin_amdgcn_readlane(__double2loint(v[7]), l);
          const uint32_t len = __builtin_amdgcn_readlane(__double2hiint(v[7]), l);
          const uint32_t lo_e = j * CC, hi_e = n - lo_e < len * CC ? n : lo_e + len * CC;
          // CC loads in flight at a time, then their adds in order (the
          // adds alone are the dependent chain).
          for (uint32_t e0 = lo_e; e0 < hi_e; e0 += CC) {
            double a[CC];
#pragma unroll
            for (uint32_t e = 0; e < CC; ++e) a[e] = e0 + e < hi_e ? p[e0 + e] : 0.0;
#pragma unroll
            for (uint32_t e = 0; e < CC; ++e)
              if (e0 + e < hi_e) s += a[e];
          }
        }
      };
      // The round first runs speculatively: a record's kind (needing or
      // segment) does not depend on s, so only a segment's transform choice
      // (by the parity of s) and its add sit on the chain; its interval test
      // is noted, not branched on. If any test failed, the round is redone
      // from its start by `record`, which branches on each test.
      auto quick = [&](const double (&v)[CD], uint32_t l, bool& bad) {
        if (__builtin_amdgcn_readlane(__double2hiint(v[6]), l)) {
          const uint32_t fi = __builtin_amdgcn_readlane(__double2loint(v[6]), l);
          const double* x = sbuf[b] + (fi - f0) * CC;
          double a[CC];
#pragma unroll
          for (uint32_t e = 0; e < CC; ++e) a[e] = x[e];
#pragma unroll
          for (uint32_t e = 0; e < CC; ++e) s += a[e];
        } else {
          const double lo0 = rl(v[0], l), hi0 = rl(v[1], l), d0 = rl(v[2], l);
          const double lo1 = rl(v[3], l), hi1 = rl(v[4], l), d1 = rl(v[5], l);
          const bool odd = __double2loint(s) & 1;
          const double lo = odd ? lo1 : lo0, hi = odd ? hi1 : hi0, dd = odd ? d1 : d0;
          bad = bad | !(lo <= s && s <= hi);
          s += dd;
        }
      };
      const uint32_t m1 = cnt < 64 ? cnt : 64;
      const double s0 = s;
      bool bad = false;
      for (uint32_t k = 0; k < m1; ++k) quick(da, k, bad);
      for (uint32_t k = 64; k < cnt; ++k) quick(db, k - 64, bad);
      if (__builtin_amdgcn_ballot_w64(bad)) {
        s = s0;
        for (uint32_t k = 0; k < m1; ++k) record(da, k);
        for (uint32_t k = 64; k < cnt; ++k) record(db, k - 64);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[axis] = s;
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
                                             const double* __restrict__ target, uint32_t n, double alpha,
                                             const double* __restrict__ alpha_dev) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  if (alpha_dev) alpha = *alpha_dev;
  v[i] += (target[i] - pos[i]) * strength[i] * alpha;
}

// simple.zig:64-86 (`a || 1e-6`, never jiggle).
__global__ __launch_bounds__(EW) void k_radial(const double* __restrict__ xs, const double* __restrict__ ys,
                                               double* __restrict__ vxs, double* __restrict__ vys,
                                               const double* __restrict__ strength,
                                               const double* __restrict__ radius, uint32_t n, double cx, double cy,
                                               double alpha, const double* __restrict__ alpha_dev) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  if (alpha_dev) alpha = *alpha_dev;
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

__global__ __launch_bounds__(EW) void k_shift_pack(const double* __restrict__ xs, const double* __restrict__ ys,
                                                   const double* __restrict__ vxs, const double* __restrict__ vys,
                                                   double2* __restrict__ pos, uint32_t n) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  pos[i] = make_double2(xs[i] + vxs[i], ys[i] + vys[i]);
}

// sim_snapshot's copy: out = vx[0..n) vy[0..n) and, when x is given, x, y.
__global__ __launch_bounds__(EW) void k_snapshot(const double* __restrict__ vxs, const double* __restrict__ vys,
                                                 const double* __restrict__ xs, const double* __restrict__ ys,
                                                 double* __restrict__ out, uint32_t n) {
  const uint32_t i = blockIdx.x * EW + threadIdx.x;
  if (i >= n) return;
  out[i] = vxs[i];
  out[(size_t)n + i] = vys[i];
  if (xs) {
    out[2 * (size_t)n + i] = xs[i];
    out[3 * (size_t)n + i] = ys[i];
  }
}

// Captured ticks: the start of a replay (one thread).
__global__ void k_tick_begin(TickState* __restrict__ st, const TickParams* __restrict__ params,
                             uint32_t* __restrict__ zero, uint32_t nzero) {
  if (threadIdx.x != 0) return;
  const TickParams p = params[st->tick];
  st->alpha = p.alpha;
  for (int k = 0; k < TICK_SEEDS; ++k) st->seed[k] = p.seed[k];
  for (uint32_t k = 0; k < nzero; ++k) zero[k] = 0;
}

// The tick's starting state into snap (vx, vy, x, y), unless an earlier
// replay of the run failed: snap then keeps the state that replay started from.
__global__ __launch_bounds__(EW) void k_tick_snapshot(const TickState* __restrict__ st, const double* __restrict__ xs,
                                                      const double* __restrict__ ys, const double* __restrict__ vxs,
                                                      const double* __restrict__ vys, uint32_t n,
                                                      double* __restrict__ out) {
  if (st->first_fail != NONE) return;
  for (uint32_t i = blockIdx.x * EW + threadIdx.x; i < n; i += gridDim.x * EW) {
    out[i] = vxs[i];
    out[(size_t)n + i] = vys[i];
    out[2 * (size_t)n + i] = xs[i];
    out[3 * (size_t)n + i] = ys[i];
  }
}

// The end of a replay: the first tick with a raised flag is kept.
__global__ void k_tick_end(TickState* __restrict__ st, TickNotes notes) {
  if (threadIdx.x != 0) return;
  uint32_t status = 0;
  for (uint32_t k = 0; k < notes.count; ++k)
    if (*notes.flag[k]) status |= notes.bit[k];
  if (status && st->first_fail == NONE) {
    st->first_fail = st->tick;
    st->status = status;
  }
  st->tick += 1;
}

}  // namespace

hipError_t snapshot_copy(const double* vx, const double* vy, const double* x, const double* y, uint32_t n,
                         double* out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_snapshot, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, vx, vy, x, y, out, n);
  return hipGetLastError();
}

hipError_t batch_center(double* x, double* y, const uint32_t* node_off, const uint32_t* node_graph, uint32_t graphs,
                        uint32_t n, double* sums, double cx, double cy, double strength, hipStream_t stream) {
  if (graphs == 0 || n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_center_sums, dim3(graphs), dim3(64), 0, stream, x, y, node_off, sums);
  hipLaunchKernelGGL(k_batch_center_apply, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, n, node_off, node_graph,
                     sums, cx, cy, strength);
  return hipGetLastError();
}

void center_work_free(CenterWork& w) {
  void* ptrs[] = {w.csum, w.start, w.ct, w.meta, w.need, w.fidx, w.segflag, w.segidx, w.segtmp, w.seglist,
                  w.round_f, w.segrange, w.staged, w.partials, w.count};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  w = CenterWork();
}

static uint32_t cs_rstride(uint32_t cap) { return cap / CR + 2; }

static hipError_t center_reserve(CenterWork& w, uint32_t nch) {
  if (nch <= w.cap) return hipSuccess;
  center_work_free(w);
  const uint32_t cap = nch + nch / 8 + 64;
  const size_t c2 = (size_t)2 * cap;
  HIPCHECK(hipMalloc(&w.csum, c2 * sizeof(double)));
  HIPCHECK(hipMalloc(&w.start, c2 * sizeof(double)));
  HIPCHECK(hipMalloc(&w.ct, c2 * 6 * sizeof(long long)));
  HIPCHECK(hipMalloc(&w.meta, c2 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.need, c2 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.fidx, c2 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.segflag, c2 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.segidx, c2 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.segtmp, c2 * CD * sizeof(double)));
  HIPCHECK(hipMalloc(&w.seglist, c2 * CD * sizeof(double)));
  HIPCHECK(hipMalloc(&w.round_f, (size_t)2 * cs_rstride(cap) * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.segrange, 4 * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.staged, c2 * CC * sizeof(double)));
  HIPCHECK(hipMalloc(&w.partials, (c2 / 4096 + 64) * sizeof(uint32_t)));
  HIPCHECK(hipMalloc(&w.count, sizeof(uint32_t)));
  w.cap = cap;
  return hipSuccess;
}

hipError_t tick_begin(TickState* st, const TickParams* params, uint32_t* zero, uint32_t nzero, hipStream_t stream) {
  hipLaunchKernelGGL(k_tick_begin, dim3(1), dim3(64), 0, stream, st, params, zero, nzero);
  return hipGetLastError();
}

hipError_t tick_snapshot(const TickState* st, const double* x, const double* y, const double* vx, const double* vy,
                         uint32_t n, double* snap, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_tick_snapshot, dim3(std::min(ceil_div(n, (uint32_t)EW), 2048u)), dim3(EW), 0, stream, st, x, y,
                     vx, vy, n, snap);
  return hipGetLastError();
}

hipError_t tick_end(TickState* st, const TickNotes& notes, hipStream_t stream) {
  hipLaunchKernelGGL(k_tick_end, dim3(1), dim3(64), 0, stream, st, notes);
  return hipGetLastError();
}

hipError_t center_raw_sum(const double* x, const double* y, uint32_t n, double* sums, CenterWork& w,
                          hipStream_t stream) {
  if (n < CHUNKED_MIN) {
    hipLaunchKernelGGL(k_center_raw_sum, dim3(1), dim3(CS_THREADS), 0, stream, x, y, n, sums);
    return hipGetLastError();
  }
  const uint32_t nch = ceil_div(n, CC);
  HIPCHECK(center_reserve(w, nch));
  const uint32_t rstride = cs_rstride(w.cap);
  const uint32_t g = ceil_div(2 * nch, (uint32_t)EW);
  hipLaunchKernelGGL(k_cs_chunk_sums, dim3(g), dim3(EW), 0, stream, x, y, n, nch, w.csum);
  hipLaunchKernelGGL(k_cs_start, dim3(2), dim3(1024), 0, stream, w.csum, nch, w.start);
  hipLaunchKernelGGL(k_cs_desc, dim3(g), dim3(EW), 0, stream, x, y, n, nch, w.start, w.ct, w.meta, w.need, w.count);
  hipLaunchKernelGGL(k_cs_segments, dim3(g), dim3(EW), 0, stream, nch, w.ct, w.meta, w.need, w.segflag, w.segtmp);
  HIPCHECK(scan_u32(w.need, w.fidx, 2 * nch, w.count, w.partials, nullptr, stream));
  HIPCHECK(scan_u32(w.segflag, w.segidx, 2 * nch, w.count, w.partials, nullptr, stream));
  hipLaunchKernelGGL(k_cs_seglist, dim3(g), dim3(EW), 0, stream, nch, w.segflag, w.segidx, w.need, w.fidx, w.segtmp,
                     w.seglist, w.round_f, w.segrange, rstride);
  hipLaunchKernelGGL(k_cs_stage, dim3(g), dim3(EW), 0, stream, x, y, n, nch, w.need, w.fidx, w.staged);
  // Each chain block asks for CS_CHAIN_LDS of LDS in all, so few other
  // workgroups share its SIMDs, yet it still fits on a CU beside the small
  // workgroups of the tick's other kernels (a block that needed a whole CU's
  // LDS would wait for a CU to drain completely, which the fast-mode kernels
  // of the tick may not leave until they end).
  static size_t pad = 0;
  if (!pad) {
    hipFuncAttributes fa;
    HIPCHECK(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(k_cs_chain)));
    pad = CS_CHAIN_LDS > fa.sharedSizeBytes ? CS_CHAIN_LDS - fa.sharedSizeBytes : 1;
    if (pad > 1)
      HIPCHECK(hipFuncSetAttribute(reinterpret_cast<const void*>(k_cs_chain),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)pad));
  }
  hipLaunchKernelGGL(k_cs_chain, dim3(2), dim3(CK_THREADS), pad > 1 ? pad : 0, stream, x, y, n, w.seglist,
                     w.staged, w.round_f, w.segrange, rstride, sums);
  return hipGetLastError();
}

hipError_t shared_streams(hipStream_t* main, hipStream_t* side, hipStream_t* aux) {
  static std::mutex lock;
  static hipStream_t sets[64][3];
  std::lock_guard<std::mutex> hold(lock);
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) HIPFAIL(hipErrorInvalidDevice, "device ordinal beyond the stream table");
  if (!sets[dev][0]) {
    hipStream_t k[3] = {nullptr, nullptr, nullptr};
    for (auto& q : k) HIPCHECK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
    for (int i = 0; i < 3; ++i) sets[dev][i] = k[i];
  }
  *main = sets[dev][0];
  *side = sets[dev][1];
  *aux = sets[dev][2];
  return hipSuccess;
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
                      double alpha, hipStream_t stream, const double* alpha_dev) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_axis, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, pos, v, strength, target, n, alpha,
                     alpha_dev);
  return hipGetLastError();
}

hipError_t radial_force(const double* x, const double* y, double* vx, double* vy, const double* strength,
                        const double* radius, uint32_t n, double cx, double cy, double alpha, hipStream_t stream,
                        const double* alpha_dev) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_radial, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, strength, radius, n, cx,
                     cy, alpha, alpha_dev);
  return hipGetLastError();
}

hipError_t shift_positions(const double* x, const double* y, const double* vx, const double* vy, double* px,
                           double* py, uint32_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, px, py, n);
  return hipGetLastError();
}

hipError_t shift_pack(const double* x, const double* y, const double* vx, const double* vy, double2* pos, uint32_t n,
                      hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_shift_pack, dim3(ceil_div(n, EW)), dim3(EW), 0, stream, x, y, vx, vy, pos, n);
  return hipGetLastError();
}

}  // namespace d3f
