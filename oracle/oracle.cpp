// CPU oracle for the d3-force tick: a serial C++17 restatement of the
// reference core (src/core/*.zig), used ONLY as the checker by tests/, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg. Nothing in the
// product links, loads or calls it.
//
// It exports the same C ABI as the GPU library (include/d3fgpu.h), under the
// prefix `oracle_`, so one host facade drives both and parity tests compare
// them bit for bit.
//
// Every function cites the reference file:line it restates. The reference's
// thread pool splits work into chunks but promises chunk-independent results
// (src/core/pool.zig:1-10); this restatement therefore runs each phase as one
// serial loop, except the optional threaded Barnes-Hut traversal used for the
// CPU baseline, which mirrors the reference's record-and-replay scheme
// (src/core/many_body.zig:31-84).
//
// Build: g++ -O2 -std=c++17 -ffp-contract=off -fno-fast-math -shared -fPIC
//        (oracle/build.sh). FP64 on x86-64 SSE2: every +, -, *, /, sqrt is
//        correctly rounded, as in wasm.
//
// Parity pins (tests/test_oracle_*.py): tree build against the reference's
// d3-quadtree traces (test/fixtures/tree-traces.json); the Zig unit-test known
// answers (src/core/tests.zig, sort.zig, hash.zig, fmm.zig:913-1000); the
// fast-mode tick against test/fixtures/fast-golden.json SHA-256 digests; the
// exact tick against an independent pure-Python restatement of d3-force's
// own manyBody/link/center/quadtree code on small inputs.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace {

constexpr uint32_t NONE = 0xffffffffu;
constexpr int KEY_LEVELS = 32;           // tree.zig:30
constexpr uint32_t MAX_DEPTH = 2200;     // tree.zig:33
const double NaN = std::numeric_limits<double>::quiet_NaN();
const double Inf = std::numeric_limits<double>::infinity();

// Status codes, root.zig:16-25.
enum Status : int32_t { OK = 0, GROWN = 1, BAD_HANDLE = -1, OUT_OF_MEMORY = -2, BAD_ARGUMENT = -3, WORKER_FAILED = -4 };

// force.zig:8, :13, :21.
enum Kind : uint32_t { K_CENTER, K_COLLIDE, K_LINK, K_MANY_BODY, K_RADIAL, K_X, K_Y, K_COUNT };
enum Scalar : uint32_t { S_STRENGTH, S_X, S_Y, S_THETA2, S_DMIN2, S_DMAX2, S_ITERATIONS, S_PRECISION, S_ORDER, S_OPENING, S_SEED, S_COUNT };
enum Array : uint32_t { A_F0, A_F1, A_F2, A_SOURCE, A_TARGET };

typedef double (*RandomFn)(void*);
RandomFn g_random = nullptr;
void* g_random_ctx = nullptr;

double js_random() {  // root.zig:139-145 (env.random)
  return g_random ? g_random(g_random_ctx) : 0.5;
}

// rng.zig:7-11
inline double jiggle(double r) { return (r - 0.5) * 1e-6; }

// ---------------------------------------------------------------- hash.zig
// splitmix64 finaliser over the pair and a salt, hash.zig:10-15.
inline uint64_t mix(uint64_t seed, uint32_t i, uint32_t j, uint32_t salt) {
  uint64_t h = seed ^ ((uint64_t)i << 32 | j) ^ ((uint64_t)salt * 0x9E3779B97F4A7C15ull);
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  return h ^ (h >> 31);
}
inline double hash_uniform(uint64_t seed, uint32_t i, uint32_t j, uint32_t salt) {  // hash.zig:18-20
  return (double)(mix(seed, i, j, salt) >> 11) * 0x1.0p-53;
}
inline double hash_jiggle(uint64_t seed, uint32_t i, uint32_t j, uint32_t salt) {  // hash.zig:23-25
  return (hash_uniform(seed, i, j, salt) - 0.5) * 1e-6;
}
inline uint64_t seed_from(double v) {  // hash.zig:28-30
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}

// Zig's @min/@max on finite floats: the first operand unless the second is
// strictly smaller (larger); NaN operands yield the other operand.
inline double zmin(double a, double b) { return (b < a || a != a) ? b : a; }
inline double zmax(double a, double b) { return (b > a || a != a) ? b : a; }

// ---------------------------------------------------------------- pool.zig
// The CPU baseline's thread pool (pool.zig:50-80 parallelFor): g_threads - 1
// persistent workers plus the calling thread take chunk indices from one
// counter. The reference promises chunk-independent results (pool.zig:1-10),
// and every threaded phase below writes each chunk's own slots, so the bits
// do not depend on the thread count.
int g_threads = 1;

class Pool {
 public:
  ~Pool() { resize(0); }
  // Runs fn(c) for every c < n on the pool; returns when all are done.
  void run(uint32_t n, const std::function<void(uint32_t)>& fn) {
    if ((int)workers_.size() != g_threads - 1) resize(g_threads - 1);
    if (workers_.empty() || n <= 1) {
      for (uint32_t c = 0; c < n; ++c) fn(c);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      job_ = &fn;
      total_ = n;
      next_.store(0);
      busy_ = (int)workers_.size();
      ++generation_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return busy_ == 0; });
    job_ = nullptr;
  }

 private:
  void drain() {
    for (;;) {
      const uint32_t c = next_.fetch_add(1);
      if (c >= total_) return;
      (*job_)(c);
    }
  }
  void resize(int k) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
      ++generation_;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    workers_.clear();
    stop_ = false;
    // Workers start from the current generation (read here, on the only
    // thread that calls run), so none can miss the next job.
    for (int i = 0; i < k; ++i) workers_.emplace_back([this, g = generation_] { loop(g); });
  }
  void loop(uint64_t seen) {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return generation_ != seen; });
        seen = generation_;
        if (stop_) return;
      }
      drain();
      std::lock_guard<std::mutex> lk(mu_);
      if (--busy_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(uint32_t)>* job_ = nullptr;
  std::atomic<uint32_t> next_{0};
  uint32_t total_ = 0;
  uint64_t generation_ = 0;
  int busy_ = 0;
  bool stop_ = false;
};
Pool g_pool;

inline uint32_t chunks(uint32_t n, uint32_t size) { return (n + size - 1) / size; }  // pool.zig:38-41

// ---------------------------------------------------------------- tree.zig
struct Cell {  // tree.zig:46-62, 88 bytes
  uint32_t child[4];
  uint32_t first, count;
  double x0, y0, x1, y1;
  double x, y, value, r;
  bool leaf() const { return child[0] == NONE && child[1] == NONE && child[2] == NONE && child[3] == NONE; }
};
static_assert(sizeof(Cell) == 88, "Cell layout");

struct Frame {  // tree.zig:64-76
  uint32_t lo, hi, depth, parent, slot;
  bool sorted;
  double x0, y0, x1, y1;
};

struct Tree {
  double extent[4] = {NaN, NaN, NaN, NaN};
  std::vector<Cell> cells;
  uint32_t len = 0, root = NONE, m = 0, depth = 0;
  std::vector<uint32_t> perm, perm_tmp, rest, diff;
  std::vector<uint64_t> keys, keys_tmp;
  // The last pooled build's job subtrees (contiguous cell ranges, tree.zig:18-22)
  // and the top cells above the cut, for the pooled accumulate.
  bool pooled = false;
  std::vector<uint32_t> job_base, job_cells, tops;

  // tree.zig:148-187 (d3-quadtree cover)
  void cover(double x, double y) {
    if (std::isnan(x) || std::isnan(y)) return;
    double x0 = extent[0], y0 = extent[1], x1 = extent[2], y1 = extent[3];
    if (std::isnan(x0)) {
      x0 = std::floor(x);
      x1 = x0 + 1;
      y0 = std::floor(y);
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
    extent[0] = x0; extent[1] = y0; extent[2] = x1; extent[3] = y1;
  }

  void reserve(uint32_t n) {
    if (perm.size() < n) {
      perm.resize(n); perm_tmp.resize(n); rest.resize(n); diff.resize(n); keys.resize(n); keys_tmp.resize(n);
    }
    if (cells.size() < 2 * (size_t)n + 64) cells.resize(2 * (size_t)n + 64);
  }

  // Box (tree.zig:316-340 with Prep.box :230-248): strict compares in index
  // order, so the first of -0 / +0 wins. Returns false with no finite point.
  bool box(const double* xs, const double* ys, uint32_t n, double b[4]) {
    double bx0 = Inf, by0 = Inf, bx1 = -Inf, by1 = -Inf;
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
      double x = xs[i], y = ys[i];
      if (x != x || y != y) continue;
      if (x < bx0) bx0 = x;
      if (x > bx1) bx1 = x;
      if (y < by0) by0 = y;
      if (y > by1) by1 = y;
      ++cnt;
    }
    m = cnt;
    if (bx0 > bx1 || by0 > by1) return false;
    b[0] = bx0; b[1] = by0; b[2] = bx1; b[3] = by1;
    return true;
  }

  // Stable sort of keys carrying perm (sort.zig:55-91 is a stable LSD radix
  // sort; a stable sort of the same keys has exactly one result).
  void stable_sort_keys(uint32_t count) {
    std::vector<uint32_t> idx(count);
    for (uint32_t k = 0; k < count; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
    std::vector<uint64_t> k2(count);
    std::vector<uint32_t> p2(count);
    for (uint32_t k = 0; k < count; ++k) { k2[k] = keys[idx[k]]; p2[k] = perm[idx[k]]; }
    std::copy(k2.begin(), k2.end(), keys.begin());
    std::copy(p2.begin(), p2.end(), perm.begin());
  }

  // tree.zig:371-388 prepare (+ exactKeys :252-284)
  bool prepare(const double* xs, const double* ys, uint32_t n) {
    reserve(n);
    extent[0] = extent[1] = extent[2] = extent[3] = NaN;
    len = 0; root = NONE; m = 0;
    double b[4];
    if (!box(xs, ys, n, b)) return false;
    cover(b[0], b[1]);
    cover(b[2], b[3]);
    uint32_t k = 0, r = 0;
    const double e0 = extent[0], e1 = extent[1], e2 = extent[2], e3 = extent[3];
    for (uint32_t i = 0; i < n; ++i) {
      double x = xs[i], y = ys[i];
      if (x != x || y != y) { rest[r++] = i; continue; }
      double x0 = e0, y0 = e1, x1 = e2, y1 = e3;
      uint64_t key = 0;
      for (int lev = 0; lev < KEY_LEVELS; ++lev) {
        double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
        bool right = x >= xm;
        if (right) x0 = xm; else x1 = xm;
        bool bottom = y >= ym;
        if (bottom) y0 = ym; else y1 = ym;
        key = key << 2 | (uint64_t)bottom << 1 | (uint64_t)right;
      }
      perm[k] = i;
      keys[k] = key;
      ++k;
    }
    stable_sort_keys(m);
    return true;
  }

  // tree.zig:392-433: diff[k] = positions <= k that differ from their predecessor.
  void run_boundaries(const double* xs, const double* ys) {
    if (m == 0) return;
    diff[0] = 0;
    uint32_t d = 0;
    for (uint32_t k = 1; k < m; ++k) {
      uint32_t p = perm[k], q = perm[k - 1];
      d += (xs[p] != xs[q] || ys[p] != ys[q]) ? 1u : 0u;
      diff[k] = d;
    }
  }

  // tree.zig:437-444
  bool all_same(const Frame& f, const double* xs, const double* ys) const {
    if (f.sorted) return diff[f.hi - 1] == diff[f.lo];
    uint32_t p0 = perm[f.lo];
    for (uint32_t k = f.lo + 1; k < f.hi; ++k) {
      uint32_t p = perm[k];
      if (xs[p] != xs[p0] || ys[p] != ys[p0]) return false;
    }
    return true;
  }

  static unsigned quadrant(double x, double y, double xm, double ym) {  // tree.zig:697-699
    return (unsigned)(y >= ym) << 1 | (unsigned)(x >= xm);
  }

  // tree.zig:449-479
  void split(const Frame& f, const double* xs, const double* ys, uint32_t out[5]) {
    out[0] = f.lo;
    out[4] = f.hi;
    if (f.sorted && f.depth < (uint32_t)KEY_LEVELS) {
      unsigned shift = 62 - 2 * f.depth;
      uint32_t k = f.lo;
      for (uint32_t q = 1; q < 4; ++q) {
        while (k < f.hi && ((keys[k] >> shift) & 3) < q) ++k;
        out[q] = k;
      }
    } else {
      double xm = (f.x0 + f.x1) / 2, ym = (f.y0 + f.y1) / 2;
      uint32_t counts[4] = {0, 0, 0, 0};
      for (uint32_t k = f.lo; k < f.hi; ++k) counts[quadrant(xs[perm[k]], ys[perm[k]], xm, ym)]++;
      uint32_t off[4], sum = f.lo;
      for (int q = 0; q < 4; ++q) { off[q] = sum; out[q] = sum; sum += counts[q]; }
      for (uint32_t k = f.lo; k < f.hi; ++k) {
        uint32_t p = perm[k];
        perm_tmp[off[quadrant(xs[p], ys[p], xm, ym)]++] = p;
      }
      std::copy(perm_tmp.begin() + f.lo, perm_tmp.begin() + f.hi, perm.begin() + f.lo);
    }
  }

  // tree.zig:482-499
  static Frame child(const Frame& f, unsigned q, uint32_t lo, uint32_t hi, uint32_t parent) {
    double xm = (f.x0 + f.x1) / 2, ym = (f.y0 + f.y1) / 2;
    bool right = (q & 1) != 0, bottom = (q & 2) != 0;
    Frame c;
    c.lo = lo; c.hi = hi; c.depth = f.depth + 1; c.parent = parent; c.slot = q;
    c.sorted = f.sorted && f.depth < (uint32_t)KEY_LEVELS;
    c.x0 = right ? xm : f.x0;
    c.y0 = bottom ? ym : f.y0;
    c.x1 = right ? f.x1 : xm;
    c.y1 = bottom ? f.y1 : ym;
    return c;
  }

  // tree.zig:572-647 emit, as the serial preorder walk the reference's job
  // decomposition reproduces (tree.zig:18-22, :621-646): the cells of the
  // subtree of `top`, numbered from 0 in preorder, into out.
  uint32_t emit_from(const Frame& top, const double* xs, const double* ys, std::vector<Cell>& out,
                     uint32_t* max_depth) {
    std::vector<Frame> stack;
    stack.push_back(top);
    stack.back().parent = NONE;
    uint32_t idx = 0;
    while (!stack.empty()) {
      Frame f = stack.back();
      stack.pop_back();
      if (out.size() <= idx) out.resize(out.size() * 2 + 64);
      Cell& c = out[idx];
      c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
      c.first = f.lo; c.count = f.hi - f.lo;
      c.x0 = f.x0; c.y0 = f.y0; c.x1 = f.x1; c.y1 = f.y1;
      c.x = c.y = c.value = c.r = 0;
      if (f.parent != NONE) out[f.parent].child[f.slot] = idx;
      uint32_t me = idx++;
      *max_depth = std::max(*max_depth, f.depth);
      if (f.depth >= MAX_DEPTH || all_same(f, xs, ys)) continue;
      uint32_t sp5[5];
      split(f, xs, ys, sp5);
      for (int q = 3; q >= 0; --q) {
        if (sp5[q] == sp5[q + 1]) continue;
        stack.push_back(child(f, (unsigned)q, sp5[q], sp5[q + 1], me));
      }
    }
    return idx;
  }

  void emit(const double* xs, const double* ys) {
    depth = 0;
    len = emit_from(Frame{0, m, 0, NONE, 0, true, extent[0], extent[1], extent[2], extent[3]}, xs, ys, cells, &depth);
    root = 0;
  }

  // tree.zig:191-196
  void build(const double* xs, const double* ys, uint32_t n) {
    pooled = false;
    if (g_threads > 1) {
      build_pooled(xs, ys, n);
      return;
    }
    if (!prepare(xs, ys, n)) return;
    run_boundaries(xs, ys);
    emit(xs, ys);
  }

  // ---- The same build on the pool, with the reference's chunking: box by
  // BOX_ROWS x CHUNK points, keys by CHUNK points, the radix sort by
  // SORT_CHUNK keys, run boundaries by RUN_CHUNK points, emit by job
  // subtrees below the cut (tree.zig:18-39, sort.zig:1-13, pool.zig:50-54).
  // Every chunk writes its own slots, so the result is the serial one.
  static constexpr uint32_t CHUNK = 1024, BOX_ROWS = 16, RUN_CHUNK = 16384, SORT_CHUNK = 16384;

  bool box_pooled(const double* xs, const double* ys, uint32_t n, double b[4]) {
    const uint32_t size = BOX_ROWS * CHUNK, nc = chunks(n, size);
    std::vector<double> part(4 * (size_t)nc);
    std::vector<uint32_t> cnt(nc);
    g_pool.run(nc, [&](uint32_t c) {
      double bx0 = Inf, by0 = Inf, bx1 = -Inf, by1 = -Inf;
      uint32_t k = 0;
      const uint32_t lo = c * size, hi = std::min(n, lo + size);
      for (uint32_t i = lo; i < hi; ++i) {
        const double x = xs[i], y = ys[i];
        if (x != x || y != y) continue;
        if (x < bx0) bx0 = x;
        if (x > bx1) bx1 = x;
        if (y < by0) by0 = y;
        if (y > by1) by1 = y;
        ++k;
      }
      part[4 * c] = bx0; part[4 * c + 1] = by0; part[4 * c + 2] = bx1; part[4 * c + 3] = by1;
      cnt[c] = k;
    });
    // Chunk partials in chunk order with the same strict compares: the first
    // of equal-comparing values (-0 / +0) still wins.
    double bx0 = Inf, by0 = Inf, bx1 = -Inf, by1 = -Inf;
    uint32_t total = 0;
    for (uint32_t c = 0; c < nc; ++c) {
      if (part[4 * c] < bx0) bx0 = part[4 * c];
      if (part[4 * c + 2] > bx1) bx1 = part[4 * c + 2];
      if (part[4 * c + 1] < by0) by0 = part[4 * c + 1];
      if (part[4 * c + 3] > by1) by1 = part[4 * c + 3];
      total += cnt[c];
    }
    m = total;
    if (bx0 > bx1 || by0 > by1) return false;
    b[0] = bx0; b[1] = by0; b[2] = bx1; b[3] = by1;
    return true;
  }

  // sort.zig:55-91: stable LSD radix sort of keys[0..m) carrying perm, 8-bit
  // digits, a pass skipped when one digit covers every key.
  void radix_sort_pooled(uint32_t bytes) {
    if (m == 0) return;
    const uint32_t nc = chunks(m, SORT_CHUNK);
    std::vector<uint32_t> hist(256 * (size_t)nc);
    for (uint32_t b = 0; b < bytes; ++b) {
      const unsigned shift = 8 * b;
      g_pool.run(nc, [&](uint32_t c) {
        uint32_t* h = hist.data() + 256 * (size_t)c;
        std::fill(h, h + 256, 0u);
        const uint32_t lo = c * SORT_CHUNK, hi = std::min(m, lo + SORT_CHUNK);
        for (uint32_t k = lo; k < hi; ++k) h[(keys[k] >> shift) & 255]++;
      });
      bool skip = false;
      for (uint32_t d = 0; d < 256 && !skip; ++d) {
        uint32_t t = 0;
        for (uint32_t c = 0; c < nc; ++c) t += hist[256 * (size_t)c + d];
        skip = t == m;
      }
      if (skip) continue;
      uint32_t run = 0;
      for (uint32_t d = 0; d < 256; ++d)
        for (uint32_t c = 0; c < nc; ++c) {
          uint32_t& h = hist[256 * (size_t)c + d];
          const uint32_t v = h;
          h = run;
          run += v;
        }
      g_pool.run(nc, [&](uint32_t c) {
        uint32_t* h = hist.data() + 256 * (size_t)c;
        const uint32_t lo = c * SORT_CHUNK, hi = std::min(m, lo + SORT_CHUNK);
        for (uint32_t k = lo; k < hi; ++k) {
          const uint32_t d = (keys[k] >> shift) & 255;
          keys_tmp[h[d]] = keys[k];
          perm_tmp[h[d]] = perm[k];
          h[d]++;
        }
      });
      keys.swap(keys_tmp);
      perm.swap(perm_tmp);
    }
  }

  void build_pooled(const double* xs, const double* ys, uint32_t n) {
    reserve(n);
    extent[0] = extent[1] = extent[2] = extent[3] = NaN;
    len = 0; root = NONE; m = 0;
    double b[4];
    if (!box_pooled(xs, ys, n, b)) return;
    cover(b[0], b[1]);
    cover(b[2], b[3]);
    // exactKeys (tree.zig:252-284) by chunks: count the finite points of
    // each chunk, then each chunk writes its keys and NaN nodes at its offsets.
    const uint32_t nc = chunks(n, CHUNK);
    std::vector<uint32_t> off(nc + 1, 0);
    g_pool.run(nc, [&](uint32_t c) {
      uint32_t k = 0;
      const uint32_t lo = c * CHUNK, hi = std::min(n, lo + CHUNK);
      for (uint32_t i = lo; i < hi; ++i) k += (xs[i] == xs[i] && ys[i] == ys[i]) ? 1u : 0u;
      off[c + 1] = k;
    });
    for (uint32_t c = 0; c < nc; ++c) off[c + 1] += off[c];
    const double e0 = extent[0], e1 = extent[1], e2 = extent[2], e3 = extent[3];
    g_pool.run(nc, [&](uint32_t c) {
      const uint32_t lo = c * CHUNK, hi = std::min(n, lo + CHUNK);
      uint32_t k = off[c], r = lo - off[c];
      for (uint32_t i = lo; i < hi; ++i) {
        const double x = xs[i], y = ys[i];
        if (x != x || y != y) { rest[r++] = i; continue; }
        double x0 = e0, y0 = e1, x1 = e2, y1 = e3;
        uint64_t key = 0;
        for (int lev = 0; lev < KEY_LEVELS; ++lev) {
          const double xm = (x0 + x1) / 2, ym = (y0 + y1) / 2;
          const bool right = x >= xm;
          if (right) x0 = xm; else x1 = xm;
          const bool bottom = y >= ym;
          if (bottom) y0 = ym; else y1 = ym;
          key = key << 2 | (uint64_t)bottom << 1 | (uint64_t)right;
        }
        perm[k] = i;
        keys[k] = key;
        ++k;
      }
    });
    radix_sort_pooled(8);
    // runBoundaries (tree.zig:392-433) by chunks: differences per chunk, then
    // each chunk writes its running count.
    if (m) {
      const uint32_t rc = chunks(m, RUN_CHUNK);
      std::vector<uint32_t> d(rc + 1, 0);
      auto differs = [&](uint32_t k) {
        const uint32_t p = perm[k], q = perm[k - 1];
        return (xs[p] != xs[q] || ys[p] != ys[q]) ? 1u : 0u;
      };
      g_pool.run(rc, [&](uint32_t c) {
        uint32_t t = 0;
        const uint32_t lo = std::max(1u, c * RUN_CHUNK), hi = std::min(m, c * RUN_CHUNK + RUN_CHUNK);
        for (uint32_t k = lo; k < hi; ++k) t += differs(k);
        d[c + 1] = t;
      });
      for (uint32_t c = 0; c < rc; ++c) d[c + 1] += d[c];
      g_pool.run(rc, [&](uint32_t c) {
        uint32_t t = d[c];
        const uint32_t lo = c * RUN_CHUNK, hi = std::min(m, lo + RUN_CHUNK);
        for (uint32_t k = lo; k < hi; ++k) {
          if (k > 0) t += differs(k);
          diff[k] = t;
        }
      });
    }
    emit_pooled(xs, ys);
    pooled = true;
  }

  // Emit by subtrees (tree.zig:572-647, the cut at max(m / 64, 64) points,
  // pool.zig:50-54): the top of the tree is walked in preorder on the calling
  // thread; a frame of at most `limit` points becomes a job whose subtree is
  // emitted on the pool; every cell then goes to the index the serial
  // preorder walk gives it.
  void emit_pooled(const double* xs, const double* ys) {
    const uint32_t limit = std::max(m / 64, 64u);
    std::vector<Cell> top_cells;
    std::vector<Frame> job_frames;
    std::vector<uint32_t> item_ref;  // per preorder item: its top cell or job
    std::vector<uint8_t> item_job;
    std::vector<Frame> stack;
    depth = 0;
    stack.push_back(Frame{0, m, 0, NONE, 0, true, extent[0], extent[1], extent[2], extent[3]});
    while (!stack.empty()) {
      Frame f = stack.back();
      stack.pop_back();
      const uint32_t item = (uint32_t)item_ref.size();
      if (f.parent != NONE) top_cells[item_ref[f.parent]].child[f.slot] = item;
      if (f.hi - f.lo <= limit) {
        item_job.push_back(1);
        item_ref.push_back((uint32_t)job_frames.size());
        job_frames.push_back(f);
        continue;
      }
      Cell c;
      c.child[0] = c.child[1] = c.child[2] = c.child[3] = NONE;
      c.first = f.lo; c.count = f.hi - f.lo;
      c.x0 = f.x0; c.y0 = f.y0; c.x1 = f.x1; c.y1 = f.y1;
      c.x = c.y = c.value = c.r = 0;
      item_job.push_back(0);
      item_ref.push_back((uint32_t)top_cells.size());
      top_cells.push_back(c);
      depth = std::max(depth, f.depth);
      if (f.depth >= MAX_DEPTH || all_same(f, xs, ys)) continue;
      uint32_t sp5[5];
      split(f, xs, ys, sp5);
      for (int q = 3; q >= 0; --q) {
        if (sp5[q] == sp5[q + 1]) continue;
        stack.push_back(child(f, (unsigned)q, sp5[q], sp5[q + 1], item));
      }
    }
    const uint32_t nj = (uint32_t)job_frames.size();
    std::vector<std::vector<Cell>> local(nj);
    std::vector<uint32_t> local_n(nj), local_depth(nj, 0);
    g_pool.run(nj, [&](uint32_t j) { local_n[j] = emit_from(job_frames[j], xs, ys, local[j], &local_depth[j]); });
    // Preorder indices: the items in order, a job taking its subtree's cells.
    std::vector<uint32_t> index(item_ref.size());
    job_base.assign(nj, 0);
    job_cells.assign(local_n.begin(), local_n.end());
    tops.clear();
    uint32_t idx = 0;
    for (size_t it = 0; it < item_ref.size(); ++it) {
      index[it] = idx;
      if (item_job[it]) {
        job_base[item_ref[it]] = idx;
        idx += local_n[item_ref[it]];
      } else {
        tops.push_back(idx);
        idx += 1;
      }
    }
    if (cells.size() < idx) cells.resize(idx);
    for (uint32_t j = 0; j < nj; ++j) depth = std::max(depth, local_depth[j]);
    for (size_t it = 0; it < item_ref.size(); ++it) {
      if (item_job[it]) continue;
      Cell c = top_cells[item_ref[it]];
      for (int q = 0; q < 4; ++q)
        if (c.child[q] != NONE) c.child[q] = index[c.child[q]];
      cells[index[it]] = c;
    }
    g_pool.run(nj, [&](uint32_t j) {
      const uint32_t base = job_base[j];
      for (uint32_t k = 0; k < local_n[j]; ++k) {
        Cell c = local[j][k];
        for (int q = 0; q < 4; ++q)
          if (c.child[q] != NONE) c.child[q] += base;
        cells[base + k] = c;
      }
    });
    root = 0;
    len = idx;
  }

  // tree.zig:652-684: d3-quadtree visit (preorder) / visitAfter (post-order) cell order.
  void order(bool after, uint32_t* out) const {
    if (root == NONE) return;
    std::vector<uint32_t> stack;
    stack.push_back(root);
    size_t w = 0;
    while (!stack.empty()) {
      uint32_t c = stack.back();
      stack.pop_back();
      out[w++] = c;
      const uint32_t* ch = cells[c].child;
      if (after) {
        for (int i = 0; i < 4; ++i) if (ch[i] != NONE) stack.push_back(ch[i]);
      } else {
        for (int i = 3; i >= 0; --i) if (ch[i] != NONE) stack.push_back(ch[i]);
      }
    }
    if (after) std::reverse(out, out + w);
  }

  // Fast-mode keys, tree.zig:287-309 + prepareFast :346-364.
  static uint64_t spread16(uint32_t v) {  // tree.zig:688-695
    uint64_t x = v & 0xFFFF;
    x = (x | (x << 8)) & 0x00FF00FF;
    x = (x | (x << 4)) & 0x0F0F0F0F;
    x = (x | (x << 2)) & 0x33333333;
    x = (x | (x << 1)) & 0x55555555;
    return x;
  }
  bool prepare_fast(const double* xs, const double* ys, uint32_t n) {
    reserve(n);
    extent[0] = extent[1] = extent[2] = extent[3] = NaN;
    len = 0; root = NONE;
    double b[4];
    if (!box(xs, ys, n, b)) return false;
    double side = zmax(b[2] - b[0], b[3] - b[1]);
    if (!(side > 0)) side = 1;
    extent[0] = b[0]; extent[1] = b[1]; extent[2] = b[0] + side; extent[3] = b[1] + side;
    double scale = 65536.0 / side;
    uint32_t k = 0, r = 0;
    for (uint32_t i = 0; i < n; ++i) {
      double x = xs[i], y = ys[i];
      if (x != x || y != y) { rest[r++] = i; continue; }
      uint32_t qx = (uint32_t)zmin(65535, std::floor((x - b[0]) * scale));
      uint32_t qy = (uint32_t)zmin(65535, std::floor((y - b[1]) * scale));
      perm[k] = i;
      keys[k] = spread16(qx) | spread16(qy) << 1;
      ++k;
    }
    stable_sort_keys(m);
    return true;
  }
};

// ---------------------------------------------------------------- force.zig / sim.zig
struct Force {  // force.zig:28-44
  Kind kind;
  double scalars[S_COUNT] = {0};
  uint32_t count = 0;
  std::vector<double> f0, f1, f2;
  std::vector<uint32_t> source, target;
  std::vector<uint32_t> csr_ptr, csr_adj;
  uint32_t csr_n = 0;
  explicit Force(Kind k) : kind(k) {}
  bool node_sized() const { return kind != K_LINK && kind != K_CENTER; }  // force.zig:50-52
  static int float_arrays(Kind k) {  // force.zig:54-61
    switch (k) {
      case K_CENTER: return 0;
      case K_COLLIDE: case K_MANY_BODY: return 1;
      case K_RADIAL: case K_X: case K_Y: return 2;
      default: return 3;
    }
  }
  // force.zig:65-82: resize every array, new float slots NaN.
  void set_count(uint32_t c) {
    uint32_t old = count;
    int k = float_arrays(kind);
    std::vector<double>* fs[3] = {&f0, &f1, &f2};
    for (int i = 0; i < k; ++i) fs[i]->resize(c, NaN);
    if (kind == K_LINK) { source.resize(c, 0); target.resize(c, 0); }
    (void)old;
    count = c;
    csr_n = 0;
  }
  double scalar(Scalar s) const { return scalars[s]; }
};

struct Fmm;

struct Sim {  // sim.zig:16-43
  uint32_t n = 0, cap = 0;
  std::vector<double> x, y, vx, vy, fx, fy, px, py, snap;
  std::vector<uint8_t> fixed;
  Tree tree;
  std::vector<Force*> forces;
  Fmm* fmm = nullptr;
  // grid scratch (grid.zig:20-36)
  std::vector<uint32_t> g_start, g_items, g_cell_of;
  std::vector<uint64_t> g_keys;
  ~Sim();

  bool set_node_count(uint32_t nn) {  // sim.zig:59-73
    n = nn;
    bool grown = false;
    for (Force* f : forces) if (f && f->node_sized() && f->count < nn) { f->set_count(nn); grown = true; }
    if (nn <= cap) return grown;
    uint32_t c = std::max(std::max(nn, cap * 2), 16u);
    for (auto* v : {&x, &y, &vx, &vy, &fx, &fy, &px, &py}) v->resize(c);
    fixed.resize(c);
    cap = c;
    return true;
  }
  void shift() {  // sim.zig:129-146
    for (uint32_t i = 0; i < n; ++i) { px[i] = x[i] + vx[i]; py[i] = y[i] + vy[i]; }
  }
};

// ---------------------------------------------------------------- many_body.zig (exact)
struct Counters { uint64_t visits = 0, accepts = 0, pairs = 0; };
Counters g_count;
bool g_counting = false;

// many_body.zig:101-135 Acc.cell.
void accumulate_cell(Tree& t, uint32_t c, const double* xs, const double* ys, const double* strengths) {
  Cell& quad = t.cells[c];
  double strength = 0;
  if (!quad.leaf()) {
    double weight = 0, x = 0, y = 0;
    for (int k = 0; k < 4; ++k) {
      uint32_t ch = quad.child[k];
      if (ch == NONE) continue;
      const Cell& q = t.cells[ch];
      double w = std::fabs(q.value);
      if (w > 0) {
        strength += q.value;
        weight += w;
        x += w * q.x;
        y += w * q.y;
      }
    }
    quad.x = x / weight;
    quad.y = y / weight;
  } else {
    uint32_t head = t.perm[quad.first + quad.count - 1];
    quad.x = xs[head];
    quad.y = ys[head];
    for (uint32_t k = quad.first + quad.count; k > quad.first;) {
      --k;
      strength += strengths[t.perm[k]];
    }
  }
  quad.value = strength;
}

void accumulate(Tree& t, const double* xs, const double* ys, const double* strengths) {
  // many_body.zig:40-46. Reverse preorder is a valid post-order: every child
  // has a larger index than its parent. After a pooled build, each job's
  // subtree (one contiguous range) runs on the pool, then the top cells, as
  // the reference does; only the 0..3 order within a cell affects the sums.
  if (t.pooled) {
    g_pool.run((uint32_t)t.job_base.size(), [&](uint32_t j) {
      for (uint32_t c = t.job_base[j] + t.job_cells[j]; c-- > t.job_base[j];) accumulate_cell(t, c, xs, ys, strengths);
    });
    for (size_t k = t.tops.size(); k-- > 0;) accumulate_cell(t, t.tops[k], xs, ys, strengths);
    return;
  }
  for (uint32_t c = t.len; c-- > 0;) accumulate_cell(t, c, xs, ys, strengths);
}

struct TravParams {
  const Tree* t;
  const double *xs, *ys, *strengths;
  double *vxs, *vys;
  double theta2, dmin2, dmax2, alpha;
};

// many_body.zig:176-263 Trav.node. With `record`, returns false without
// writing at the first point where d3 would call random.
bool trav_node(const TravParams& p, bool record, uint32_t i, std::vector<uint32_t>& stack, Counters* cnt) {
  const Cell* cells = p.t->cells.data();
  const uint32_t* perm = p.t->perm.data();
  const double nx = p.xs[i], ny = p.ys[i];
  double vx = p.vxs[i], vy = p.vys[i];
  stack.clear();
  stack.push_back(p.t->root);
  while (!stack.empty()) {
    const Cell& quad = cells[stack.back()];
    stack.pop_back();
    const double value = quad.value;
    if (!(value != 0 && value == value)) continue;
    if (cnt) cnt->visits++;
    double x = quad.x - nx, y = quad.y - ny;
    const double w = quad.x1 - quad.x0;
    double l = x * x + y * y;
    if (w * w / p.theta2 < l) {
      if (l < p.dmax2) {
        if (record && (x == 0 || y == 0)) return false;
        if (x == 0) { x = jiggle(js_random()); l += x * x; }
        if (y == 0) { y = jiggle(js_random()); l += y * y; }
        if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);
        vx += x * value * p.alpha / l;
        vy += y * value * p.alpha / l;
        if (cnt) cnt->accepts++;
      }
      continue;
    }
    if (!quad.leaf()) {
      for (int q = 3; q >= 0; --q) if (quad.child[q] != NONE) stack.push_back(quad.child[q]);
      continue;
    }
    if (l >= p.dmax2) continue;
    uint32_t head = perm[quad.first + quad.count - 1];
    if (head != i || quad.count > 1) {
      if (record && (x == 0 || y == 0)) return false;
      if (x == 0) { x = jiggle(js_random()); l += x * x; }
      if (y == 0) { y = jiggle(js_random()); l += y * y; }
      if (l < p.dmin2) l = std::sqrt(p.dmin2 * l);
    }
    for (uint32_t k = quad.first + quad.count; k > quad.first;) {
      --k;
      uint32_t j = perm[k];
      if (j != i) {
        double wj = p.strengths[j] * p.alpha / l;
        vx += x * wj;
        vy += y * wj;
        if (cnt) cnt->pairs++;
      }
    }
  }
  p.vxs[i] = vx;
  p.vys[i] = vy;
  return true;
}

void many_body_exact(Sim& s, Force& f, double alpha) {  // many_body.zig:31-84
  Tree& t = s.tree;
  t.build(s.x.data(), s.y.data(), s.n);
  if (t.root == NONE) return;
  accumulate(t, s.x.data(), s.y.data(), f.f0.data());
  TravParams p{&t, s.x.data(), s.y.data(), f.f0.data(), s.vx.data(), s.vy.data(),
               f.scalar(S_THETA2), f.scalar(S_DMIN2), f.scalar(S_DMAX2), alpha};
  const uint32_t n = s.n;
  if (g_threads <= 1 || n < 4096) {
    // d3's own order: node 0..n-1, each drawing from random as it goes.
    std::vector<uint32_t> stack;
    Counters* c = g_counting ? &g_count : nullptr;
    for (uint32_t i = 0; i < n; ++i) trav_node(p, false, i, stack, c);
    return;
  }
  // Threaded: record pass in chunks of Morton order, then replay flagged
  // nodes in ascending index with the real random (many_body.zig:65-83).
  std::vector<uint8_t> flags(n, 0);
  const uint32_t CHUNK = 32;  // many_body.zig:25
  g_pool.run(chunks(n, CHUNK), [&](uint32_t c) {
    thread_local std::vector<uint32_t> stack;
    const uint32_t lo = c * CHUNK, hi = std::min(n, lo + CHUNK);
    for (uint32_t pos = lo; pos < hi; ++pos) {
      uint32_t i = pos < t.m ? t.perm[pos] : t.rest[pos - t.m];
      if (!trav_node(p, true, i, stack, nullptr)) flags[i] = 1;
    }
  });
  std::vector<uint32_t> stack;
  for (uint32_t i = 0; i < n; ++i) if (flags[i]) trav_node(p, false, i, stack, nullptr);
}

// ---------------------------------------------------------------- link.zig (exact)
int32_t link_exact(Sim& s, Force& f, double alpha) {
  for (uint32_t e = 0; e < f.count; ++e)
    if (f.source[e] >= s.n || f.target[e] >= s.n) return BAD_ARGUMENT;
  const double iterations = f.scalar(S_ITERATIONS);
  double* xs = s.x.data(); double* ys = s.y.data();
  double* vxs = s.vx.data(); double* vys = s.vy.data();
  for (double k = 0; k < iterations; k += 1) {
    for (uint32_t i = 0; i < f.count; ++i) {
      uint32_t src = f.source[i], tgt = f.target[i];
      double x = xs[tgt] + vxs[tgt] - xs[src] - vxs[src];
      if (x == 0 || x != x) x = jiggle(js_random());
      double y = ys[tgt] + vys[tgt] - ys[src] - vys[src];
      if (y == 0 || y != y) y = jiggle(js_random());
      double l = std::sqrt(x * x + y * y);
      l = (l - f.f1[i]) / l * alpha * f.f0[i];
      x *= l;
      y *= l;
      double b = f.f2[i];
      vxs[tgt] -= x * b;
      vys[tgt] -= y * b;
      b = 1 - b;
      vxs[src] += x * b;
      vys[src] += y * b;
    }
  }
  return OK;
}

// ---------------------------------------------------------------- collide.zig (exact)
// Each iteration: tree on (x + vx, y + vy) (collide.zig:32-33), the radius of
// every cell bottom-up (Prep, collide.zig:104-130: a leaf takes its chain
// head's radius, an internal cell the largest child radius above 0; cells are
// numbered in preorder, so descending index visits children first), then the
// Gauss-Seidel pass over i (collide.zig:45-101).
void collide_exact(Sim& s, Force& f) {
  const double* radii = f.f0.data();
  const double strength = f.scalar(S_STRENGTH);
  const double iterations = f.scalar(S_ITERATIONS);
  const double* xs = s.x.data(); const double* ys = s.y.data();
  double* vxs = s.vx.data(); double* vys = s.vy.data();
  std::vector<uint32_t> stack;
  for (double k = 0; k < iterations; k += 1) {
    s.shift();
    s.tree.build(s.px.data(), s.py.data(), s.n);
    Tree& t = s.tree;
    if (t.root == NONE) continue;
    Cell* cells = t.cells.data();
    const uint32_t* perm = t.perm.data();
    for (uint32_t c = t.len; c-- > 0;) {
      Cell& q = cells[c];
      if (q.leaf()) {
        q.r = radii[perm[q.first + q.count - 1]];
      } else {
        q.r = 0;
        for (uint32_t ch : q.child)
          if (ch != NONE && cells[ch].r > q.r) q.r = cells[ch].r;
      }
    }
    for (uint32_t i = 0; i < s.n; ++i) {
      const double ri = radii[i], ri2 = ri * ri;
      const double xi = xs[i] + vxs[i], yi = ys[i] + vys[i];
      stack.assign(1, t.root);
      while (!stack.empty()) {
        const Cell& q = cells[stack.back()];
        stack.pop_back();
        double rj = q.r, r = ri + rj;
        if (q.leaf()) {
          const uint32_t j = perm[q.first + q.count - 1];
          if (j > i) {
            double x = xi - xs[j] - vxs[j], y = yi - ys[j] - vys[j];
            double l = x * x + y * y;
            if (l < r * r) {
              if (x == 0) { x = jiggle(js_random()); l += x * x; }
              if (y == 0) { y = jiggle(js_random()); l += y * y; }
              l = std::sqrt(l);
              l = (r - l) / l * strength;
              x *= l;
              y *= l;
              rj *= rj;
              r = rj / (ri2 + rj);
              vxs[i] += x * r;
              vys[i] += y * r;
              r = 1 - r;
              vxs[j] -= x * r;
              vys[j] -= y * r;
            }
          }
          continue;
        }
        if (q.x0 > xi + r || q.x1 < xi - r || q.y0 > yi + r || q.y1 < yi - r) continue;
        for (unsigned c = 4; c-- > 0;)
          if (q.child[c] != NONE) stack.push_back(q.child[c]);
      }
    }
  }
}

// ---------------------------------------------------------------- simple.zig
void center(Sim& s, Force& f) {  // simple.zig:14-31
  const uint32_t n = s.n;
  double sx = 0, sy = 0;
  for (uint32_t i = 0; i < n; ++i) { sx += s.x[i]; sy += s.y[i]; }
  const double nf = (double)n;
  const double strength = f.scalar(S_STRENGTH);
  sx = (sx / nf - f.scalar(S_X)) * strength;
  sy = (sy / nf - f.scalar(S_Y)) * strength;
  for (uint32_t i = 0; i < n; ++i) { s.x[i] -= sx; s.y[i] -= sy; }
}

void axis(Sim& s, Force& f, double alpha, int dim) {  // simple.zig:51-62
  const double* pos = dim == 0 ? s.x.data() : s.y.data();
  double* v = dim == 0 ? s.vx.data() : s.vy.data();
  for (uint32_t i = 0; i < s.n; ++i) v[i] += (f.f1[i] - pos[i]) * f.f0[i] * alpha;
}

void radial(Sim& s, Force& f, double alpha) {  // simple.zig:64-86
  const double cx = f.scalar(S_X), cy = f.scalar(S_Y);
  for (uint32_t i = 0; i < s.n; ++i) {
    double dx = s.x[i] - cx;
    if (dx == 0 || dx != dx) dx = 1e-6;
    double dy = s.y[i] - cy;
    if (dy == 0 || dy != dy) dy = 1e-6;
    double l = std::sqrt(dx * dx + dy * dy);
    double k = (f.f1[i] - l) * f.f0[i] * alpha / l;
    s.vx[i] += dx * k;
    s.vy[i] += dy * k;
  }
}

void integrate(Sim& s, double decay) {  // simple.zig:90-123, by chunks on the pool
  const uint32_t size = 16384;
  g_pool.run(chunks(s.n, size), [&](uint32_t c) {
    const uint32_t lo = c * size, hi = std::min(s.n, lo + size);
    for (uint32_t i = lo; i < hi; ++i) {
      uint8_t fixed = s.fixed[i];
      if ((fixed & 1) == 0) { s.vx[i] *= decay; s.x[i] += s.vx[i]; }
      else { s.x[i] = s.fx[i]; s.vx[i] = 0; }
      if ((fixed & 2) == 0) { s.vy[i] *= decay; s.y[i] += s.vy[i]; }
      else { s.y[i] = s.fy[i]; s.vy[i] = 0; }
    }
  });
}

// ---------------------------------------------------------------- link_fast.zig
constexpr uint32_t TARGET_BIT = 0x80000000u;

void build_csr(Force& f, uint32_t n) {  // link_fast.zig:105-130
  const uint32_t m = f.count;
  f.csr_ptr.assign(n + 1, 0);
  f.csr_adj.assign(2 * (size_t)m, 0);
  uint32_t* row = f.csr_ptr.data();
  for (uint32_t e = 0; e < m; ++e) { row[f.source[e] + 1]++; row[f.target[e] + 1]++; }
  for (uint32_t i = 1; i < n + 1; ++i) row[i] += row[i - 1];
  uint32_t* fill = f.csr_adj.data();
  for (uint32_t e = 0; e < m; ++e) {
    uint32_t src = f.source[e], tgt = f.target[e];
    fill[row[src + 1] - 1] = e; row[src + 1]--;
    fill[row[tgt + 1] - 1] = e | TARGET_BIT; row[tgt + 1]--;
  }
  for (uint32_t i = 0; i < n; ++i) row[i] = row[i + 1];
  row[n] = 2 * m;
}

int32_t link_fast(Sim& s, Force& f, double alpha) {  // link_fast.zig:22-99
  const uint32_t n = s.n, m = f.count;
  for (uint32_t e = 0; e < m; ++e)
    if (f.source[e] >= n || f.target[e] >= n) return BAD_ARGUMENT;
  if (f.csr_n != n) { build_csr(f, n); f.csr_n = n; }
  const uint64_t seed = seed_from(f.scalar(S_SEED));
  const double iterations = f.scalar(S_ITERATIONS);
  const uint32_t* row = f.csr_ptr.data();
  const uint32_t* adj = f.csr_adj.data();
  for (double k = 0; k < iterations; k += 1) {
    s.shift();
    for (uint32_t i = 0; i < n; ++i) {
      double vx = s.vx[i], vy = s.vy[i];
      const double pxi = s.px[i], pyi = s.py[i];
      for (uint32_t a = row[i]; a < row[i + 1]; ++a) {
        uint32_t entry = adj[a];
        uint32_t e = entry & ~TARGET_BIT;
        bool is_target = (entry & TARGET_BIT) != 0;
        uint32_t j = is_target ? f.source[e] : f.target[e];
        double x = is_target ? pxi - s.px[j] : s.px[j] - pxi;
        double y = is_target ? pyi - s.py[j] : s.py[j] - pyi;
        if (x == 0 || x != x) x = hash_jiggle(seed, e, 0, 0);
        if (y == 0 || y != y) y = hash_jiggle(seed, e, 0, 1);
        double l = std::sqrt(x * x + y * y);
        l = (l - f.f1[e]) / l * alpha * f.f0[e];
        x *= l;
        y *= l;
        double b = f.f2[e];
        if (is_target) { vx -= x * b; vy -= y * b; }
        else { vx += x * (1 - b); vy += y * (1 - b); }
      }
      s.vx[i] = vx;
      s.vy[i] = vy;
    }
  }
  return OK;
}

// ---------------------------------------------------------------- grid.zig (fast collide)
constexpr uint32_t NO_CELL = 0xffffffffu;

void grid_collide(Sim& s, Force& f) {  // grid.zig:198-286
  const uint32_t n = s.n;
  const double strength = f.scalar(S_STRENGTH);
  const uint64_t seed = seed_from(f.scalar(S_SEED));
  const double* radii = f.f0.data();
  const double iterations = f.scalar(S_ITERATIONS);
  s.g_items.resize(std::max<size_t>(s.g_items.size(), n));
  s.g_keys.resize(std::max<size_t>(s.g_keys.size(), n));
  s.g_cell_of.resize(std::max<size_t>(s.g_cell_of.size(), n));
  for (double k = 0; k < iterations; k += 1) {
    // bounds (grid.zig:66-91), combined as the serial scan.
    double x0 = Inf, y0 = Inf, x1 = -Inf, y1 = -Inf, rmax = 0;
    uint32_t m = 0;
    for (uint32_t i = 0; i < n; ++i) {
      double x = s.x[i] + s.vx[i], y = s.y[i] + s.vy[i];
      s.px[i] = x; s.py[i] = y;
      if (x != x || y != y) continue;
      ++m;
      if (x < x0) x0 = x;
      if (x > x1) x1 = x;
      if (y < y0) y0 = y;
      if (y > y1) y1 = y;
      double rad = radii[i];
      if (rad > rmax) rmax = rad;
    }
    if (m < 2 || !(rmax > 0)) continue;
    double h = 2 * rmax;
    uint32_t gx = (uint32_t)(std::floor((x1 - x0) / h) + 1);
    uint32_t gy = (uint32_t)(std::floor((y1 - y0) / h) + 1);
    while ((uint64_t)gx * gy > 4 * (uint64_t)n + 16) {
      h *= 2;
      gx = (uint32_t)(std::floor((x1 - x0) / h) + 1);
      gy = (uint32_t)(std::floor((y1 - y0) / h) + 1);
    }
    const uint32_t cells = gx * gy;
    s.g_start.resize(std::max<size_t>(s.g_start.size(), (size_t)cells + 1));
    const double inv_h = 1 / h;
    // bin (grid.zig:95-111): position k holds node n-1-k.
    for (uint32_t kk = 0; kk < n; ++kk) {
      uint32_t i = n - 1 - kk;
      uint32_t cell = cells;
      if (s.px[i] == s.px[i] && s.py[i] == s.py[i]) {
        uint32_t cx = (uint32_t)std::floor((s.px[i] - x0) * inv_h);
        uint32_t cy = (uint32_t)std::floor((s.py[i] - y0) * inv_h);
        cell = cy * gx + cx;
      }
      s.g_cell_of[i] = cell == cells ? NO_CELL : cell;
      s.g_keys[kk] = cell;
      s.g_items[kk] = i;
    }
    // stable sort by cell (sort.zig via grid.zig:392)
    {
      std::vector<uint32_t> idx(n);
      for (uint32_t q = 0; q < n; ++q) idx[q] = q;
      std::stable_sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return s.g_keys[a] < s.g_keys[b]; });
      std::vector<uint64_t> k2(n);
      std::vector<uint32_t> i2(n);
      for (uint32_t q = 0; q < n; ++q) { k2[q] = s.g_keys[idx[q]]; i2[q] = s.g_items[idx[q]]; }
      std::copy(k2.begin(), k2.end(), s.g_keys.begin());
      std::copy(i2.begin(), i2.end(), s.g_items.begin());
    }
    // starts (grid.zig:115-127)
    uint32_t* start = s.g_start.data();
    for (uint32_t kk = 0; kk < n; ++kk) {
      uint32_t lo = kk == 0 ? 0 : (uint32_t)(s.g_keys[kk - 1] + 1);
      uint32_t hi = (uint32_t)std::min<uint64_t>(s.g_keys[kk], cells);
      if (kk == 0 || s.g_keys[kk] != s.g_keys[kk - 1]) {
        for (uint32_t cell = lo; cell <= hi; ++cell) start[cell] = kk;
      }
    }
    uint32_t cell = s.g_keys[n - 1] == cells ? cells + 1 : (uint32_t)(s.g_keys[n - 1] + 1);
    for (; cell <= cells; ++cell) start[cell] = m;
    start[cells] = m;
    // sweep (grid.zig:129-195)
    for (uint32_t i = 0; i < n; ++i) {
      uint32_t ci = s.g_cell_of[i];
      if (ci == NO_CELL) continue;
      const double ri = radii[i], ri2 = ri * ri;
      const double xi = s.px[i], yi = s.py[i];
      double vx = s.vx[i], vy = s.vy[i];
      uint32_t cxi = ci % gx, cyi = ci / gx;
      uint32_t xa = cxi > 0 ? cxi - 1 : 0, xb = std::min(cxi + 1, gx - 1);
      uint32_t ya = cyi > 0 ? cyi - 1 : 0, yb = std::min(cyi + 1, gy - 1);
      for (uint32_t cy = ya; cy <= yb; ++cy) {
        for (uint32_t cx = xa; cx <= xb; ++cx) {
          uint32_t c = cy * gx + cx;
          for (uint32_t a = start[c]; a < start[c + 1]; ++a) {
            uint32_t j = s.g_items[a];
            if (j == i) continue;
            double rj = radii[j];
            double rr = ri + rj;
            double x = xi - s.px[j], y = yi - s.py[j];
            double l = x * x + y * y;
            if (l < rr * rr) {
              if (x == 0 || y == 0) {
                uint32_t lo = std::min(i, j), hi = std::max(i, j);
                double sign = i == lo ? 1.0 : -1.0;
                if (x == 0) { x = sign * hash_jiggle(seed, lo, hi, 0); l += x * x; }
                if (y == 0) { y = sign * hash_jiggle(seed, lo, hi, 1); l += y * y; }
              }
              l = std::sqrt(l);
              l = (rr - l) / l * strength;
              x *= l;
              y *= l;
              rr = rj * rj / (ri2 + rj * rj);
              vx += x * rr;
              vy += y * rr;
            }
          }
        }
      }
      s.vx[i] = vx;
      s.vy[i] = vy;
    }
  }
}

// ---------------------------------------------------------------- fmm.zig (fast many-body)
constexpr uint32_t MAX_ORDER = 24, DEFAULT_ORDER = 8;
constexpr double DEFAULT_OPENING = 1.4;
constexpr uint32_t MIN_N = 2000, LEAF = 32, LEVELS = 16;
constexpr double MIN_SEP2 = 1e-28;

struct FCell { double cx, cy, r; uint32_t first, count, child0, nchild, depth, pad; };  // fmm.zig:68-78

struct Fmm {
  std::vector<FCell> cells;
  std::vector<double> ux, uy, px, py, ps, bx0, by0, bx1, by1, m, l, binom;
  size_t bstride = 0;
  double fact[2 * MAX_ORDER + 2];
  uint32_t p = DEFAULT_ORDER;
  size_t stride = 0;
  const uint32_t* perm = nullptr;
  double *vx = nullptr, *vy = nullptr;
  double inv_w = 0, mac_far = 0, dmax_scaled = 0, dmin_scaled = 0, dmin2 = 0, dmax2 = 0;
  bool ring = false, has_dmax = false;
  uint64_t seed = 0;
  double dre[MAX_ORDER + 1], dim[MAX_ORDER + 1];
  uint64_t n_m2l = 0, n_p2p_pairs = 0;

  void binomials(uint32_t pp) {  // fmm.zig:223-242
    size_t nmax = 2 * pp + 2;
    bstride = nmax;
    binom.assign(nmax * nmax, 0.0);
    for (size_t n = 0; n < nmax; ++n) {
      binom[n * nmax] = 1;
      for (size_t k = 1; k < n + 1; ++k)
        binom[n * nmax + k] = binom[(n - 1) * nmax + k - 1] + (k < n ? binom[(n - 1) * nmax + k] : 0.0);
    }
    fact[0] = 1;
    for (size_t k = 1; k < nmax; ++k) fact[k] = fact[k - 1] * (double)k;
  }
  double bin(size_t n, size_t k) const { return binom[n * bstride + k]; }

  void shift_powers(double re, double im) {  // fmm.zig:614-623
    dre[0] = 1; dim[0] = 0;
    for (uint32_t k = 1; k < p + 1; ++k) {
      dre[k] = dre[k - 1] * re - dim[k - 1] * im;
      dim[k] = dre[k - 1] * im + dim[k - 1] * re;
    }
  }

  // d3's pair formula, fmm.zig:529-553
  void pair(uint32_t ti, uint32_t tj, double xi, double yi, double& vxv, double& vyv) {
    if (tj == ti) return;
    double sj = ps[tj];
    double x = px[tj] - xi, y = py[tj] - yi;
    double ll = x * x + y * y;
    if (ll >= dmax2) return;
    if (x == 0 || y == 0) {
      uint32_t i = perm[ti], j = perm[tj];
      if (x == 0) { x = hash_jiggle(seed, i, j, 0); ll += x * x; }
      if (y == 0) { y = hash_jiggle(seed, i, j, 1); ll += y * y; }
    }
    if (ll < dmin2) ll = std::sqrt(dmin2 * ll);
    double wj = sj / ll;
    vxv += x * wj;
    vyv += y * wj;
  }

  // fmm.zig:479-526: four sources per step on two f64 lanes.
  void target(bool has_dm, uint32_t ti, uint32_t b0, uint32_t b1) {
    uint32_t i = perm[ti];
    double xi = px[ti], yi = py[ti];
    double vxv = vx[i], vyv = vy[i];
    double ax0 = 0, ax1 = 0, ay0 = 0, ay1 = 0;
    uint32_t k = b0;
    while (k + 4 <= b1) {
      double sa0 = ps[k], sa1 = ps[k + 1], sb0 = ps[k + 2], sb1 = ps[k + 3];
      double xa0 = px[k] - xi, xa1 = px[k + 1] - xi;
      double ya0 = py[k] - yi, ya1 = py[k + 1] - yi;
      double xb0 = px[k + 2] - xi, xb1 = px[k + 3] - xi;
      double yb0 = py[k + 2] - yi, yb1 = py[k + 3] - yi;
      double la0 = xa0 * xa0 + ya0 * ya0, la1 = xa1 * xa1 + ya1 * ya1;
      double lb0 = xb0 * xb0 + yb0 * yb0, lb1 = xb1 * xb1 + yb1 * yb1;
      bool zero = (xa0 * ya0 * xb0 * yb0 == 0) || (xa1 * ya1 * xb1 * yb1 == 0);
      double mn = zmin(zmin(la0, lb0), zmin(la1, lb1));
      if (zero || !(mn >= dmin2) || (ti >= k && ti < k + 4)) {
        for (uint32_t q = 0; q < 4; ++q) pair(ti, k + q, xi, yi, vxv, vyv);
        k += 4;
        continue;
      }
      double pa0 = la0 * la1, pa1 = la1 * la0;
      double pb0 = lb0 * lb1, pb1 = lb1 * lb0;
      double inv0 = 1 / (pa0 * pb0), inv1 = 1 / (pa1 * pb1);
      double wa0 = sa0 * la1 * pb0 * inv0, wa1 = sa1 * la0 * pb1 * inv1;
      double wb0 = sb0 * lb1 * pa0 * inv0, wb1 = sb1 * lb0 * pa1 * inv1;
      if (has_dm) {
        wa0 = la0 < dmax2 ? wa0 : 0; wa1 = la1 < dmax2 ? wa1 : 0;
        wb0 = lb0 < dmax2 ? wb0 : 0; wb1 = lb1 < dmax2 ? wb1 : 0;
      }
      ax0 += xa0 * wa0 + xb0 * wb0; ax1 += xa1 * wa1 + xb1 * wb1;
      ay0 += ya0 * wa0 + yb0 * wb0; ay1 += ya1 * wa1 + yb1 * wb1;
      k += 4;
    }
    while (k < b1) { pair(ti, k, xi, yi, vxv, vyv); ++k; }
    vx[i] = vxv + (ax0 + ax1);
    vy[i] = vyv + (ay0 + ay1);
  }

  void p2p(const FCell& ca, const FCell& cb) {  // fmm.zig:457-475
    n_p2p_pairs += (uint64_t)ca.count * cb.count;
    for (uint32_t ti = ca.first; ti < ca.first + ca.count; ++ti) target(has_dmax, ti, cb.first, cb.first + cb.count);
  }

  // fmm.zig:561-610
  void m2l(uint32_t A, uint32_t B, double dx, double dy) {
    ++n_m2l;
    const uint32_t q = p + 1;
    double d2 = dx * dx + dy * dy;
    double ire = dx / d2, iim = -dy / d2;
    double tre[2 * MAX_ORDER + 1], tim[2 * MAX_ORDER + 1];
    double are = ire, aim = iim;
    tre[0] = are; tim[0] = aim;
    for (uint32_t mm = 1; mm < 2 * p + 1; ++mm) {
      double nre = are * ire - aim * iim;
      double nim = are * iim + aim * ire;
      are = nre; aim = nim;
      tre[mm] = are * fact[mm];
      tim[mm] = aim * fact[mm];
    }
    const double* mre = &m[B * stride];
    const double* mim = &m[B * stride + q];
    double* lre = &l[A * stride];
    double* lim = &l[A * stride + q];
    for (uint32_t ll = 0; ll < q; ++ll) {
      double vre0 = 0, vre1 = 0, vim0 = 0, vim1 = 0;
      uint32_t k = 0;
      while (k + 2 <= q) {
        double a0 = mre[k], a1 = mre[k + 1], b0 = mim[k], b1 = mim[k + 1];
        double tr0 = tre[ll + k], tr1 = tre[ll + k + 1], ti0 = tim[ll + k], ti1 = tim[ll + k + 1];
        vre0 += a0 * tr0 - b0 * ti0; vre1 += a1 * tr1 - b1 * ti1;
        vim0 += a0 * ti0 + b0 * tr0; vim1 += a1 * ti1 + b1 * tr1;
        k += 2;
      }
      double sre = vre0 + vre1, sim = vim0 + vim1;
      if (k < q) {
        sre += mre[k] * tre[ll + k] - mim[k] * tim[ll + k];
        sim += mre[k] * tim[ll + k] + mim[k] * tre[ll + k];
      }
      lre[ll] += sre;
      lim[ll] += sim;
    }
  }

  // fmm.zig:400-453, the serial recursion (the reference's top traversal and
  // job queues reproduce it exactly, fmm.zig:12-21).
  void interact(uint32_t A, uint32_t B) {
    const FCell& ca = cells[A];
    if (A == B) {
      if (ca.nchild == 0) {
        p2p(ca, ca);
      } else {
        for (uint32_t i = 0; i < ca.nchild; ++i)
          for (uint32_t j = 0; j < ca.nchild; ++j) interact(ca.child0 + i, ca.child0 + j);
      }
      return;
    }
    const FCell& cb = cells[B];
    const double ra = ca.r, rb = cb.r;
    const double dx = ca.cx - cb.cx, dy = ca.cy - cb.cy;
    const double d2 = dx * dx + dy * dy;
    const double rmax = zmax(ra, rb), rmin = zmin(ra, rb);
    const double thr = mac_far * rmax + rmin;
    if (d2 > MIN_SEP2 && d2 >= thr * thr) {
      bool separated = true;
      if (ring) {
        double gap = std::sqrt(d2) - ra - rb;
        if (gap >= dmax_scaled) return;
        if (gap + 2 * (ra + rb) > dmax_scaled || gap < dmin_scaled) separated = false;
      }
      if (separated) { m2l(A, B, dx, dy); return; }
    }
    const bool la = ca.nchild == 0, lb = cb.nchild == 0;
    if (la && lb) {
      p2p(ca, cb);
    } else if (la || (!lb && rb > ra)) {
      for (uint32_t j = 0; j < cb.nchild; ++j) interact(A, cb.child0 + j);
    } else {
      for (uint32_t i = 0; i < ca.nchild; ++i) interact(ca.child0 + i, B);
    }
  }

  void geometry_cell(uint32_t i) {  // fmm.zig:721-754
    FCell& cell = cells[i];
    double x0 = Inf, y0 = Inf, x1 = -Inf, y1 = -Inf;
    if (cell.nchild == 0) {
      for (uint32_t k = cell.first; k < cell.first + cell.count; ++k) {
        x0 = zmin(x0, ux[k]); x1 = zmax(x1, ux[k]);
        y0 = zmin(y0, uy[k]); y1 = zmax(y1, uy[k]);
      }
    } else {
      for (uint32_t ch = cell.child0; ch < cell.child0 + cell.nchild; ++ch) {
        x0 = zmin(x0, bx0[ch]); x1 = zmax(x1, bx1[ch]);
        y0 = zmin(y0, by0[ch]); y1 = zmax(y1, by1[ch]);
      }
    }
    bx0[i] = x0; by0[i] = y0; bx1[i] = x1; by1[i] = y1;
    cell.cx = (x0 + x1) / 2;
    cell.cy = (y0 + y1) / 2;
    double dx = x1 - x0, dy = y1 - y0;
    cell.r = 0.5 * std::sqrt(dx * dx + dy * dy);
  }

  void upward_cell(uint32_t i) {  // fmm.zig:757-806
    const FCell& cell = cells[i];
    const uint32_t q = p + 1;
    double* mre = &m[i * stride];
    double* mim = &m[i * stride + q];
    if (cell.nchild == 0) {
      for (uint32_t k = cell.first; k < cell.first + cell.count; ++k) {
        double sv = ps[k];
        if (sv != sv) continue;
        double zre = ux[k] - cell.cx, zim = uy[k] - cell.cy;
        double pre = 1, pim = 0;
        mre[0] += sv;
        for (uint32_t o = 1; o < q; ++o) {
          double nre = pre * zre - pim * zim;
          double nim = pre * zim + pim * zre;
          pre = nre; pim = nim;
          mre[o] += sv * pre;
          mim[o] += sv * pim;
        }
      }
    } else {
      for (uint32_t ch = cell.child0; ch < cell.child0 + cell.nchild; ++ch) {
        const FCell& c = cells[ch];
        shift_powers(c.cx - cell.cx, c.cy - cell.cy);
        const double* cre = &m[ch * stride];
        const double* cim = &m[ch * stride + q];
        for (uint32_t nn = 0; nn < q; ++nn) {
          double sre = 0, sim = 0;
          for (uint32_t k = 0; k < nn + 1; ++k) {
            double b = bin(nn, k);
            double xre = cre[k] * b, xim = cim[k] * b;
            sre += xre * dre[nn - k] - xim * dim[nn - k];
            sim += xre * dim[nn - k] + xim * dre[nn - k];
          }
          mre[nn] += sre;
          mim[nn] += sim;
        }
      }
    }
  }

  void l2l(uint32_t i, uint32_t ch) {  // fmm.zig:823-849
    const uint32_t q = p + 1;
    const FCell& cell = cells[i];
    const double* lre = &l[i * stride];
    const double* lim = &l[i * stride + q];
    const FCell& c = cells[ch];
    shift_powers(c.cx - cell.cx, c.cy - cell.cy);
    double* cre = &l[ch * stride];
    double* cim = &l[ch * stride + q];
    for (uint32_t mm = 0; mm < q; ++mm) {
      double sre = 0, sim = 0;
      for (uint32_t ll = mm; ll < q; ++ll) {
        double b = bin(ll, mm);
        double xre = lre[ll] * b, xim = lim[ll] * b;
        sre += xre * dre[ll - mm] - xim * dim[ll - mm];
        sim += xre * dim[ll - mm] + xim * dre[ll - mm];
      }
      cre[mm] += sre;
      cim[mm] += sim;
    }
  }

  void downward_cell(uint32_t i) {  // fmm.zig:853-886
    const FCell& cell = cells[i];
    const uint32_t q = p + 1;
    const double* lre = &l[i * stride];
    const double* lim = &l[i * stride + q];
    if (cell.nchild == 0) {
      for (uint32_t k = cell.first; k < cell.first + cell.count; ++k) {
        uint32_t j = perm[k];
        double wre = ux[k] - cell.cx, wim = uy[k] - cell.cy;
        double are = lre[p], aim = lim[p];
        for (uint32_t ll = p; ll > 0;) {
          --ll;
          double nre = are * wre - aim * wim + lre[ll];
          double nim = are * wim + aim * wre + lim[ll];
          are = nre; aim = nim;
        }
        vx[j] -= are * inv_w;
        vy[j] += aim * inv_w;
      }
    } else {
      for (uint32_t ch = cell.child0; ch < cell.child0 + cell.nchild; ++ch) l2l(i, ch);
    }
  }
};

Sim::~Sim() { delete fmm; for (Force* f : forces) delete f; }

// fmm.zig:685-711
void build_cells(Fmm& fm, const Tree& t) {
  const uint32_t m = t.m;
  fm.cells.clear();
  fm.cells.reserve(std::max<uint32_t>(64, 2 * (m / LEAF + 1)));
  fm.cells.push_back(FCell{0, 0, 0, 0, m, 0, 0, 0, 0});
  for (uint32_t c = 0; c < fm.cells.size(); ++c) {
    const uint32_t d = fm.cells[c].depth, lo = fm.cells[c].first, hi = lo + fm.cells[c].count;
    if (fm.cells[c].count <= LEAF || d >= LEVELS) continue;
    const unsigned shift = 2 * (LEVELS - 1 - d);
    uint32_t k = lo;
    const uint32_t c0 = (uint32_t)fm.cells.size();
    uint32_t nc = 0;
    for (uint32_t q = 0; q < 4; ++q) {
      uint32_t st = k;
      while (k < hi && ((t.keys[k] >> shift) & 3) == q) ++k;
      if (k > st) { fm.cells.push_back(FCell{0, 0, 0, st, k - st, 0, 0, d + 1, 0}); ++nc; }
    }
    fm.cells[c].child0 = c0;
    fm.cells[c].nchild = nc;
  }
}

// fmm.zig:247-354. Returns false when the extent is degenerate.
bool fmm_run(Sim& s, Force& f, double alpha) {
  Tree& t = s.tree;
  const uint32_t n = s.n;
  if (!t.prepare_fast(s.x.data(), s.y.data(), n)) return true;
  const double* e = t.extent;
  const double w = e[2] - e[0];
  if (!(w > 0 && w < Inf)) return false;
  uint32_t p = DEFAULT_ORDER;
  double po = f.scalar(S_ORDER);
  if (po >= 1 && po <= (double)MAX_ORDER) p = (uint32_t)po;
  double kappa = f.scalar(S_OPENING);
  if (!(kappa > 1)) kappa = DEFAULT_OPENING;
  if (!s.fmm) s.fmm = new Fmm();
  Fmm& fm = *s.fmm;
  fm.p = p;
  fm.binomials(p);
  build_cells(fm, t);
  const uint32_t nc = (uint32_t)fm.cells.size();
  fm.stride = 2 * (p + 1);
  fm.m.assign(nc * fm.stride, 0.0);
  fm.l.assign(nc * fm.stride, 0.0);
  fm.bx0.assign(nc, 0); fm.by0.assign(nc, 0); fm.bx1.assign(nc, 0); fm.by1.assign(nc, 0);
  double dmax_scaled = Inf;
  const double dmax2 = f.scalar(S_DMAX2);
  if (dmax2 < Inf) dmax_scaled = std::sqrt(dmax2) / w;
  const double dmin2 = f.scalar(S_DMIN2);
  const double dmin_scaled = dmin2 > 0 ? std::sqrt(dmin2) / w : 0;
  const double inv_w = 1 / w;
  fm.perm = t.perm.data();
  fm.vx = s.vx.data(); fm.vy = s.vy.data();
  fm.inv_w = inv_w;
  fm.mac_far = 2 * kappa - 1;
  fm.ring = dmin_scaled > 0 || dmax_scaled < Inf;
  fm.has_dmax = dmax_scaled < Inf;
  fm.dmax_scaled = dmax_scaled;
  fm.dmin_scaled = dmin_scaled;
  fm.dmin2 = dmin2;
  fm.dmax2 = dmax2;
  fm.seed = seed_from(f.scalar(S_SEED));
  const double ox = (e[0] + e[2]) / 2, oy = (e[1] + e[3]) / 2;
  // gather (fmm.zig:668-682)
  const uint32_t m = t.m;
  fm.ux.resize(m); fm.uy.resize(m); fm.px.resize(m); fm.py.resize(m); fm.ps.resize(m);
  for (uint32_t j = 0; j < m; ++j) {
    uint32_t i = t.perm[j];
    double x = s.x[i], y = s.y[i];
    fm.px[j] = x; fm.py[j] = y;
    double st = f.f0[i] * alpha;
    fm.ps[j] = st != st ? 0 : st;
    fm.ux[j] = (x - ox) * inv_w;
    fm.uy[j] = (y - oy) * inv_w;
  }
  // upward, children before parents (level order reversed)
  for (uint32_t c = nc; c-- > 0;) { fm.geometry_cell(c); fm.upward_cell(c); }
  // scale M_k by 1/k! (fmm.zig:629-642)
  const uint32_t q = p + 1;
  for (uint32_t i = 0; i < nc; ++i) {
    double* mre = &fm.m[i * fm.stride];
    double* mim = &fm.m[i * fm.stride + q];
    for (uint32_t o = 2; o < q; ++o) {
      double inv = 1 / fm.fact[o];
      mre[o] *= inv;
      mim[o] *= inv;
    }
  }
  fm.interact(0, 0);
  // unscale L (fmm.zig:645-657)
  for (uint32_t i = 0; i < nc; ++i) {
    double* lre = &fm.l[i * fm.stride];
    double* lim = &fm.l[i * fm.stride + q];
    double sign = 1;
    for (uint32_t o = 0; o < q; ++o) {
      double inv = sign / fm.fact[o];
      lre[o] *= inv;
      lim[o] *= inv;
      sign = -sign;
    }
  }
  // downward in level order, parents before children
  for (uint32_t i = 0; i < nc; ++i) fm.downward_cell(i);
  return true;
}

// ---------------------------------------------------------------- handles (root.zig)
std::vector<Sim*> g_sims;

Sim* get(uint32_t h) {
  if (h == 0 || h > g_sims.size()) return nullptr;
  return g_sims[h - 1];
}
Force* get_force(Sim* s, uint32_t fh) {
  if (!s || fh == 0 || fh > s->forces.size()) return nullptr;
  return s->forces[fh - 1];
}
template <class T> void* addr(std::vector<T>& v) { return v.empty() ? nullptr : (void*)v.data(); }

bool fast_mode(const Sim& s, const Force& f) {  // root.zig:246-248
  return f.scalar(S_PRECISION) == 1 && s.n >= MIN_N;
}

}  // namespace

// ---------------------------------------------------------------- exported ABI
EXPORT void oracle_core_set_random(RandomFn fn, void* ctx) { g_random = fn; g_random_ctx = ctx; }

EXPORT uint32_t oracle_sim_create() {  // root.zig:65-78
  Sim* s = new (std::nothrow) Sim();
  if (!s) return 0;
  for (size_t i = 0; i < g_sims.size(); ++i) if (!g_sims[i]) { g_sims[i] = s; return (uint32_t)i + 1; }
  g_sims.push_back(s);
  return (uint32_t)g_sims.size();
}

EXPORT int32_t oracle_sim_destroy(uint32_t h) {  // root.zig:80-86
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  delete s;
  g_sims[h - 1] = nullptr;
  return OK;
}

EXPORT int32_t oracle_sim_set_node_count(uint32_t h, uint32_t n) {  // root.zig:88-92
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  return s->set_node_count(n) ? GROWN : OK;
}

EXPORT void* oracle_sim_ptr(uint32_t h, uint32_t which) {  // root.zig:94-98, sim.zig:75-88
  Sim* s = get(h);
  if (!s) return nullptr;
  switch (which) {
    case 0: return addr(s->x);
    case 1: return addr(s->y);
    case 2: return addr(s->vx);
    case 3: return addr(s->vy);
    case 4: return addr(s->fx);
    case 5: return addr(s->fy);
    case 6: return addr(s->fixed);
    case 7: return addr(s->tree.perm);
    case 8: return nullptr;
    case 9: return addr(s->snap);
    default: return nullptr;
  }
}

EXPORT int32_t oracle_sim_sync(uint32_t h, uint32_t) { return get(h) ? OK : BAD_HANDLE; }

EXPORT int32_t oracle_tree_build(uint32_t h, uint32_t shifted) {  // root.zig:102-108
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const Cell* c0 = s->tree.cells.data();
  const uint32_t* p0 = s->tree.perm.data();
  if (shifted) {
    s->shift();
    s->tree.build(s->px.data(), s->py.data(), s->n);
  } else {
    s->tree.build(s->x.data(), s->y.data(), s->n);
  }
  return (s->tree.cells.data() != c0 || s->tree.perm.data() != p0) ? GROWN : OK;
}

EXPORT uint64_t oracle_tree_info(uint32_t h, uint32_t which) {  // root.zig:113-126
  Sim* s = get(h);
  if (!s) return 0;
  Tree& t = s->tree;
  switch (which) {
    case 0: return t.len;
    case 1: return t.root;
    case 2: return (uint64_t)(uintptr_t)t.cells.data();
    case 3: return (uint64_t)(uintptr_t)t.extent;
    case 4: return (uint64_t)(uintptr_t)t.perm.data();
    case 5: return sizeof(Cell);
    case 6: return t.m;
    default: return 0;
  }
}

EXPORT int32_t oracle_tree_order(uint32_t h, uint32_t after, uint32_t* out) {  // root.zig:130-137
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  if (!out) return BAD_ARGUMENT;
  s->tree.order(after != 0, out);
  return OK;
}

EXPORT uint32_t oracle_force_create(uint32_t h, uint32_t kind) {  // root.zig:147-161
  Sim* s = get(h);
  if (!s || kind >= K_COUNT) return 0;
  Force* f = new (std::nothrow) Force((Kind)kind);
  if (!f) return 0;
  if (f->node_sized()) f->set_count(s->n);
  for (size_t i = 0; i < s->forces.size(); ++i) if (!s->forces[i]) { s->forces[i] = f; return (uint32_t)i + 1; }
  s->forces.push_back(f);
  return (uint32_t)s->forces.size();
}

EXPORT int32_t oracle_force_destroy(uint32_t h, uint32_t fh) {  // root.zig:163-170
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  delete f;
  s->forces[fh - 1] = nullptr;
  return OK;
}

EXPORT int32_t oracle_force_set_count(uint32_t h, uint32_t fh, uint32_t count) {  // root.zig:174-179
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  f->set_count(count);
  return GROWN;
}

EXPORT void* oracle_force_ptr(uint32_t h, uint32_t fh, uint32_t which) {  // root.zig:181-186
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return nullptr;
  switch (which) {
    case A_F0: return addr(f->f0);
    case A_F1: return addr(f->f1);
    case A_F2: return addr(f->f2);
    case A_SOURCE: return addr(f->source);
    case A_TARGET: return addr(f->target);
    default: return nullptr;
  }
}

EXPORT int32_t oracle_force_set_scalar(uint32_t h, uint32_t fh, uint32_t which, double v) {  // root.zig:188-194
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (which >= S_COUNT) return BAD_ARGUMENT;
  f->scalars[which] = v;
  return OK;
}

EXPORT double oracle_force_get_scalar(uint32_t h, uint32_t fh, uint32_t which) {  // root.zig:196-200
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return 0;
  return which < S_COUNT ? f->scalars[which] : 0;
}

EXPORT int32_t oracle_sim_snapshot(uint32_t h, uint32_t positions) {  // root.zig:205-209, sim.zig:91-100
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const uint32_t n = s->n;
  if (s->snap.size() < 4 * (size_t)n) s->snap.resize(4 * (size_t)s->cap);
  std::copy(s->vx.begin(), s->vx.begin() + n, s->snap.begin());
  std::copy(s->vy.begin(), s->vy.begin() + n, s->snap.begin() + n);
  if (positions) {
    std::copy(s->x.begin(), s->x.begin() + n, s->snap.begin() + 2 * n);
    std::copy(s->y.begin(), s->y.begin() + n, s->snap.begin() + 3 * n);
  }
  return OK;
}

EXPORT int32_t oracle_force_run(uint32_t h, uint32_t fh, double alpha) {  // root.zig:211-242
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (f->node_sized() && f->count < s->n) return BAD_ARGUMENT;
  switch (f->kind) {
    case K_CENTER: center(*s, *f); return OK;
    case K_X: axis(*s, *f, alpha, 0); return OK;
    case K_Y: axis(*s, *f, alpha, 1); return OK;
    case K_RADIAL: radial(*s, *f, alpha); return OK;
    case K_LINK:
      return fast_mode(*s, *f) ? link_fast(*s, *f, alpha) : link_exact(*s, *f, alpha);
    case K_MANY_BODY:
      if (fast_mode(*s, *f) && fmm_run(*s, *f, alpha)) return OK;
      many_body_exact(*s, *f, alpha);
      return OK;
    case K_COLLIDE:
      if (fast_mode(*s, *f)) { grid_collide(*s, *f); return OK; }
      collide_exact(*s, *f);
      return OK;
    default: return BAD_ARGUMENT;
  }
}

EXPORT int32_t oracle_sim_integrate(uint32_t h, double velocity_decay) {  // root.zig:250-254
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  integrate(*s, velocity_decay);
  return OK;
}

// --- oracle-only instrumentation (not part of the product ABI) ---
// Traversal counters for the roofline's algorithmic work: non-empty cell
// visits, accepted far-field cells, leaf pair interactions.
EXPORT void oracle_counters(int on, uint64_t* out) {
  if (out) { out[0] = g_count.visits; out[1] = g_count.accepts; out[2] = g_count.pairs; }
  g_counting = on != 0;
  g_count = Counters();
}
EXPORT void oracle_set_threads(int k) { g_threads = k < 1 ? 1 : k; }
EXPORT void oracle_fmm_stats(uint32_t h, uint64_t* out) {
  Sim* s = get(h);
  if (!s || !s->fmm) { out[0] = out[1] = out[2] = 0; return; }
  out[0] = s->fmm->n_m2l; out[1] = s->fmm->n_p2p_pairs; out[2] = s->fmm->cells.size();
  s->fmm->n_m2l = s->fmm->n_p2p_pairs = 0;
}
// The sorted Morton keys of the last exact tree build (for parity of the GPU sort).
EXPORT void* oracle_tree_keys(uint32_t h) {
  Sim* s = get(h);
  return s ? (void*)s->tree.keys.data() : nullptr;
}
