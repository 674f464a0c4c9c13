// Host runtime and C ABI of the d3fgpu library (include/d3fgpu.h): the
// replacement for the reference core's exported ABI (src/core/root.zig) with
// the simulation state resident in HBM.
//
// Ownership: each simulation owns pinned host arrays (what sim_ptr returns;
// the facade reads and writes them) and their device mirrors; sim_sync moves
// node arrays between the two. Each force owns pinned parameter arrays and
// device copies, uploaded on the first force_run after the facade could have
// written them (any force_ptr or force_set_count call).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/d3fgpu.h"
#include "buffers.h"
#include "common.h"

#define EXPORT extern "C" __attribute__((visibility("default")))

namespace d3f {

namespace {
thread_local char g_noted[256];
thread_local bool g_has_note = false;
thread_local char g_last_error[256];
}  // namespace

void note_failure(hipError_t e, const char* what, const char* file, int line) {
  if (g_has_note) return;  // the first failure is the cause; its callers pass it on
  const char* base = std::strrchr(file, '/');
  std::snprintf(g_noted, sizeof g_noted, "%s at %s:%d: %s", hipGetErrorName(e), base ? base + 1 : file, line, what);
  g_has_note = true;
}

void publish_failure(hipError_t e) {
  if (e != hipSuccess) {
    if (g_has_note)
      std::memcpy(g_last_error, g_noted, sizeof g_last_error);
    else
      std::snprintf(g_last_error, sizeof g_last_error, "%s", hipGetErrorName(e));
  }
  g_has_note = false;
}

namespace {

enum Status : int32_t { OK = 0, GROWN = 1, BAD_HANDLE = -1, OUT_OF_MEMORY = -2, BAD_ARGUMENT = -3, WORKER_FAILED = -4 };
enum Kind : uint32_t { K_CENTER, K_COLLIDE, K_LINK, K_MANY_BODY, K_RADIAL, K_X, K_Y, K_COUNT };
enum Scalar : uint32_t { S_STRENGTH, S_X, S_Y, S_THETA2, S_DMIN2, S_DMAX2, S_ITERATIONS, S_PRECISION, S_ORDER,
                         S_OPENING, S_SEED, S_COUNT };
constexpr uint32_t MIN_N = 2000;  // fmm.zig:43, the fast-mode crossover

d3f_random_fn g_random = nullptr;
void* g_random_ctx = nullptr;
bool g_timing = false;

double draw() { return g_random ? g_random(g_random_ctx) : 0.5; }
double jiggle_host(double r) { return (r - 0.5) * 1e-6; }  // rng.zig:7-11

int32_t status_of(hipError_t e) {
  publish_failure(e);
  if (e == hipSuccess) return OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return OUT_OF_MEMORY;
  return WORKER_FAILED;
}

// Failure injection for tests of the failover path: after g_fail_after more
// steps (force_run, sim_integrate), the next g_fail_count steps return
// g_fail_status before any device work.
uint32_t g_fail_after = 0, g_fail_count = 0;
int32_t g_fail_status = WORKER_FAILED;
bool injected_failure(int32_t* status) {
  if (g_fail_count == 0) return false;
  if (g_fail_after > 0) {
    --g_fail_after;
    return false;
  }
  --g_fail_count;
  std::snprintf(g_last_error, sizeof g_last_error, "injected failure (core_fail_next): no device work was done");
  *status = g_fail_status;
  return true;
}

struct Force {
  Kind kind;
  double scalars[S_COUNT] = {0};
  uint32_t count = 0;
  Pinned<double> f[3];
  Pinned<uint32_t> source, target;
  Device<double> df[3];
  Device<uint32_t> dsrc, dtgt;
  bool dirty = true;
  // exact link schedule, level-major
  bool sched_valid = false;
  double collide_scale = 0.5;  // exact collide: box scale of the last parallel pass
  uint32_t sched_n = 0;
  std::vector<uint32_t> level_off;
  Device<uint32_t> lsrc, ltgt;
  Device<double> lf[3];
  // fast link adjacency rows (link_fast.zig:105-130)
  bool csr_valid = false;
  uint32_t csr_n = 0;
  Device<uint32_t> row, adj, owner, rows_long;
  Device<LinkEntry> entries;
  Device<double> terms;
  uint32_t long_count = 0;
  LinkCsrArgs csr_args(uint32_t n) const {
    LinkCsrArgs a;
    a.src = dsrc.p;
    a.tgt = dtgt.p;
    a.strength = df[0].p;
    a.distance = df[1].p;
    a.bias = df[2].p;
    a.m = count;
    a.n = n;
    a.row = row.p;
    a.adj = adj.p;
    a.owner = owner.p;
    a.rows_long = rows_long.p;
    a.entries = entries.p;
    a.terms = terms.p;
    return a;
  }

  explicit Force(Kind k) : kind(k) {}
  ~Force() {
    for (auto& a : f) a.release();
    source.release();
    target.release();
    for (auto& a : df) a.release();
    dsrc.release();
    dtgt.release();
    lsrc.release();
    ltgt.release();
    for (auto& a : lf) a.release();
    row.release();
    adj.release();
    owner.release();
    rows_long.release();
    entries.release();
    terms.release();
  }
  bool node_sized() const { return kind != K_LINK && kind != K_CENTER; }  // force.zig:50-52
  int float_arrays() const {                                              // force.zig:54-61
    switch (kind) {
      case K_CENTER: return 0;
      case K_COLLIDE: case K_MANY_BODY: return 1;
      case K_RADIAL: case K_X: case K_Y: return 2;
      default: return 3;
    }
  }
  // force.zig:65-82: new float slots NaN, new index slots 0.
  hipError_t set_count(uint32_t c) {
    const int k = float_arrays();
    for (int i = 0; i < k; ++i) {
      HIPCHECK(f[i].reserve(c > 0 ? c : 1, count));
      for (uint32_t j = count; j < c; ++j) f[i].p[j] = __builtin_nan("");
    }
    if (kind == K_LINK) {
      HIPCHECK(source.reserve(c > 0 ? c : 1, count));
      HIPCHECK(target.reserve(c > 0 ? c : 1, count));
      for (uint32_t j = count; j < c; ++j) source.p[j] = target.p[j] = 0;
    }
    count = c;
    dirty = true;
    sched_valid = false;
    csr_valid = false;
    return hipSuccess;
  }
  double scalar(Scalar s) const { return scalars[s]; }
};

struct Sim {
  uint32_t n = 0, cap = 0;
  Pinned<double> x, y, vx, vy, fx, fy, snap;
  Device<double> dsnap;     // sim_snapshot's copy (vx, vy, x, y), brought to `snap` on demand
  bool snap_pending = false;  // dsnap is newer than snap
  bool snap_positions = false;
  Pinned<uint8_t> fixed;
  Device<double> dx, dy, dvx, dvy, dfx, dfy, dpx, dpy;
  Device<double2> dpos;      // x + vx, y + vy as pairs (fast link sweep)
  Device<LinkNode> dlnodes;  // the exact link pass's packed node state
  Device<uint8_t> dfixed;
  TreeBuffers tree;
  FmmBuffers fmm;
  // fast-mode scratch: radix sort ping-pong, grid of the fast collide
  Device<uint64_t> sk0, sk1;
  Device<uint32_t> sv0, sv1, shist, sglob, scount;
  Device<uint32_t> gstart;
  Device<double4> gcand;
  Device<uint8_t> gpartials;
  GridParams* gparams = nullptr;
  hipStream_t stream = nullptr;
  std::vector<Force*> forces;
  // scratch
  Device<uint32_t> flagged, order, draws_count, draw_off;
  Device<double> draws;
  uint32_t* dscalar_u = nullptr;   // [0] flagged count, [1] link zero flag
  double* dscalar_d = nullptr;     // [0..1] center sums computed on the main stream
  // forceCenter's sums of x and y depend on x, y alone, and only center,
  // integrate and an H2D sync write them. So after each of those writes (when
  // a center force exists) the sums for the next center run are started on a
  // side stream, where they overlap the many-body and link work of the tick.
  hipStream_t side = nullptr;
  hipStream_t aux = nullptr;  // a force's own concurrent work (the FMM's multipoles)
  hipEvent_t ev_x_ready = nullptr, ev_sums_done = nullptr;
  double* dsums_side = nullptr;
  CenterWork cw_side, cw_main;  // workspaces of the sums on each stream
  bool charges_pending = false;  // tcells hold many-body charges the cells' x, y, value lack
  bool side_valid = false;   // dsums_side holds the sums of the current x, y
  bool side_busy = false;    // the side stream may still be reading x, y
  uint32_t centers = 0;      // center forces of this simulation
  LinkSerialState* dlink = nullptr;
  uint32_t* hscalar_u = nullptr;   // pinned mirrors
  LinkSerialState* hlink = nullptr;
  // exact collide: per-node candidate lists, pair ranks, dataflow counters,
  // the velocities at the iteration's start and the serial pass's state
  Device<uint32_t> c_count, c_off, c_cand, c_leaf, c_top, c_run, c_rank, c_indeg, c_done, c_part,
      c_stack, c_fin, c_list;
  Device<double> c_svx, c_svy;
  Device<CRec> c_rec;
  CollideSerialState* dcol = nullptr;
  CollideSerialState* hcol = nullptr;
  // host copies for tree introspection
  Pinned<Cell> hcells;
  Pinned<uint32_t> hperm;
  double hextent[4];
  d3f_stats stats{};
  unsigned long long* dcounters = nullptr;  // traversal visits, accepts, pairs
  struct Timer { int phase; hipEvent_t a, b; };
  hipEvent_t side_open = nullptr;
  std::vector<Timer> pending;
  std::vector<hipEvent_t> spare;
  hipEvent_t open_event = nullptr;
  int open_phase = 0;

  ~Sim() {
    // the streams are shared with the device's other simulations: wait for
    // this one's work, keep the streams
    if (stream) (void)hipStreamSynchronize(stream);
    if (side) (void)hipStreamSynchronize(side);
    if (aux) (void)hipStreamSynchronize(aux);
    for (Force* f : forces) delete f;
    for (auto* a : {&x, &y, &vx, &vy, &fx, &fy, &snap}) a->release();
    fixed.release();
    for (auto* a : {&dx, &dy, &dvx, &dvy, &dfx, &dfy, &dpx, &dpy, &dsnap}) a->release();
    dpos.release();
    dlnodes.release();
    dfixed.release();
    tree_free(tree);
    fmm_free(fmm);
    for (auto* a : {&sk0, &sk1}) a->release();
    for (auto* a : {&sv0, &sv1, &shist, &sglob, &scount, &gstart}) a->release();
    gcand.release();
    gpartials.release();
    if (gparams) (void)hipFree(gparams);
    flagged.release();
    order.release();
    draws_count.release();
    draw_off.release();
    draws.release();
    if (dscalar_u) (void)hipFree(dscalar_u);
    if (dscalar_d) (void)hipFree(dscalar_d);
    if (dlink) (void)hipFree(dlink);
    if (hscalar_u) (void)hipHostFree(hscalar_u);
    if (hlink) (void)hipHostFree(hlink);
    for (auto* a : {&c_count, &c_off, &c_cand, &c_leaf, &c_top, &c_run, &c_rank, &c_indeg, &c_done,
                    &c_part, &c_stack, &c_fin, &c_list})
      a->release();
    c_svx.release();
    c_svy.release();
    c_rec.release();
    if (dcol) (void)hipFree(dcol);
    if (hcol) (void)hipHostFree(hcol);
    hcells.release();
    hperm.release();
    for (auto& t : pending) {
      (void)hipEventDestroy(t.a);
      (void)hipEventDestroy(t.b);
    }
    for (auto e : spare) (void)hipEventDestroy(e);
    if (open_event) (void)hipEventDestroy(open_event);
    if (dcounters) (void)hipFree(dcounters);
    if (side) (void)hipStreamSynchronize(side);
    if (dsums_side) (void)hipFree(dsums_side);
    center_work_free(cw_side);
    center_work_free(cw_main);
    if (ev_x_ready) (void)hipEventDestroy(ev_x_ready);
    if (ev_sums_done) (void)hipEventDestroy(ev_sums_done);
  }

  hipError_t init() {
    HIPCHECK(shared_streams(&stream, &side, &aux));
    HIPCHECK(hipMalloc(&dscalar_u, 16 * sizeof(uint32_t)));
    HIPCHECK(hipMalloc(&dscalar_d, 16 * sizeof(double)));
    HIPCHECK(hipMalloc(&dlink, sizeof(LinkSerialState)));
    HIPCHECK(hipHostMalloc((void**)&hscalar_u, 16 * sizeof(uint32_t), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc((void**)&hlink, sizeof(LinkSerialState), hipHostMallocDefault));
    HIPCHECK(hipMalloc(&dcol, sizeof(CollideSerialState)));
    HIPCHECK(hipHostMalloc((void**)&hcol, sizeof(CollideSerialState), hipHostMallocDefault));
    HIPCHECK(hipEventCreateWithFlags(&ev_x_ready, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&ev_sums_done, hipEventDisableTiming));
    HIPCHECK(hipMalloc(&dsums_side, 2 * sizeof(double)));
    HIPCHECK(hipMalloc(&dcounters, 4 * sizeof(unsigned long long)));
    HIPCHECK(hipMalloc(&gparams, sizeof(GridParams)));
    HIPCHECK(sglob.reserve(8 * 256));
    HIPCHECK(scount.reserve(4));
    HIPCHECK(hipMemset(dcounters, 0, 4 * sizeof(unsigned long long)));
    for (int k = 0; k < 4; ++k) hextent[k] = __builtin_nan("");
    return hipSuccess;
  }

  // sim.zig:59-73
  hipError_t set_node_count(uint32_t nn, bool* grown) {
    HIPCHECK(before_x_write());
    side_valid = false;
    const uint32_t old_n = n;
    n = nn;
    *grown = false;
    for (Force* f : forces)
      if (f && f->node_sized() && f->count < nn) {
        HIPCHECK(f->set_count(nn));
        *grown = true;
      }
    if (nn <= cap) return hipSuccess;
    const uint32_t c = std::max(std::max(nn, cap * 2), 16u);
    for (auto* a : {&x, &y, &vx, &vy, &fx, &fy}) HIPCHECK(a->reserve(c, old_n));
    HIPCHECK(fixed.reserve(c, old_n));
    for (auto* a : {&dx, &dy, &dvx, &dvy, &dfx, &dfy}) HIPCHECK(a->reserve(c, old_n, stream));
    for (auto* a : {&dpx, &dpy}) HIPCHECK(a->reserve(c, 0, stream));
    HIPCHECK(dpos.reserve(c, 0, stream));
    HIPCHECK(dfixed.reserve(c, old_n, stream));
    HIPCHECK(flagged.reserve(c, 0, stream));
    HIPCHECK(order.reserve(c, 0, stream));
    HIPCHECK(draws_count.reserve(c, 0, stream));
    HIPCHECK(draw_off.reserve(c, 0, stream));
    cap = c;
    *grown = true;
    return hipSuccess;
  }

  // The main stream must not write x, y while the side stream reads them.
  hipError_t before_x_write() {
    if (side_busy) {
      HIPCHECK(hipStreamWaitEvent(stream, ev_sums_done, 0));
      side_busy = false;
    }
    side_valid = false;
    return hipSuccess;
  }
  // Starts the sums of the current x, y on the side stream.
  hipError_t speculate_center() {
    if (centers == 0 || n == 0) return hipSuccess;
    HIPCHECK(hipEventRecord(ev_x_ready, stream));
    HIPCHECK(hipStreamWaitEvent(side, ev_x_ready, 0));
    hipEvent_t a = g_timing ? take_event() : nullptr;
    if (a) HIPCHECK(hipEventRecord(a, side));
    HIPCHECK(center_raw_sum(dx.p, dy.p, n, dsums_side, cw_side, side));
    if (a) {
      hipEvent_t b = take_event();
      if (b) {
        HIPCHECK(hipEventRecord(b, side));
        pending.push_back(Timer{D3F_PHASE_CENTER_SUM, a, b});
      }
    }
    HIPCHECK(hipEventRecord(ev_sums_done, side));
    side_valid = true;
    side_busy = true;
    return hipSuccess;
  }

  // Radix-sort scratch for `items` entries.
  hipError_t sort_buffers(size_t items, SortBuffers* out) {
    const size_t c = items > 0 ? items : 1;
    HIPCHECK(sk0.reserve(c, 0, stream));
    HIPCHECK(sk1.reserve(c, 0, stream));
    HIPCHECK(sv0.reserve(c, 0, stream));
    HIPCHECK(sv1.reserve(c, 0, stream));
    HIPCHECK(shist.reserve(sort_hist_words((uint32_t)c), 0, stream));
    out->k0 = sk0.p;
    out->k1 = sk1.p;
    out->v0 = sv0.p;
    out->v1 = sv1.p;
    out->hist = shist.p;
    out->glob = sglob.p;
    out->count = scount.p;
    out->cap = (uint32_t)c;
    return hipSuccess;
  }

  hipEvent_t take_event() {
    hipEvent_t e = nullptr;
    if (!spare.empty()) {
      e = spare.back();
      spare.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      e = nullptr;
    }
    return e;
  }
  // Brackets a phase's launches with events on the stream; no synchronisation.
  void timer_start(int phase) {
    if (!g_timing) return;
    open_phase = phase;
    open_event = take_event();
    if (open_event) (void)hipEventRecord(open_event, stream);
  }
  void timer_stop() {
    if (!g_timing || !open_event) return;
    hipEvent_t b = take_event();
    if (!b) return;
    (void)hipEventRecord(b, stream);
    pending.push_back(Timer{open_phase, open_event, b});
    open_event = nullptr;
  }
  hipError_t collect() {
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipStreamSynchronize(side));
    for (auto& t : pending) {
      float ms = 0;
      HIPCHECK(hipEventElapsedTime(&ms, t.a, t.b));
      stats.ms[t.phase] += ms;
      stats.launches[t.phase] += 1;
      spare.push_back(t.a);
      spare.push_back(t.b);
    }
    pending.clear();
    unsigned long long c[4];
    HIPCHECK(hipMemcpy(c, dcounters, sizeof(c), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemset(dcounters, 0, sizeof(c)));
    stats.visits += c[0];
    stats.accepts += c[1];
    stats.pairs += c[2];
    return hipSuccess;
  }
};

std::vector<Sim*> g_sims;

Sim* get(uint32_t h) {
  if (h == 0 || h > g_sims.size()) return nullptr;
  return g_sims[h - 1];
}
Force* get_force(Sim* s, uint32_t fh) {
  if (!s || fh == 0 || fh > s->forces.size()) return nullptr;
  return s->forces[fh - 1];
}

bool fast_mode(const Sim& s, const Force& f) {  // root.zig:246-248
  return f.scalar(S_PRECISION) == 1 && s.n >= MIN_N;
}

// Uploads a force's parameter arrays when the facade may have written them.
hipError_t upload(Sim& s, Force& f) {
  if (!f.dirty) return hipSuccess;
  const size_t c = f.count > 0 ? f.count : 1;
  for (int i = 0; i < f.float_arrays(); ++i) {
    HIPCHECK(f.df[i].reserve(c, 0, s.stream));
    if (f.count) HIPCHECK(hipMemcpyAsync(f.df[i].p, f.f[i].p, f.count * sizeof(double), hipMemcpyHostToDevice, s.stream));
  }
  if (f.kind == K_LINK) {
    HIPCHECK(f.dsrc.reserve(c, 0, s.stream));
    HIPCHECK(f.dtgt.reserve(c, 0, s.stream));
    if (f.count) {
      HIPCHECK(hipMemcpyAsync(f.dsrc.p, f.source.p, f.count * sizeof(uint32_t), hipMemcpyHostToDevice, s.stream));
      HIPCHECK(hipMemcpyAsync(f.dtgt.p, f.target.p, f.count * sizeof(uint32_t), hipMemcpyHostToDevice, s.stream));
    }
  }
  HIPCHECK(hipStreamSynchronize(s.stream));
  f.dirty = false;
  f.sched_valid = false;
  return hipSuccess;
}

// ---------------------------------------------------------------- many-body (exact)
int32_t run_many_body_exact(Sim& s, Force& f, double alpha) {
  TreeBuffers& t = s.tree;
  s.timer_start(D3F_PHASE_TREE);
  s.charges_pending = false;
  hipError_t e = tree_build_exact(t, s.dx.p, s.dy.p, s.n, s.stream);
  if (e != hipSuccess) return status_of(e);
  s.timer_stop();
  if (!t.meta_host->valid) return OK;  // many_body.zig:33
  s.timer_start(D3F_PHASE_ACCUMULATE);
  e = mb_accumulate(t, s.dx.p, s.dy.p, f.df[0].p, f.scalar(S_THETA2), s.stream);
  if (e != hipSuccess) return status_of(e);
  s.charges_pending = true;
  s.timer_stop();
  TravArgs a{};
  a.tc = t.tcells;
  a.order = t.perm;
  a.count = t.meta_host->m;
  a.perm = t.perm;
  a.x = s.dx.p;
  a.y = s.dy.p;
  a.s = f.df[0].p;
  a.vx = s.dvx.p;
  a.vy = s.dvy.p;
  a.dmin2 = f.scalar(S_DMIN2);
  a.dmax2 = f.scalar(S_DMAX2);
  a.alpha = alpha;
  a.flagged = s.flagged.p;
  a.n_flagged = s.dscalar_u;
  a.len = t.meta_host->len;
  a.counters = s.dcounters;
  if ((e = hipMemsetAsync(s.dscalar_u, 0, sizeof(uint32_t), s.stream)) != hipSuccess) return status_of(e);
  s.timer_start(D3F_PHASE_TRAVERSAL);
  if ((e = mb_traverse(a, TRAV_RECORD, s.stream)) != hipSuccess) return status_of(e);
  s.timer_stop();
  if ((e = mb_nan_nodes(t, s.n, s.dvx.p, s.dvy.p, s.stream)) != hipSuccess) return status_of(e);
  if ((e = hipMemcpyAsync(s.hscalar_u, s.dscalar_u, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
    return status_of(e);
  if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return status_of(e);
  const uint32_t nf = s.hscalar_u[0];
  s.stats.flagged += nf;
  if (nf == 0) return OK;
  // Jiggle replay (many_body.zig:77-83): ascending node order, real draws.
  std::vector<uint32_t> list(nf), counts(nf), offs(nf);
  if ((e = hipMemcpy(list.data(), s.flagged.p, nf * sizeof(uint32_t), hipMemcpyDeviceToHost)) != hipSuccess)
    return status_of(e);
  std::sort(list.begin(), list.end());
  if ((e = hipMemcpy(s.order.p, list.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return status_of(e);
  a.order = s.order.p;
  a.count = nf;
  a.draws_count = s.draws_count.p;
  if ((e = mb_traverse(a, TRAV_COUNT, s.stream)) != hipSuccess) return status_of(e);
  if ((e = hipMemcpyAsync(counts.data(), s.draws_count.p, nf * sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream)) !=
      hipSuccess)
    return status_of(e);
  if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return status_of(e);
  uint32_t total = 0;
  for (uint32_t k = 0; k < nf; ++k) {
    offs[k] = total;
    total += counts[k];
  }
  std::vector<double> values(total > 0 ? total : 1);
  for (uint32_t k = 0; k < total; ++k) values[k] = draw();
  if ((e = s.draws.reserve(values.size(), 0, s.stream)) != hipSuccess) return status_of(e);
  if ((e = hipMemcpy(s.draws.p, values.data(), values.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return status_of(e);
  if ((e = hipMemcpy(s.draw_off.p, offs.data(), nf * sizeof(uint32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return status_of(e);
  a.draw_off = s.draw_off.p;
  a.draws = s.draws.p;
  if ((e = mb_traverse(a, TRAV_REPLAY, s.stream)) != hipSuccess) return status_of(e);
  return status_of(hipStreamSynchronize(s.stream));
}

// ---------------------------------------------------------------- link (exact)
// The dependency-level schedule (see link.hip), built on the host when the
// links change: a serial O(m) scan, as the reference's CSR build is.
hipError_t build_schedule(Sim& s, Force& f) {
  const uint32_t m = f.count, n = s.n;
  std::vector<uint32_t> last(n, 0), level(m);
  uint32_t levels = 0;
  for (uint32_t e = 0; e < m; ++e) {
    const uint32_t a = f.source.p[e], b = f.target.p[e];
    const uint32_t l = std::max(last[a], last[b]) + 1;
    last[a] = last[b] = l;
    level[e] = l;
    levels = std::max(levels, l);
  }
  f.level_off.assign(levels + 2, 0);
  for (uint32_t e = 0; e < m; ++e) f.level_off[level[e] + 1]++;
  for (uint32_t l = 1; l < levels + 2; ++l) f.level_off[l] += f.level_off[l - 1];
  std::vector<uint32_t> pos(f.level_off.begin(), f.level_off.end());
  std::vector<uint32_t> ls(m), lt(m);
  std::vector<double> l0(m), l1(m), l2(m);
  for (uint32_t e = 0; e < m; ++e) {
    const uint32_t p = pos[level[e]]++;
    ls[p] = f.source.p[e];
    lt[p] = f.target.p[e];
    l0[p] = f.f[0].p[e];
    l1[p] = f.f[1].p[e];
    l2[p] = f.f[2].p[e];
  }
  const size_t c = m > 0 ? m : 1;
  HIPCHECK(f.lsrc.reserve(c, 0, s.stream));
  HIPCHECK(f.ltgt.reserve(c, 0, s.stream));
  for (auto& a : f.lf) HIPCHECK(a.reserve(c, 0, s.stream));
  if (m) {
    HIPCHECK(hipMemcpy(f.lsrc.p, ls.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(f.ltgt.p, lt.data(), m * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(f.lf[0].p, l0.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(f.lf[1].p, l1.data(), m * sizeof(double), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(f.lf[2].p, l2.data(), m * sizeof(double), hipMemcpyHostToDevice));
  }
  f.sched_valid = true;
  f.sched_n = n;
  return hipSuccess;
}

int32_t run_link_exact(Sim& s, Force& f, double alpha) {
  const uint32_t m = f.count;
  if (!f.sched_valid || f.sched_n != s.n) {
    for (uint32_t e = 0; e < m; ++e)  // link.zig:10-12
      if (f.source.p[e] >= s.n || f.target.p[e] >= s.n) return BAD_ARGUMENT;
    hipError_t e = build_schedule(s, f);
    if (e != hipSuccess) return status_of(e);
  }
  s.stats.link_levels = f.level_off.size() >= 2 ? f.level_off.size() - 2 : 0;
  const double iterations = f.scalar(S_ITERATIONS);
  if (m == 0 || !(0 < iterations)) return OK;
  hipError_t e;
  if ((e = s.dlnodes.reserve(s.n > 0 ? s.n : 1, 0, s.stream)) != hipSuccess) return status_of(e);
  if ((e = hipMemsetAsync(s.dscalar_u + 1, 0, sizeof(uint32_t), s.stream)) != hipSuccess) return status_of(e);
  LinkLevelArgs a{f.lsrc.p, f.ltgt.p, f.lf[0].p, f.lf[1].p, f.lf[2].p, s.dx.p, s.dy.p, s.dvx.p, s.dvy.p,
                  s.dlnodes.p, s.n, alpha, s.dscalar_u + 1};
  const uint32_t levels = (uint32_t)f.level_off.size() - 2;
  s.timer_start(D3F_PHASE_LINK);
  if ((e = link_pack(a, s.stream)) != hipSuccess) return status_of(e);
  for (double k = 0; k < iterations; k += 1) {
    for (uint32_t l = 1; l <= levels; ++l) {
      if ((e = link_level(a, f.level_off[l], f.level_off[l + 1], s.stream)) != hipSuccess) return status_of(e);
    }
  }
  if ((e = link_unpack(a, s.stream)) != hipSuccess) return status_of(e);
  if ((e = hipMemcpyAsync(s.hscalar_u + 1, s.dscalar_u + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream)) !=
      hipSuccess)
    return status_of(e);
  s.timer_stop();
  if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return status_of(e);
  if (s.hscalar_u[1] == 0) return OK;
  // A zero or NaN delta: d3 draws in edge order. vx, vy still hold their
  // values from before the pass; run it serially.
  for (double k = 0; k < iterations; k += 1) {
    LinkSerialState* st = s.hlink;
    st->edge = 0;
    st->need = 0;
    st->has_jig = 0;
    for (;;) {
      if ((e = hipMemcpyAsync(s.dlink, st, sizeof(LinkSerialState), hipMemcpyHostToDevice, s.stream)) != hipSuccess)
        return status_of(e);
      if ((e = link_serial(f.dsrc.p, f.dtgt.p, f.df[0].p, f.df[1].p, f.df[2].p, m, s.dx.p, s.dy.p, s.dvx.p, s.dvy.p,
                           alpha, s.dlink, s.stream)) != hipSuccess)
        return status_of(e);
      if ((e = hipMemcpyAsync(st, s.dlink, sizeof(LinkSerialState), hipMemcpyDeviceToHost, s.stream)) != hipSuccess)
        return status_of(e);
      if ((e = hipStreamSynchronize(s.stream)) != hipSuccess) return status_of(e);
      if (st->edge >= m) break;
      if (st->need & 1) st->jig[0] = jiggle_host(draw());
      if (st->need & 2) st->jig[1] = jiggle_host(draw());
      st->has_jig = 1;
    }
  }
  return OK;
}

// ---------------------------------------------------------------- collide (exact)
// One iteration of collide.zig:27-101: the tree on x + vx, the cell radii,
// then the pass over i as a dataflow over the candidate conflicts
// (collide.hip); when the parallel pass cannot guarantee the reference's
// result (see collide.hip) the velocities are restored and the iteration
// runs on the serial kernel, which takes jiggles from the host's random in
// pass order.
// Jiggles drawn in one iteration, replayed in order if the iteration restarts.
struct JiggleReplay {
  std::vector<double> values;
  size_t pos = 0;
  double next() {
    if (pos == values.size()) values.push_back(jiggle_host(draw()));
    return values[pos++];
  }
};

// Runs k_collide_serial to the end of the list (all nodes when list is
// null), drawing jiggles on the host when it stops for them. *moved reports a
// push outside the boxes of a continuation.
hipError_t collide_serial_run(Sim& s, const CollideArgs& a, const uint32_t* list, uint32_t count,
                              JiggleReplay& replay, bool* moved) {
  HIPCHECK(s.c_stack.reserve(3 * (size_t)s.tree.meta_host->depth + 8, 0, s.stream));
  CollideSerialState* st = s.hcol;
  std::memset(st, 0, sizeof(*st));
  if (moved) *moved = false;
  for (;;) {
    HIPCHECK(hipMemcpyAsync(s.dcol, st, sizeof(*st), hipMemcpyHostToDevice, s.stream));
    HIPCHECK(collide_serial(a, s.c_stack.p, s.dcol, list, count, s.stream));
    HIPCHECK(hipMemcpyAsync(st, s.dcol, sizeof(*st), hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    if (st->node >= count) return hipSuccess;
    if (st->need & COLLIDE_NEED_MOVED) {
      if (moved) *moved = true;
      return hipSuccess;
    }
    if (st->need & 1) st->jig[0] = replay.next();
    if (st->need & 2) st->jig[1] = replay.next();
    st->has_jig = 1;
  }
}

hipError_t collide_iteration(Sim& s, Force& f) {
  const uint32_t n = s.n;
  JiggleReplay replay;
  TreeBuffers& t = s.tree;
  s.charges_pending = false;
  HIPCHECK(shift_positions(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, s.dpx.p, s.dpy.p, n, s.stream));
  HIPCHECK(tree_build_exact(t, s.dpx.p, s.dpy.p, n, s.stream));
  if (!t.meta_host->valid) return hipSuccess;  // collide.zig:34
  HIPCHECK(collide_radii(t, f.df[0].p, s.stream));
  const size_t c = n > 0 ? n : 1;
  for (auto* b : {&s.c_count, &s.c_off, &s.c_run, &s.c_indeg, &s.c_done, &s.c_fin, &s.c_list})
    HIPCHECK(b->reserve(c, 0, s.stream));
  HIPCHECK(s.c_cand.reserve(c * COLLIDE_K, 0, s.stream));
  HIPCHECK(s.c_leaf.reserve(c * COLLIDE_K, 0, s.stream));
  HIPCHECK(s.c_top.reserve(c * COLLIDE_K, 0, s.stream));
  HIPCHECK(s.c_part.reserve(c / 4096 + 2, 0, s.stream));
  HIPCHECK(s.c_svx.reserve(c, 0, s.stream));
  HIPCHECK(s.c_svy.reserve(c, 0, s.stream));
  uint32_t* w = s.dscalar_u;  // [4] pairs, [5] status, [6] ticket, [7] n
  CollideArgs a{};
  HIPCHECK(s.c_rec.reserve(t.meta_host->len, 0, s.stream));
  a.cells = t.cells;
  a.rec = s.c_rec.p;
  a.meta = t.meta;
  a.perm = t.perm;
  a.px = s.dpx.p;
  a.py = s.dpy.p;
  a.x = s.dx.p;
  a.y = s.dy.p;
  a.vx = s.dvx.p;
  a.vy = s.dvy.p;
  a.radii = f.df[0].p;
  a.strength = f.scalar(S_STRENGTH);
  a.n = n;
  a.m = t.meta_host->m;
  a.key_shift = n <= 2 ? 1 : 32 - __builtin_clz(n - 1);
  a.ccount = s.c_count.p;
  a.coff = s.c_off.p;
  a.cand = s.c_cand.p;
  a.cleaf = s.c_leaf.p;
  a.ctop = s.c_top.p;
  a.run_start = s.c_run.p;
  a.indeg = s.c_indeg.p;
  a.done = s.c_done.p;
  a.fin = s.c_fin.p;
  a.status = w + 5;
  a.ticket = w + 6;
  a.stop = w + 8;
  HIPCHECK(collide_records(a, t.meta_host->len, s.stream));
  HIPCHECK(hipMemsetD32Async((hipDeviceptr_t)(w + 7), (int)n, 1, s.stream));
  HIPCHECK(hipMemcpyAsync(s.c_svx.p, s.dvx.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
  HIPCHECK(hipMemcpyAsync(s.c_svy.p, s.dvy.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
  // A pass whose pushes leave their boxes is tried again with boxes three
  // times wider (up to three passes); any other stop goes to the serial pass.
  bool serial = !std::isfinite(a.strength);
  for (int attempt = 0; attempt < 3 && !serial; ++attempt) {
    a.scale = f.collide_scale;
    HIPCHECK(hipMemsetAsync(w + 4, 0, 3 * sizeof(uint32_t), s.stream));
    HIPCHECK(collide_candidates(a, s.stream));
    HIPCHECK(scan_u32(a.ccount, a.coff, n, w + 7, s.c_part.p, w + 4, s.stream));
    HIPCHECK(hipMemcpyAsync(s.hscalar_u + 4, w + 4, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    const uint32_t pairs = s.hscalar_u[4];
    if (s.hscalar_u[5] != 0) {
      s.stats.collide_status |= s.hscalar_u[5];
      serial = true;
      break;
    }
    if (pairs == 0) return hipSuccess;  // no pair can overlap: the pass pushes nothing
    SortBuffers sb;
    HIPCHECK(s.sort_buffers(pairs, &sb));
    HIPCHECK(s.c_rank.reserve(pairs, 0, s.stream));
    a.pkeys = sb.k0;
    a.pvals = sb.v0;
    a.rank = s.c_rank.p;
    HIPCHECK(collide_pairs(a, s.stream));
    HIPCHECK(sort_digits(sb.hist, sb.glob, sb.k0, sb.v0, sb.k1, sb.v1, pairs, w + 4, 0, (2 * a.key_shift + 7) / 8,
                         s.stream));
    HIPCHECK(hipMemsetAsync(a.indeg, 0, (size_t)n * sizeof(uint32_t), s.stream));
    HIPCHECK(hipMemsetAsync(a.done, 0, (size_t)n * sizeof(uint32_t), s.stream));
    HIPCHECK(hipMemsetAsync(a.fin, 0, (size_t)n * sizeof(uint32_t), s.stream));
    HIPCHECK(hipMemsetD32Async((hipDeviceptr_t)a.stop, (int)NONE, 1, s.stream));
    HIPCHECK(collide_ranks(a, pairs, s.stream));
    HIPCHECK(collide_apply(a, s.stream));
    HIPCHECK(hipMemcpyAsync(s.hscalar_u + 5, w + 5, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipMemcpyAsync(s.hscalar_u + 8, w + 8, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    const uint32_t status = s.hscalar_u[5];
    if (status == 0) {
      f.collide_scale = std::max(0.25, 0.5 * f.collide_scale);
      if (s.hscalar_u[8] == NONE) return hipSuccess;
      // A node stopped for a jiggle: the unfinished nodes run serially, in
      // index order, after the finished ones.
      s.stats.collide_status |= COLLIDE_JIGGLE;
      HIPCHECK(collide_unfinished(a, s.c_run.p, s.c_indeg.p, s.c_list.p, w + 7, s.c_part.p, w + 9, s.stream));
      HIPCHECK(hipMemcpyAsync(s.hscalar_u + 9, w + 9, sizeof(uint32_t), hipMemcpyDeviceToHost, s.stream));
      HIPCHECK(hipStreamSynchronize(s.stream));
      bool moved = false;
      HIPCHECK(collide_serial_run(s, a, s.c_list.p, s.hscalar_u[9], replay, &moved));
      if (!moved) return hipSuccess;
      // A push left its box, which the finished nodes assumed: the whole
      // iteration again, serially, replaying the jiggles already drawn.
      HIPCHECK(hipMemcpyAsync(s.dvx.p, s.c_svx.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
      HIPCHECK(hipMemcpyAsync(s.dvy.p, s.c_svy.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
      serial = true;
      break;
    }
    s.stats.collide_status |= status;
    HIPCHECK(hipMemcpyAsync(s.dvx.p, s.c_svx.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    HIPCHECK(hipMemcpyAsync(s.dvy.p, s.c_svy.p, (size_t)n * sizeof(double), hipMemcpyDeviceToDevice, s.stream));
    if (status != COLLIDE_MOVED || f.collide_scale >= 8) {
      serial = true;
      break;
    }
    f.collide_scale *= 3;
    serial = attempt == 2;
  }
  ++s.stats.collide_serial;
  replay.pos = 0;
  return collide_serial_run(s, a, nullptr, n, replay, nullptr);
}

int32_t run_collide_exact(Sim& s, Force& f) {  // collide.zig:17-102
  const double iterations = f.scalar(S_ITERATIONS);
  s.timer_start(D3F_PHASE_COLLIDE);
  for (double k = 0; k < iterations; k += 1) {
    hipError_t e = collide_iteration(s, f);
    if (e != hipSuccess) return status_of(e);
  }
  s.timer_stop();
  return OK;
}

// ---------------------------------------------------------------- fast mode
uint64_t seed_of(const Force& f) {  // hash.zig:28-30: the seed's bits
  const double v = f.scalar(S_SEED);
  uint64_t u;
  std::memcpy(&u, &v, 8);
  return u;
}

int32_t run_link_fast(Sim& s, Force& f, double alpha) {  // link_fast.zig:22-99
  const uint32_t m = f.count, n = s.n;
  hipError_t e;
  if (!f.csr_valid || f.csr_n != n) {
    for (uint32_t k = 0; k < m; ++k)  // link_fast.zig:24-26
      if (f.source.p[k] >= n || f.target.p[k] >= n) return BAD_ARGUMENT;
    SortBuffers sb;
    if ((e = s.sort_buffers(2 * (size_t)m, &sb)) != hipSuccess) return status_of(e);
    if ((e = f.row.reserve((size_t)n + 1, 0, s.stream)) != hipSuccess) return status_of(e);
    if ((e = f.adj.reserve(2 * (size_t)m > 0 ? 2 * (size_t)m : 1, 0, s.stream)) != hipSuccess) return status_of(e);
    const size_t c = m > 0 ? m : 1;
    if ((e = f.owner.reserve(2 * c, 0, s.stream)) != hipSuccess) return status_of(e);
    if ((e = f.entries.reserve(2 * c, 0, s.stream)) != hipSuccess) return status_of(e);
    if ((e = f.terms.reserve(4 * c, 0, s.stream)) != hipSuccess) return status_of(e);
    if ((e = f.rows_long.reserve(n > 0 ? n : 1, 0, s.stream)) != hipSuccess) return status_of(e);
    if ((e = link_csr(f.csr_args(n), sb, &f.long_count, s.stream)) != hipSuccess) return status_of(e);
    f.csr_valid = true;
    f.csr_n = n;
  }
  const double iterations = f.scalar(S_ITERATIONS);
  const uint64_t seed = seed_of(f);
  s.timer_start(D3F_PHASE_LINK);
  for (double k = 0; k < iterations; k += 1) {
    e = shift_pack(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, s.dpos.p, n, s.stream);
    if (e == hipSuccess) e = link_sweep(f.csr_args(n), f.long_count, s.dpos.p, s.dvx.p, s.dvy.p, alpha, seed, s.stream);
    if (e != hipSuccess) return status_of(e);
  }
  s.timer_stop();
  return OK;
}

int32_t run_collide_fast(Sim& s, Force& f) {  // grid.zig:198-286
  const uint32_t n = s.n;
  hipError_t e;
  SortBuffers sb;
  if ((e = s.sort_buffers(n, &sb)) != hipSuccess) return status_of(e);
  const size_t c = n > 0 ? n : 1;
  if ((e = s.gstart.reserve(4 * c + 17, 0, s.stream)) != hipSuccess) return status_of(e);
  if ((e = s.gcand.reserve(c, 0, s.stream)) != hipSuccess) return status_of(e);
  if ((e = s.gpartials.reserve(grid_partials_bytes(), 0, s.stream)) != hipSuccess) return status_of(e);
  GridBuffers g;
  g.px = s.dpx.p;
  g.py = s.dpy.p;
  g.start = s.gstart.p;
  g.cand = s.gcand.p;
  g.partials = s.gpartials.p;
  g.params = s.gparams;
  const double iterations = f.scalar(S_ITERATIONS);
  const uint64_t seed = seed_of(f);
  s.timer_start(D3F_PHASE_COLLIDE);
  for (double k = 0; k < iterations; k += 1) {
    e = grid_collide_iteration(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, f.df[0].p, n, f.scalar(S_STRENGTH), seed, g, sb,
                               s.stream);
    if (e != hipSuccess) return status_of(e);
  }
  s.timer_stop();
  return OK;
}

int32_t run_many_body_fast(Sim& s, Force& f, double alpha) {  // fmm.zig:247-354
  FmmArgs a;
  a.x = s.dx.p;
  a.y = s.dy.p;
  a.strength = f.df[0].p;
  a.vx = s.dvx.p;
  a.vy = s.dvy.p;
  a.n = s.n;
  const double order = f.scalar(S_ORDER);
  a.order = order >= 1 && order <= (double)FMM_MAX_ORDER ? (uint32_t)order : 8;
  a.alpha = alpha;
  a.opening = f.scalar(S_OPENING);
  a.dmin2 = f.scalar(S_DMIN2);
  a.dmax2 = f.scalar(S_DMAX2);
  const uint64_t seed = seed_of(f);
  const uint32_t node_off[2] = {0, s.n};
  a.graphs = 1;
  a.node_off = node_off;
  a.node_graph = nullptr;
  a.seeds = &seed;
  a.aux = s.aux;
  s.timer_start(D3F_PHASE_FMM);
  hipError_t e = fmm_run(s.fmm, s.tree, a, s.stream);
  s.timer_stop();
  if (e == hipErrorNotSupported) return run_many_body_exact(s, f, alpha);  // fmm.zig:255-257
  if (e != hipSuccess) return status_of(e);
  const FmmParams& P = *s.fmm.params_host;
  s.stats.m2l += P.m2l;
  s.stats.p2p_pairs += P.pairs;
  s.stats.fmm_cells = P.level_off[FMM_LEVELS + 1];
  return OK;
}

int32_t run_center(Sim& s, Force& f) {  // simple.zig:14-31
  hipError_t e = hipSuccess;
  s.timer_start(D3F_PHASE_CENTER);
  const double* sums = s.dsums_side;
  if (s.side_valid) {
    e = hipStreamWaitEvent(s.stream, s.ev_sums_done, 0);
    s.side_busy = false;
  } else {
    sums = s.dscalar_d;
    e = s.before_x_write();
    if (e == hipSuccess) e = center_raw_sum(s.dx.p, s.dy.p, s.n, s.dscalar_d, s.cw_main, s.stream);
  }
  s.side_valid = false;
  if (e == hipSuccess)
    e = center_apply(s.dx.p, s.dy.p, s.n, sums, f.scalar(S_X), f.scalar(S_Y), f.scalar(S_STRENGTH), s.stream);
  s.timer_stop();
  return status_of(e);
}

// The snapshot's host copy (sim_ptr(h, D3F_PSNAP)): vx, vy, and x, y when
// the snapshot took positions, n entries each.
hipError_t fetch_snapshot(Sim& s) {
  if (!s.snap_pending) return hipSuccess;
  const size_t b = (size_t)s.n * sizeof(double) * (s.snap_positions ? 4 : 2);
  HIPCHECK(hipMemcpyAsync(s.snap.p, s.dsnap.p, b, hipMemcpyDeviceToHost, s.stream));
  HIPCHECK(hipStreamSynchronize(s.stream));
  s.snap_pending = false;
  return hipSuccess;
}

hipError_t refresh_tree_host(Sim& s) {
  TreeBuffers& t = s.tree;
  if (!t.meta) {
    for (int k = 0; k < 4; ++k) s.hextent[k] = __builtin_nan("");
    return hipSuccess;
  }
  HIPCHECK(hipStreamSynchronize(s.stream));
  HIPCHECK(hipMemcpy(t.meta_host, t.meta, sizeof(TreeMeta), hipMemcpyDeviceToHost));
  if (!t.meta_host->valid) t.meta_host->len = 0;
  if (s.charges_pending) {  // the last many-body run's charges, as the reference's cells hold them
    HIPCHECK(mb_publish_charges(t, s.stream));
    HIPCHECK(hipStreamSynchronize(s.stream));
    s.charges_pending = false;
  }
  for (int k = 0; k < 4; ++k) s.hextent[k] = t.meta_host->extent[k];
  const uint32_t len = t.meta_host->len, m = t.meta_host->m;
  HIPCHECK(s.hcells.reserve(len > 0 ? len : 1, 0));
  HIPCHECK(s.hperm.reserve(m > 0 ? m : 1, 0));
  if (len) HIPCHECK(hipMemcpy(s.hcells.p, t.cells, (size_t)len * sizeof(Cell), hipMemcpyDeviceToHost));
  if (m) HIPCHECK(hipMemcpy(s.hperm.p, t.perm, (size_t)m * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return hipSuccess;
}

}  // namespace
}  // namespace d3f

using namespace d3f;

// ---------------------------------------------------------------- exported ABI

EXPORT void core_set_random(d3f_random_fn fn, void* ctx) {
  g_random = fn;
  g_random_ctx = ctx;
}

EXPORT void core_set_timing(int on) { g_timing = on != 0; }

EXPORT int32_t core_synchronize(void) { return status_of(hipDeviceSynchronize()); }

EXPORT const char* core_last_error(void) { return g_last_error; }

EXPORT int32_t core_fail_next(int32_t status, uint32_t after, uint32_t count) {
  if (status != WORKER_FAILED && status != OUT_OF_MEMORY) return BAD_ARGUMENT;
  g_fail_status = status;
  g_fail_after = after;
  g_fail_count = count;
  return OK;
}

EXPORT uint32_t sim_create(void) {  // root.zig:65-78
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return 0;
  Sim* s = new (std::nothrow) Sim();
  if (!s) return 0;
  if (s->init() != hipSuccess) {
    delete s;
    return 0;
  }
  for (size_t i = 0; i < g_sims.size(); ++i)
    if (!g_sims[i]) {
      g_sims[i] = s;
      return (uint32_t)i + 1;
    }
  g_sims.push_back(s);
  return (uint32_t)g_sims.size();
}

EXPORT int32_t sim_destroy(uint32_t h) {  // root.zig:80-86
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  (void)hipStreamSynchronize(s->stream);
  (void)hipStreamSynchronize(s->side);
  delete s;
  g_sims[h - 1] = nullptr;
  return OK;
}

EXPORT int32_t sim_set_node_count(uint32_t h, uint32_t n) {  // root.zig:88-92
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  bool grown = false;
  hipError_t e = s->set_node_count(n, &grown);
  if (e != hipSuccess) return OUT_OF_MEMORY;
  return grown ? GROWN : OK;
}

EXPORT void* sim_ptr(uint32_t h, uint32_t which) {  // root.zig:94-98, sim.zig:75-88
  Sim* s = get(h);
  if (!s) return nullptr;
  switch (which) {
    case D3F_PX: return s->cap ? s->x.p : nullptr;
    case D3F_PY: return s->cap ? s->y.p : nullptr;
    case D3F_PVX: return s->cap ? s->vx.p : nullptr;
    case D3F_PVY: return s->cap ? s->vy.p : nullptr;
    case D3F_PFX: return s->cap ? s->fx.p : nullptr;
    case D3F_PFY: return s->cap ? s->fy.p : nullptr;
    case D3F_PFIXED: return s->cap ? s->fixed.p : nullptr;
    case D3F_PPERM:
      if (refresh_tree_host(*s) != hipSuccess) return nullptr;
      return s->tree.meta_host && s->tree.meta_host->m ? s->hperm.p : nullptr;
    case D3F_PSNAP: return fetch_snapshot(*s) == hipSuccess && s->snap.cap ? s->snap.p : nullptr;
    default: return nullptr;
  }
}

EXPORT int32_t sim_sync(uint32_t h, uint32_t direction) {
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const uint32_t n = s->n;
  if (n == 0) return OK;
  hipError_t e = hipSuccess;
  const size_t b = (size_t)n * sizeof(double);
  if (direction == D3F_SYNC_H2D) {
    e = s->before_x_write();
    struct { double* d; double* h; } pairs[] = {{s->dx.p, s->x.p}, {s->dy.p, s->y.p}, {s->dvx.p, s->vx.p},
                                                {s->dvy.p, s->vy.p}, {s->dfx.p, s->fx.p}, {s->dfy.p, s->fy.p}};
    for (auto& p : pairs)
      if (e == hipSuccess) e = hipMemcpyAsync(p.d, p.h, b, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s->dfixed.p, s->fixed.p, n, hipMemcpyHostToDevice, s->stream);
  } else if (direction == D3F_SYNC_D2H) {
    struct { double* h; double* d; } pairs[] = {{s->x.p, s->dx.p}, {s->y.p, s->dy.p}, {s->vx.p, s->dvx.p},
                                                {s->vy.p, s->dvy.p}};
    for (auto& p : pairs)
      if (e == hipSuccess) e = hipMemcpyAsync(p.h, p.d, b, hipMemcpyDeviceToHost, s->stream);
  } else {
    return BAD_ARGUMENT;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess && direction == D3F_SYNC_H2D) e = s->speculate_center();
  return status_of(e);
}

EXPORT int32_t sim_snapshot(uint32_t h, uint32_t positions) {  // root.zig:205-209, sim.zig:91-100
  // GPU: the copy stays in HBM (one copy kernel on the stream); sim_ptr(h,
  // D3F_PSNAP) brings it to the host array when asked for it.
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const uint32_t n = s->n;
  const size_t c = 4 * (size_t)(s->cap > 0 ? s->cap : 1);
  if (s->snap.reserve(c, 0) != hipSuccess || s->dsnap.reserve(c, 0, s->stream) != hipSuccess) return OUT_OF_MEMORY;
  hipError_t e = hipSuccess;
  if (n) e = snapshot_copy(s->dvx.p, s->dvy.p, positions ? s->dx.p : nullptr, positions ? s->dy.p : nullptr, n,
                           s->dsnap.p, s->stream);
  s->snap_pending = n > 0;
  s->snap_positions = positions != 0;
  return status_of(e);
}

EXPORT int32_t sim_integrate(uint32_t h, double velocity_decay) {  // root.zig:250-254
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  int32_t injected;
  if (injected_failure(&injected)) return injected;
  hipError_t e = s->before_x_write();
  s->timer_start(D3F_PHASE_INTEGRATE);
  if (e == hipSuccess)
    e = integrate(s->dx.p, s->dy.p, s->dvx.p, s->dvy.p, s->dfx.p, s->dfy.p, s->dfixed.p, s->n, velocity_decay,
                  s->stream);
  s->timer_stop();
  if (e == hipSuccess) e = s->speculate_center();
  return status_of(e);
}

EXPORT int32_t tree_build(uint32_t h, uint32_t shifted) {  // root.zig:102-108
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const void* c0 = s->hcells.p;
  const void* p0 = s->hperm.p;
  hipError_t e;
  s->charges_pending = false;
  if (shifted) {
    e = shift_positions(s->dx.p, s->dy.p, s->dvx.p, s->dvy.p, s->dpx.p, s->dpy.p, s->n, s->stream);
    if (e == hipSuccess) e = tree_build_exact(s->tree, s->dpx.p, s->dpy.p, s->n, s->stream);
  } else {
    e = tree_build_exact(s->tree, s->dx.p, s->dy.p, s->n, s->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess) e = refresh_tree_host(*s);
  if (e != hipSuccess) return status_of(e);
  return (s->hcells.p != c0 || s->hperm.p != p0) ? GROWN : OK;
}

EXPORT uint64_t tree_info(uint32_t h, uint32_t which) {  // root.zig:113-126
  Sim* s = get(h);
  if (!s) return 0;
  if (refresh_tree_host(*s) != hipSuccess) return 0;
  const TreeMeta* m = s->tree.meta_host;
  const bool valid = m && m->valid;
  switch (which) {
    case 0: return valid ? m->len : 0;
    case 1: return valid ? 0 : NONE;
    case 2: return (uint64_t)(uintptr_t)s->hcells.p;
    case 3: return (uint64_t)(uintptr_t)s->hextent;
    case 4: return (uint64_t)(uintptr_t)s->hperm.p;
    case 5: return sizeof(Cell);
    case 6: return valid ? m->m : 0;
    default: return 0;
  }
}

EXPORT int32_t tree_order(uint32_t h, uint32_t after, uint32_t* out) {  // root.zig:130-137, tree.zig:652-684
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  if (!out) return BAD_ARGUMENT;
  if (refresh_tree_host(*s) != hipSuccess) return WORKER_FAILED;
  const TreeMeta* m = s->tree.meta_host;
  if (!m || !m->valid) return OK;
  std::vector<uint32_t> stack;
  stack.push_back(0);
  size_t w = 0;
  while (!stack.empty()) {
    const uint32_t c = stack.back();
    stack.pop_back();
    out[w++] = c;
    const uint32_t* ch = s->hcells.p[c].child;
    if (after) {
      for (int i = 0; i < 4; ++i) if (ch[i] != NONE) stack.push_back(ch[i]);
    } else {
      for (int i = 3; i >= 0; --i) if (ch[i] != NONE) stack.push_back(ch[i]);
    }
  }
  if (after) std::reverse(out, out + w);
  return OK;
}

EXPORT uint32_t force_create(uint32_t h, uint32_t kind) {  // root.zig:147-161
  Sim* s = get(h);
  if (!s || kind >= K_COUNT) return 0;
  Force* f = new (std::nothrow) Force((Kind)kind);
  if (!f) return 0;
  if (f->node_sized() && f->set_count(s->n) != hipSuccess) {
    delete f;
    return 0;
  }
  if (kind == K_CENTER) s->centers++;
  for (size_t i = 0; i < s->forces.size(); ++i)
    if (!s->forces[i]) {
      s->forces[i] = f;
      return (uint32_t)i + 1;
    }
  s->forces.push_back(f);
  return (uint32_t)s->forces.size();
}

EXPORT int32_t force_destroy(uint32_t h, uint32_t fh) {  // root.zig:163-170
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  (void)hipStreamSynchronize(s->stream);
  if (f->kind == K_CENTER) s->centers--;
  delete f;
  s->forces[fh - 1] = nullptr;
  return OK;
}

EXPORT int32_t force_set_count(uint32_t h, uint32_t fh, uint32_t count) {  // root.zig:174-179
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (f->set_count(count) != hipSuccess) return OUT_OF_MEMORY;
  return GROWN;
}

EXPORT void* force_ptr(uint32_t h, uint32_t fh, uint32_t which) {  // root.zig:181-186
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f || f->count == 0) return nullptr;
  f->dirty = true;
  f->sched_valid = false;
  f->csr_valid = false;
  switch (which) {
    case D3F_F0: return f->float_arrays() > 0 ? f->f[0].p : nullptr;
    case D3F_F1: return f->float_arrays() > 1 ? f->f[1].p : nullptr;
    case D3F_F2: return f->float_arrays() > 2 ? f->f[2].p : nullptr;
    case D3F_SOURCE: return f->kind == K_LINK ? f->source.p : nullptr;
    case D3F_TARGET: return f->kind == K_LINK ? f->target.p : nullptr;
    default: return nullptr;
  }
}

EXPORT int32_t force_set_scalar(uint32_t h, uint32_t fh, uint32_t which, double v) {  // root.zig:188-194
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (which >= S_COUNT) return BAD_ARGUMENT;
  f->scalars[which] = v;
  return OK;
}

EXPORT double force_get_scalar(uint32_t h, uint32_t fh, uint32_t which) {  // root.zig:196-200
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return 0;
  return which < S_COUNT ? f->scalars[which] : 0;
}

EXPORT int32_t force_run(uint32_t h, uint32_t fh, double alpha) {  // root.zig:211-242
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (f->node_sized() && f->count < s->n) return BAD_ARGUMENT;
  int32_t injected;
  if (injected_failure(&injected)) return injected;
  hipError_t e = upload(*s, *f);
  if (e != hipSuccess) return status_of(e);
  switch (f->kind) {
    case K_CENTER: return run_center(*s, *f);
    case K_X:
      return status_of(axis_force(s->dx.p, s->dvx.p, f->df[0].p, f->df[1].p, s->n, alpha, s->stream));
    case K_Y:
      return status_of(axis_force(s->dy.p, s->dvy.p, f->df[0].p, f->df[1].p, s->n, alpha, s->stream));
    case K_RADIAL:
      return status_of(radial_force(s->dx.p, s->dy.p, s->dvx.p, s->dvy.p, f->df[0].p, f->df[1].p, s->n,
                                    f->scalar(S_X), f->scalar(S_Y), alpha, s->stream));
    case K_LINK:
      return fast_mode(*s, *f) ? run_link_fast(*s, *f, alpha) : run_link_exact(*s, *f, alpha);
    case K_MANY_BODY:
      return fast_mode(*s, *f) ? run_many_body_fast(*s, *f, alpha) : run_many_body_exact(*s, *f, alpha);
    case K_COLLIDE:
      return fast_mode(*s, *f) ? run_collide_fast(*s, *f) : run_collide_exact(*s, *f);
    default:
      return BAD_ARGUMENT;
  }
}

EXPORT int32_t force_direct_sum(uint32_t h, uint32_t fh, double alpha, const uint32_t* sample, uint32_t k,
                                double* out) {
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (f->kind != K_MANY_BODY || f->count < s->n || (k && (!sample || !out))) return BAD_ARGUMENT;
  for (uint32_t q = 0; q < k; ++q)
    if (sample[q] >= s->n) return BAD_ARGUMENT;
  hipError_t e = upload(*s, *f);
  Device<uint32_t> dsample;
  Device<double> dout;
  if (e == hipSuccess) e = dsample.reserve(k > 0 ? k : 1);
  if (e == hipSuccess) e = dout.reserve(2 * (size_t)(k > 0 ? k : 1));
  if (e == hipSuccess && k) e = hipMemcpyAsync(dsample.p, sample, k * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess)
    e = direct_sum(s->dx.p, s->dy.p, f->df[0].p, s->n, dsample.p, k, alpha, f->scalar(S_DMIN2), f->scalar(S_DMAX2),
                   dout.p, s->stream);
  if (e == hipSuccess && k) e = hipMemcpyAsync(out, dout.p, 2 * k * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  dsample.release();
  dout.release();
  return status_of(e);
}

// ---- force_direct_sum without allocation: the staging of the sample
// indices and the outputs lives in `scratch`, a file-scope device array of
// fixed size (each device holds its own copy until the process ends). Host
// code never passes `scratch` itself to the device: its device address is
// taken with hipGetSymbolAddress once per device and kept in g_scratch_dev.
constexpr uint32_t STATIC_SAMPLES = 16384;  // the most samples one call takes
constexpr size_t SCRATCH_OUT = (size_t)STATIC_SAMPLES * 2 * sizeof(double);  // outputs first (256-aligned)
static __device__ __align__(256) unsigned char scratch[SCRATCH_OUT + (size_t)STATIC_SAMPLES * sizeof(uint32_t)];
constexpr int MAX_DEVICES = 64;
unsigned char* g_scratch_dev[MAX_DEVICES] = {nullptr};

// The scratch array's device address on the current device (set up on the
// first call there).
hipError_t scratch_address(unsigned char** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= MAX_DEVICES) HIPFAIL(hipErrorInvalidDevice, "device index beyond the scratch table");
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    HIPCHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = static_cast<unsigned char*>(ptr);
  }
  *out = g_scratch_dev[dev];
  return hipSuccess;
}

// upload() without allocation: the force's device arrays must already hold
// `count` entries (a force_run of the force sizes them).
int32_t upload_in_place(Sim& s, Force& f) {
  if (!f.dirty) return OK;
  for (int i = 0; i < f.float_arrays(); ++i)
    if (f.df[i].cap < f.count || !f.df[i].p) {
      std::snprintf(g_last_error, sizeof g_last_error,
                    "force_direct_sum_static: the force's parameters are not on the device yet (run the force "
                    "once, or call force_direct_sum, after changing its arrays)");
      return BAD_ARGUMENT;
    }
  for (int i = 0; i < f.float_arrays(); ++i) {
    if (!f.count) continue;
    hipError_t e = hipMemcpyAsync(f.df[i].p, f.f[i].p, f.count * sizeof(double), hipMemcpyHostToDevice, s.stream);
    if (e != hipSuccess) return status_of(e);
  }
  hipError_t e = hipStreamSynchronize(s.stream);
  if (e != hipSuccess) return status_of(e);
  f.dirty = false;
  f.sched_valid = false;
  return OK;
}

EXPORT int32_t force_direct_sum_static(uint32_t h, uint32_t fh, double alpha, const uint32_t* sample, uint32_t k,
                                       double* out) {
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  if (f->kind != K_MANY_BODY || f->count < s->n || (k && (!sample || !out))) return BAD_ARGUMENT;
  if (k > STATIC_SAMPLES) {
    std::snprintf(g_last_error, sizeof g_last_error,
                  "force_direct_sum_static: %u samples exceed its fixed scratch of %u (use force_direct_sum)", k,
                  STATIC_SAMPLES);
    return BAD_ARGUMENT;
  }
  for (uint32_t q = 0; q < k; ++q)
    if (sample[q] >= s->n) return BAD_ARGUMENT;
  const int32_t st = upload_in_place(*s, *f);
  if (st != OK) return st;
  if (k == 0) return OK;
  unsigned char* dev = nullptr;
  hipError_t e = scratch_address(&dev);
  double* dout = reinterpret_cast<double*>(dev);
  uint32_t* dsample = reinterpret_cast<uint32_t*>(dev + SCRATCH_OUT);
  if (e == hipSuccess) e = hipMemcpyAsync(dsample, sample, k * sizeof(uint32_t), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess)
    e = direct_sum(s->dx.p, s->dy.p, f->df[0].p, s->n, dsample, k, alpha, f->scalar(S_DMIN2), f->scalar(S_DMAX2),
                   dout, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out, dout, 2 * k * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  return status_of(e);
}

EXPORT int32_t sim_stats(uint32_t h, d3f_stats* out) {
  Sim* s = get(h);
  if (!s || !out) return BAD_HANDLE;
  hipError_t e = s->collect();
  if (e != hipSuccess) return status_of(e);
  *out = s->stats;
  s->stats = d3f_stats{};
  return OK;
}
