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
#include <deque>
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

// Values a run of captured ticks drew ahead of the eager order (seeds of
// ticks it then had to redo eagerly): the next draws take them first.
std::deque<double> g_pending;

double draw() {
  if (!g_pending.empty()) {
    const double v = g_pending.front();
    g_pending.pop_front();
    return v;
  }
  return g_random ? g_random(g_random_ctx) : 0.5;
}
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
  // captured ticks (sim_ticks): one tick as a HIP graph, replayed with
  // alpha and the seeds from tparams[tick]
  uint64_t epoch = 1;  // advanced by anything that may change what a captured tick holds
  TickState* dtick = nullptr;
  TickState* htick = nullptr;  // pinned
  Device<TickParams> tparams;
  Pinned<TickParams> hparams;
  Device<double> tsnap;  // the state at the start of the first failing replay
  hipGraph_t tgraph = nullptr;
  hipGraphExec_t texec = nullptr;
  hipEvent_t ev_tfork = nullptr, ev_tjoin = nullptr;
  std::vector<uint32_t> tforces;  // what the captured tick was captured for
  double tdecay = 0;
  uint64_t tepoch = 0;
  uint32_t tlevels = 0;
  uint64_t graph_ticks = 0, graph_redos = 0, graph_captures = 0;
  void drop_graph() {
    if (texec) (void)hipGraphExecDestroy(texec);
    if (tgraph) (void)hipGraphDestroy(tgraph);
    texec = nullptr;
    tgraph = nullptr;
  }

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
    drop_graph();
    if (ev_tfork) (void)hipEventDestroy(ev_tfork);
    if (ev_tjoin) (void)hipEventDestroy(ev_tjoin);
    if (dtick) (void)hipFree(dtick);
    if (htick) (void)hipHostFree(htick);
    tparams.release();
    hparams.release();
    tsnap.release();
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
    ++epoch;
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
  ++s.epoch;
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
  a.counters = g_timing ? s.dcounters : nullptr;  // work counters while timing (atomics on three words)
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
  for (double k = 0; k < iterations; k += 1)
    if ((e = link_levels(a, f.level_off.data(), levels, s.stream)) != hipSuccess) return status_of(e);
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

// One force's run (root.zig:211-242 after the argument checks).
int32_t run_force(Sim& s, Force& f, double alpha) {
  hipError_t e = upload(s, f);
  if (e != hipSuccess) return status_of(e);
  switch (f.kind) {
    case K_CENTER: return run_center(s, f);
    case K_X:
      return status_of(axis_force(s.dx.p, s.dvx.p, f.df[0].p, f.df[1].p, s.n, alpha, s.stream));
    case K_Y:
      return status_of(axis_force(s.dy.p, s.dvy.p, f.df[0].p, f.df[1].p, s.n, alpha, s.stream));
    case K_RADIAL:
      return status_of(radial_force(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, f.df[0].p, f.df[1].p, s.n, f.scalar(S_X),
                                    f.scalar(S_Y), alpha, s.stream));
    case K_LINK:
      return fast_mode(s, f) ? run_link_fast(s, f, alpha) : run_link_exact(s, f, alpha);
    case K_MANY_BODY:
      return fast_mode(s, f) ? run_many_body_fast(s, f, alpha) : run_many_body_exact(s, f, alpha);
    case K_COLLIDE:
      return fast_mode(s, f) ? run_collide_fast(s, f) : run_collide_exact(s, f);
    default:
      return BAD_ARGUMENT;
  }
}

hipError_t integrate_step(Sim& s, double velocity_decay) {  // simple.zig:90-123
  HIPCHECK(s.before_x_write());
  s.timer_start(D3F_PHASE_INTEGRATE);
  HIPCHECK(integrate(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, s.dfx.p, s.dfy.p, s.dfixed.p, s.n, velocity_decay, s.stream));
  s.timer_stop();
  return s.speculate_center();
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

// ---------------------------------------------------------------- captured ticks
// sim_ticks runs whole ticks (simulation.js:63-82 with the facade's forces
// in order, then integrate) as one captured HIP graph per tick. What a
// replay cannot finish (jiggle draws in the exact traversal or link pass, a
// tree or FMM beyond the captured buffers, an FMM that declines) is noted on
// the device; the host then restores the state from the start of the first
// such tick and runs that tick eagerly, drawing from the random source in
// the eager order.

enum TickFlag : uint32_t { TF_FMM_ABORT = 1, TF_FMM_DECLINED = 2, TF_TREE_STOP = 4, TF_JIGGLE = 8, TF_LINK_ZERO = 16 };

// Seeds are drawn per run for the seeded kinds in fast precision (core.js:213-231).
bool seeded(const Force& f) {
  return f.scalar(S_PRECISION) == 1 && (f.kind == K_MANY_BODY || f.kind == K_LINK || f.kind == K_COLLIDE);
}

// The eager tick: each force's run as force_run does it, then integrate.
int32_t eager_tick(Sim& s, const std::vector<Force*>& fs, double alpha, double decay) {
  for (Force* f : fs) {
    if (seeded(*f)) f->scalars[S_SEED] = draw();
    const int32_t r = run_force(s, *f, alpha);
    if (r < 0) return r;
  }
  return status_of(integrate_step(s, decay));
}

FmmArgs fast_args(Sim& s, Force& f, double alpha) {
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
  a.graphs = 1;
  a.node_graph = nullptr;
  return a;
}

// The device part of a tick in capturable form (no host synchronisation,
// nothing read back), on s.stream with forceCenter's sums of the starting
// positions on the side stream.
hipError_t graph_tick(Sim& s, const std::vector<Force*>& fs, double decay) {
  hipStream_t st = s.stream;
  const uint32_t n = s.n;
  const double* alpha_dev = &s.dtick->alpha;
  TickNotes notes{};
  auto note = [&](const uint32_t* flag, uint32_t bit) {
    notes.flag[notes.count] = flag;
    notes.bit[notes.count] = bit;
    ++notes.count;
  };
  HIPCHECK(tick_begin(s.dtick, s.tparams.p, s.dscalar_u, 2, st));
  HIPCHECK(tick_snapshot(s.dtick, s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, n, s.tsnap.p, st));
  // Only center and integrate write x, y: the first center's sums are those
  // of the tick's starting positions.
  bool side_sums = false;
  for (Force* f : fs) side_sums = side_sums || f->kind == K_CENTER;
  if (side_sums) {
    HIPCHECK(hipEventRecord(s.ev_tfork, st));
    HIPCHECK(hipStreamWaitEvent(s.side, s.ev_tfork, 0));
    HIPCHECK(center_raw_sum(s.dx.p, s.dy.p, n, s.dsums_side, s.cw_side, s.side));
    HIPCHECK(hipEventRecord(s.ev_tjoin, s.side));
  }
  int seed_slot = 0;
  for (Force* fp : fs) {
    Force& f = *fp;
    const uint64_t* seed_dev = seeded(f) ? &s.dtick->seed[seed_slot++] : nullptr;
    const bool fast = fast_mode(s, f);
    switch (f.kind) {
      case K_MANY_BODY:
        if (fast) {
          FmmArgs a = fast_args(s, f, 0);
          const uint32_t node_off[2] = {0, n};
          a.node_off = node_off;
          a.alpha_dev = alpha_dev;
          a.seeds_dev = seed_dev;
          a.graph = true;
          a.aux = s.aux;
          HIPCHECK(fmm_run(s.fmm, s.tree, a, st));
          note(&s.fmm.params->abort, TF_FMM_ABORT);
          note(&s.fmm.params->declined, TF_FMM_DECLINED);
        } else {
          TreeBuffers& t = s.tree;
          HIPCHECK(tree_build_exact(t, s.dx.p, s.dy.p, n, st, true));
          HIPCHECK(mb_accumulate(t, s.dx.p, s.dy.p, f.df[0].p, f.scalar(S_THETA2), st, true));
          TravArgs a{};
          a.tc = t.tcells;
          a.order = t.perm;
          a.count = n;  // bounds the grid; the walks take m from the tree
          a.perm = t.perm;
          a.x = s.dx.p;
          a.y = s.dy.p;
          a.s = f.df[0].p;
          a.vx = s.dvx.p;
          a.vy = s.dvy.p;
          a.dmin2 = f.scalar(S_DMIN2);
          a.dmax2 = f.scalar(S_DMAX2);
          a.alpha = 0;
          a.flagged = s.flagged.p;
          a.n_flagged = s.dscalar_u;
          a.len = t.cell_cap;
          a.counters = nullptr;  // the work counters are taken by the ticks run force by force
          a.meta = t.meta;
          a.alpha_dev = alpha_dev;
          HIPCHECK(mb_traverse(a, TRAV_RECORD, st));
          HIPCHECK(mb_nan_nodes_graph(t, n, s.dvx.p, s.dvy.p, st));
          note(&t.meta->stop, TF_TREE_STOP);
          note(s.dscalar_u, TF_JIGGLE);
        }
        break;
      case K_LINK: {
        const double iterations = f.scalar(S_ITERATIONS);
        if (f.count == 0 || !(0 < iterations)) break;
        if (fast) {
          for (double k = 0; k < iterations; k += 1) {
            HIPCHECK(shift_pack(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, s.dpos.p, n, st));
            HIPCHECK(link_sweep(f.csr_args(n), f.long_count, s.dpos.p, s.dvx.p, s.dvy.p, 0, 0, st, alpha_dev,
                                seed_dev));
          }
        } else {
          LinkLevelArgs a{f.lsrc.p, f.ltgt.p, f.lf[0].p, f.lf[1].p, f.lf[2].p, s.dx.p, s.dy.p, s.dvx.p, s.dvy.p,
                          s.dlnodes.p, n, 0, s.dscalar_u + 1};
          a.alpha_dev = alpha_dev;
          const uint32_t levels = (uint32_t)f.level_off.size() - 2;
          HIPCHECK(link_pack(a, st));
          for (double k = 0; k < iterations; k += 1)
            HIPCHECK(link_levels(a, f.level_off.data(), levels, st));
          HIPCHECK(link_unpack(a, st));
          note(s.dscalar_u + 1, TF_LINK_ZERO);
        }
        break;
      }
      case K_CENTER: {
        const double* sums = s.dsums_side;
        if (side_sums) {
          HIPCHECK(hipStreamWaitEvent(st, s.ev_tjoin, 0));
          side_sums = false;
        } else {
          sums = s.dscalar_d;
          HIPCHECK(center_raw_sum(s.dx.p, s.dy.p, n, s.dscalar_d, s.cw_main, st));
        }
        HIPCHECK(center_apply(s.dx.p, s.dy.p, n, sums, f.scalar(S_X), f.scalar(S_Y), f.scalar(S_STRENGTH), st));
        break;
      }
      case K_X:
        HIPCHECK(axis_force(s.dx.p, s.dvx.p, f.df[0].p, f.df[1].p, n, 0, st, alpha_dev));
        break;
      case K_Y:
        HIPCHECK(axis_force(s.dy.p, s.dvy.p, f.df[0].p, f.df[1].p, n, 0, st, alpha_dev));
        break;
      case K_RADIAL:
        HIPCHECK(radial_force(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, f.df[0].p, f.df[1].p, n, f.scalar(S_X),
                              f.scalar(S_Y), 0, st, alpha_dev));
        break;
      default:
        HIPFAIL(hipErrorInvalidValue, "a force kind the captured tick does not hold");
    }
  }
  HIPCHECK(integrate(s.dx.p, s.dy.p, s.dvx.p, s.dvy.p, s.dfx.p, s.dfy.p, s.dfixed.p, n, decay, st));
  return tick_end(s.dtick, notes, st);
}

// Sizes every buffer a replay uses from the eager tick just run (with room
// to grow), then captures the tick.
hipError_t capture_tick(Sim& s, const std::vector<Force*>& fs, const std::vector<uint32_t>& fhs, double decay) {
  s.drop_graph();
  const uint32_t n = s.n;
  if (!s.dtick) {
    HIPCHECK(hipMalloc(&s.dtick, sizeof(TickState)));
    HIPCHECK(hipHostMalloc((void**)&s.htick, sizeof(TickState), hipHostMallocDefault));
    HIPCHECK(hipEventCreateWithFlags(&s.ev_tfork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&s.ev_tjoin, hipEventDisableTiming));
  }
  HIPCHECK(s.tparams.reserve(TICK_RUN, 0, s.stream));
  HIPCHECK(s.hparams.reserve(TICK_RUN, 0));
  HIPCHECK(s.tsnap.reserve(4 * (size_t)s.cap, 0, s.stream));
  for (Force* f : fs) {
    if (f->kind != K_MANY_BODY) continue;
    if (fast_mode(s, *f)) {
      FmmArgs a = fast_args(s, *f, 0);
      const uint32_t node_off[2] = {0, n};
      a.node_off = node_off;
      a.seeds_dev = &s.dtick->seed[0];  // the replays' seeds (nothing taken from the host)
      const uint32_t cells = 2 * s.fmm.params_host->level_off[FMM_LEVELS + 1] + 4096;
      HIPCHECK(fmm_reserve_capacity(s.fmm, s.tree, a, cells, s.stream));
    } else {
      const uint32_t len = s.tree.meta_host ? s.tree.meta_host->len : 0;
      HIPCHECK(tree_reserve_cells(s.tree, std::max(len + len / 2 + 4096, 2 * n + 64)));
    }
  }
  // both center workspaces sized (the capture may use either)
  if (s.centers) {
    HIPCHECK(center_raw_sum(s.dx.p, s.dy.p, n, s.dscalar_d, s.cw_main, s.stream));
    HIPCHECK(center_raw_sum(s.dx.p, s.dy.p, n, s.dscalar_d, s.cw_side, s.stream));
  }
  HIPCHECK(hipStreamSynchronize(s.stream));
  HIPCHECK(hipStreamSynchronize(s.side));
  s.side_busy = false;
  s.side_valid = false;
  HIPCHECK(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = graph_tick(s, fs, decay);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(s.stream, &g);
  if (e != hipSuccess || e2 != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    return e != hipSuccess ? e : e2;
  }
  s.tgraph = g;
  HIPCHECK(hipGraphInstantiate(&s.texec, s.tgraph, nullptr, nullptr, 0));
  s.tforces = fhs;
  s.tdecay = decay;
  s.tepoch = s.epoch;
  s.tlevels = s.fmm.build_levels;
  ++s.graph_captures;
  return hipSuccess;
}

bool graph_matches(const Sim& s, const std::vector<uint32_t>& fhs, double decay) {
  return s.texec && s.tepoch == s.epoch && s.tforces == fhs && std::memcmp(&s.tdecay, &decay, sizeof decay) == 0 &&
         s.tlevels == s.fmm.build_levels;
}

// Runs `ticks` ticks; returns NOT_CAPTURED before doing anything when the
// forces or the library's state call for the per-force path.
constexpr int32_t NOT_CAPTURED = 2;
int32_t run_ticks(Sim& s, const uint32_t* fhs_in, uint32_t nf, uint32_t ticks, const double* alphas, double decay,
                  uint32_t* done) {
  if (g_timing || g_fail_count || s.n < 2 || nf == 0 || nf > 16) return NOT_CAPTURED;
  std::vector<uint32_t> fhs(fhs_in, fhs_in + nf);
  std::vector<Force*> fs;
  uint32_t bodies = 0, links = 0, seeds = 0;
  for (uint32_t fh : fhs) {
    Force* f = get_force(&s, fh);
    if (!f) return BAD_HANDLE;
    if (f->node_sized() && f->count < s.n) return BAD_ARGUMENT;
    if (f->kind == K_COLLIDE) return NOT_CAPTURED;
    bodies += f->kind == K_MANY_BODY;
    links += f->kind == K_LINK;
    seeds += seeded(*f);
    fs.push_back(f);
  }
  if (bodies > 1 || links > 1 || seeds > (uint32_t)TICK_SEEDS) return NOT_CAPTURED;
  int32_t r;
  uint32_t t = 0;
  bool use_graph = true;
  bool fresh = false;  // captured after the last eager tick, not yet replayed through
  std::vector<double> drawn;
  auto finish = [&](int32_t status) {
    g_pending.clear();
    if (done) *done = t;
    return status;
  };
  // After an eager tick: a capture sized by it (also after the run's last
  // tick, for the next call), unless this run has gone eager.
  auto recapture = [&]() -> int32_t {
    if (!use_graph) return OK;
    const hipError_t e = capture_tick(s, fs, fhs, decay);
    fresh = true;
    if (e == hipSuccess) return OK;
    s.drop_graph();
    use_graph = false;  // the rest of this run goes eagerly
    if (e == hipErrorOutOfMemory) return status_of(e);
    (void)hipGetLastError();
    return OK;
  };
  while (t < ticks) {
    if (!use_graph || !graph_matches(s, fhs, decay)) {
      if ((r = eager_tick(s, fs, alphas[t], decay)) < 0) return finish(r);
      ++t;
      if ((r = recapture()) < 0) return finish(r);
      continue;
    }
    const uint32_t chunk = std::min(ticks - t, TICK_RUN);
    drawn.clear();
    for (uint32_t k = 0; k < chunk; ++k) {
      TickParams& p = s.hparams.p[k];
      p.alpha = alphas[t + k];
      int slot = 0;
      for (Force* f : fs)
        if (seeded(*f)) {
          const double v = draw();
          drawn.push_back(v);
          std::memcpy(&p.seed[slot++], &v, sizeof v);
        }
    }
    TickState& h = *s.htick;
    h.tick = 0;
    h.first_fail = NONE;
    h.status = 0;
    hipError_t e = hipMemcpyAsync(s.tparams.p, s.hparams.p, chunk * sizeof(TickParams), hipMemcpyHostToDevice,
                                  s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.dtick, s.htick, sizeof(TickState), hipMemcpyHostToDevice, s.stream);
    for (uint32_t k = 0; k < chunk && e == hipSuccess; ++k) e = hipGraphLaunch(s.texec, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.htick, s.dtick, sizeof(TickState), hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) return finish(status_of(e));
    if (h.first_fail == NONE) {
      s.graph_ticks += chunk;
      t += chunk;
      fresh = false;
      continue;
    }
    // A replay could not finish its tick: the state from its start, then
    // that tick eagerly. The seeds drawn for it and the ticks after it go
    // back in front of the random source, so the eager draws see them in
    // the eager order.
    const uint32_t tf = h.first_fail;
    s.graph_ticks += tf;
    ++s.graph_redos;
    for (size_t k = drawn.size(); k-- > (size_t)tf * seeds;) g_pending.push_front(drawn[k]);
    const size_t bn = (size_t)s.n * sizeof(double);
    double* to[4] = {s.dvx.p, s.dvy.p, s.dx.p, s.dy.p};
    for (int k = 0; k < 4 && e == hipSuccess; ++k)
      e = hipMemcpyAsync(to[k], s.tsnap.p + (size_t)k * s.n, bn, hipMemcpyDeviceToDevice, s.stream);
    if (e != hipSuccess) return finish(status_of(e));
    t += tf;
    s.side_busy = false;
    s.side_valid = false;
    s.drop_graph();  // the eager tick may grow buffers the graph holds
    if ((r = eager_tick(s, fs, alphas[t], decay)) < 0) return finish(r);
    ++t;
    // A tick a fresh capture could not replay at once would likely fail
    // again: the rest of the run goes eagerly.
    if (fresh && tf == 0) use_graph = false;
    if ((r = recapture()) < 0) return finish(r);
  }
  if (s.texec && graph_matches(s, fhs, decay)) {
    // as after an eager integrate: the next center's sums start now
    s.side_busy = false;
    s.side_valid = false;
    const hipError_t e = s.speculate_center();
    if (e != hipSuccess) return finish(status_of(e));
    for (Force* f : fs)
      if (f->kind == K_MANY_BODY && !fast_mode(s, *f)) s.charges_pending = true;
  }
  return finish(OK);
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

EXPORT int32_t sim_ticks(uint32_t h, const uint32_t* forces, uint32_t nforces, uint32_t ticks, const double* alphas,
                         double velocity_decay, uint32_t* done) {  // simulation.js:63-82
  Sim* s = get(h);
  if (done) *done = 0;
  if (!s) return BAD_HANDLE;
  if (ticks == 0) return OK;
  if (!forces || !alphas) return BAD_ARGUMENT;
  return run_ticks(*s, forces, nforces, ticks, alphas, velocity_decay, done);
}

EXPORT int32_t sim_tick_counters(uint32_t h, uint64_t* out) {
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  if (!out) return BAD_ARGUMENT;
  out[0] = s->graph_captures;
  out[1] = s->graph_ticks;
  out[2] = s->graph_redos;
  return OK;
}

EXPORT int32_t sim_integrate(uint32_t h, double velocity_decay) {  // root.zig:250-254
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  int32_t injected;
  if (injected_failure(&injected)) return injected;
  ++s->epoch;
  return status_of(integrate_step(*s, velocity_decay));
}

EXPORT int32_t tree_build(uint32_t h, uint32_t shifted) {  // root.zig:102-108
  Sim* s = get(h);
  if (!s) return BAD_HANDLE;
  const void* c0 = s->hcells.p;
  const void* p0 = s->hperm.p;
  hipError_t e;
  ++s->epoch;
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
  ++s->epoch;
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
  ++s->epoch;
  delete f;
  s->forces[fh - 1] = nullptr;
  return OK;
}

EXPORT int32_t force_set_count(uint32_t h, uint32_t fh, uint32_t count) {  // root.zig:174-179
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f) return BAD_HANDLE;
  ++s->epoch;
  if (f->set_count(count) != hipSuccess) return OUT_OF_MEMORY;
  return GROWN;
}

EXPORT void* force_ptr(uint32_t h, uint32_t fh, uint32_t which) {  // root.zig:181-186
  Sim* s = get(h);
  Force* f = get_force(s, fh);
  if (!f || f->count == 0) return nullptr;
  ++s->epoch;
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
  // a captured tick holds every scalar but the seed (it reads the seed on the device)
  if (which != S_SEED && std::memcmp(&f->scalars[which], &v, sizeof v) != 0) ++s->epoch;
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
  ++s->epoch;
  return run_force(*s, *f, alpha);
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
