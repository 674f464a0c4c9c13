// Batched fast mode: many independent d3 simulations ("graphs") in one node
// arena and one link arena, ticked together on one MI355X. The reference has
// no counterpart (it runs one simulation per core instance); per graph, a
// tick is the facade's fast tick with forceManyBody, forceLink and
// forceCenter in that order, then integrate (simulation.js:49-89):
//  - many-body: one forest FMM over every graph (fmm.hip): one key sort, one
//    level-by-level build and one dual-tree traversal started from every
//    root; a graph's cells, interaction lists and sums are the ones a run of
//    that graph alone builds;
//  - link: the adjacency sweep (link_fast.zig) over the arena's rows;
//  - center: each graph's serial sums (simple.zig:14-31), one wave a graph;
//  - integrate: per node.
// A graph below the fast-mode crossover (n < 2000, root.zig:246-248) runs
// manyBody and link with the exact kernels instead, as the reference does,
// one workgroup per graph (batch_exact.hip). Each graph has its own random
// source, d3's lcg (lcg.js), from which the facade's draws are taken: per
// tick the manyBody seed and the link seed (core.js:218-222) and, for a small
// graph, the exact kernels' jiggle draws; so a graph's result does not depend
// on the other graphs of the batch.
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

enum BatchStatus : int32_t { B_OK = 0, B_BAD_HANDLE = -1, B_OUT_OF_MEMORY = -2, B_BAD_ARGUMENT = -3, B_FAILED = -4 };

int32_t bstatus(hipError_t e) {
  publish_failure(e);
  if (e == hipSuccess) return B_OK;
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return B_OUT_OF_MEMORY;
  return B_FAILED;
}

constexpr uint32_t BATCH_MIN_N = 2000;  // fmm.zig:43, the fast-mode crossover

// d3's lcg (lcg.js): s = (1664525 s + 1013904223) mod 2^32, random() = s / 2^32.
__device__ inline double batch_lcg(uint32_t& s) {
  s = 1664525u * s + 1013904223u;
  return (double)s / 4294967296.0;
}

// Per-tick scalars on the device (so that a captured tick reads them):
// alpha (simulation.js:64) and the tick counter, advanced by k_batch_step at
// the start of every tick with the host's arithmetic; `used` is the tick the
// seeds are mixed with.
struct BatchState {
  double alpha;
  unsigned long long tick, used;
};

__global__ void k_batch_step(BatchState* __restrict__ st, double target, double decay) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  st->alpha += (target - st->alpha) * decay;
  st->used = st->tick;
  st->tick += 1;
}

// The fast-mode seeds of one tick for the graphs at or above the crossover:
// manyBody's then link's, each one draw of the graph's random source taken as
// the seed's bits (core.js:218-222, hash.zig:28-30). The small graphs' draws
// are taken by the exact kernels, in the facade's order.
__global__ __launch_bounds__(256) void k_batch_seeds(const uint32_t* __restrict__ node_off, uint32_t graphs,
                                                     uint32_t* __restrict__ lcg, uint64_t* __restrict__ mb,
                                                     uint64_t* __restrict__ link) {
  const uint32_t g = blockIdx.x * 256 + threadIdx.x;
  if (g >= graphs || node_off[g + 1] - node_off[g] < BATCH_MIN_N) return;
  uint32_t s = lcg[g];
  mb[g] = __builtin_bit_cast(uint64_t, batch_lcg(s));
  link[g] = __builtin_bit_cast(uint64_t, batch_lcg(s));
  lcg[g] = s;
}

struct Batch {
  uint32_t graphs = 0, n = 0, m = 0;
  std::vector<uint32_t> node_off, link_off, node_graph;
  Pinned<double> x, y, vx, vy, charge, lstrength, ldistance, lbias;
  Pinned<uint32_t> source, target;
  Pinned<uint64_t> seed_ids, seeds_mb, seeds_link;
  Device<double> dx, dy, dvx, dvy, dcharge, dls, dld, dlb, dzero, dsums;
  Device<double2> dpos;  // x + vx, y + vy as pairs (link sweep)
  Device<uint8_t> dfixed;
  Device<uint32_t> dsrc, dtgt, dnode_graph, dnode_off, dlink_off;
  Device<uint64_t> dseeds_link, dseeds_mb;
  Device<uint32_t> dlcg, snap_lcg;  // per graph: the random state, and its copy from before the tick
  // the graphs below the crossover, ticked by the exact kernels
  std::vector<SmallJob> small;
  Device<SmallJob> dsmall;
  Device<uint8_t> scells;
  Device<TCell> stcells;
  Device<double> vout_x, vout_y;
  Device<SmallLink> slinks;     // their link schedules (k_small_schedule)
  Device<uint32_t> slevels, sorder;
  uint32_t* derror = nullptr;  // device: a small graph's tree outgrew its kernel
  uint32_t* herror = nullptr;  // pinned mirror
  BatchState* dstate = nullptr;
  bool state_dirty = true;     // the host's alpha / tick must be uploaded before the next tick
  // captured tick (D3F_BS_GRAPH)
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool graph_valid = false;
  Device<double> snap;         // x, y, vx, vy from before the tick (replays; eager ticks with small graphs)
  uint64_t graph_replays = 0, graph_redos = 0;
  // link rows
  bool csr_valid = false;
  Device<uint32_t> row, adj, owner, rows_long;
  Device<LinkEntry> entries;
  Device<double> terms;
  uint32_t long_count = 0;
  Device<uint64_t> sk0, sk1;
  Device<uint32_t> sv0, sv1, shist, sglob, scount;
  TreeBuffers tree;
  FmmBuffers fmm;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // the small graphs' kernel, beside the FMM
  hipEvent_t fork = nullptr, join = nullptr;
  double scalars[D3F_BS_COUNT] = {0};
  d3f_stats stats{};
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> pending;

  ~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto* a : {&x, &y, &vx, &vy, &charge, &lstrength, &ldistance, &lbias}) a->release();
    for (auto* a : {&source, &target}) a->release();
    for (auto* a : {&seed_ids, &seeds_mb, &seeds_link}) a->release();
    for (auto* a : {&dx, &dy, &dvx, &dvy, &dcharge, &dls, &dld, &dlb, &dzero, &dsums, &terms}) a->release();
    dpos.release();
    dfixed.release();
    for (auto* a : {&dsrc, &dtgt, &dnode_graph, &dnode_off, &dlink_off, &row, &adj, &owner, &rows_long, &sv0, &sv1,
                    &shist, &sglob, &scount})
      a->release();
    for (auto* a : {&dseeds_link, &dseeds_mb, &sk0, &sk1}) a->release();
    for (auto* a : {&dlcg, &snap_lcg}) a->release();
    dsmall.release();
    scells.release();
    stcells.release();
    for (auto* a : {&vout_x, &vout_y}) a->release();
    slinks.release();
    for (auto* a : {&slevels, &sorder}) a->release();
    if (derror) (void)hipFree(derror);
    if (herror) (void)hipHostFree(herror);
    snap.release();
    drop_graph();
    if (dstate) (void)hipFree(dstate);
    entries.release();
    tree_free(tree);
    fmm_free(fmm);
    for (auto& p : pending) {
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_valid = false;
  }

  hipError_t init() {
    HIPCHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    HIPCHECK(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
    HIPCHECK(hipEventCreateWithFlags(&join, hipEventDisableTiming));
    HIPCHECK(hipMalloc(&dstate, sizeof(BatchState)));
    HIPCHECK(hipMalloc(&derror, sizeof(uint32_t)));
    HIPCHECK(hipHostMalloc((void**)&herror, sizeof(uint32_t), hipHostMallocDefault));
    HIPCHECK(hipMemset(derror, 0, sizeof(uint32_t)));
    scalars[D3F_BS_THETA2] = 0.81;  // manyBody.js:9, theta 0.9
    HIPCHECK(sglob.reserve(8 * 256));
    HIPCHECK(scount.reserve(4));
    // d3's defaults (simulation.js:18-31, manyBody.js:5-12, center.js:4-7)
    scalars[D3F_BS_ALPHA] = 1;
    scalars[D3F_BS_ALPHA_DECAY] = 1 - std::pow(0.001, 1.0 / 300);
    scalars[D3F_BS_ALPHA_TARGET] = 0;
    scalars[D3F_BS_VELOCITY_DECAY] = 0.6;
    scalars[D3F_BS_DISTANCE_MIN2] = 1;
    scalars[D3F_BS_DISTANCE_MAX2] = __builtin_huge_val();
    scalars[D3F_BS_ORDER] = 8;
    scalars[D3F_BS_OPENING] = 1.5;
    scalars[D3F_BS_CENTER_STRENGTH] = 1;
    scalars[D3F_BS_LINK_ITERATIONS] = 1;
    return hipSuccess;
  }

  hipError_t set_graphs(uint32_t g, const uint32_t* noff, const uint32_t* loff) {
    HIPCHECK(hipStreamSynchronize(stream));
    graphs = g;
    node_off.assign(noff, noff + g + 1);
    link_off.assign(loff, loff + g + 1);
    n = node_off[g];
    m = link_off[g];
    node_graph.resize(n);
    for (uint32_t k = 0; k < g; ++k)
      for (uint32_t i = node_off[k]; i < node_off[k + 1]; ++i) node_graph[i] = k;
    const size_t cn = n > 0 ? n : 1, cm = m > 0 ? m : 1, cg = g > 0 ? g : 1;
    for (auto* a : {&x, &y, &vx, &vy, &charge}) HIPCHECK(a->reserve(cn, 0));
    for (auto* a : {&lstrength, &ldistance, &lbias}) HIPCHECK(a->reserve(cm, 0));
    for (auto* a : {&source, &target}) HIPCHECK(a->reserve(cm, 0));
    for (auto* a : {&seed_ids, &seeds_mb, &seeds_link}) HIPCHECK(a->reserve(cg, 0));
    for (uint32_t k = 0; k < g; ++k) seed_ids.p[k] = 1;  // lcg.js starts at s = 1
    for (auto* a : {&dx, &dy, &dvx, &dvy, &dcharge, &dzero}) HIPCHECK(a->reserve(cn, 0, stream));
    HIPCHECK(dpos.reserve(cn, 0, stream));
    for (auto* a : {&dls, &dld, &dlb}) HIPCHECK(a->reserve(cm, 0, stream));
    HIPCHECK(dsums.reserve(2 * cg, 0, stream));
    HIPCHECK(dfixed.reserve(cn, 0, stream));
    HIPCHECK(dsrc.reserve(cm, 0, stream));
    HIPCHECK(dtgt.reserve(cm, 0, stream));
    HIPCHECK(dnode_graph.reserve(cn, 0, stream));
    HIPCHECK(dnode_off.reserve(cg + 1, 0, stream));
    HIPCHECK(dlink_off.reserve(cg + 1, 0, stream));
    HIPCHECK(dseeds_link.reserve(cg, 0, stream));
    HIPCHECK(dseeds_mb.reserve(cg, 0, stream));
    HIPCHECK(dlcg.reserve(cg, 0, stream));
    HIPCHECK(snap_lcg.reserve(cg, 0, stream));
    // the exact kernels' jobs: graphs below the crossover, 3 n + 128 cells each
    // and their link schedules (built with the link rows): m records and
    // m + 1 level starts each
    small.clear();
    uint32_t cells = 0, recs = 0, levs = 0;
    for (uint32_t k = 0; k < g; ++k) {
      const uint32_t nk = node_off[k + 1] - node_off[k];
      if (nk >= BATCH_MIN_N) continue;
      const uint32_t mk = link_off[k + 1] - link_off[k];
      SmallJob j{node_off[k], nk, link_off[k], mk, cells, 3 * nk + 128, k, recs, levs, 0, {0, 0}};
      small.push_back(j);
      cells += j.cell_cap;
      recs += mk;
      levs += mk + 1;
    }
    HIPCHECK(dsmall.reserve(small.empty() ? 1 : small.size(), 0, stream));
    HIPCHECK(scells.reserve((size_t)(cells > 0 ? cells : 1) * small_cell_bytes(), 0, stream));
    HIPCHECK(stcells.reserve(cells > 0 ? cells : 1, 0, stream));
    HIPCHECK(slinks.reserve(recs > 0 ? recs : 1, 0, stream));
    HIPCHECK(sorder.reserve(recs > 0 ? recs : 1, 0, stream));
    HIPCHECK(slevels.reserve(levs > 0 ? levs : 1, 0, stream));
    for (auto* a : {&vout_x, &vout_y}) HIPCHECK(a->reserve(cn, 0, stream));
    if (!small.empty())
      HIPCHECK(hipMemcpyAsync(dsmall.p, small.data(), small.size() * sizeof(SmallJob), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemsetAsync(dfixed.p, 0, cn, stream));
    HIPCHECK(hipMemsetAsync(dzero.p, 0, cn * sizeof(double), stream));
    if (n) HIPCHECK(hipMemcpyAsync(dnode_graph.p, node_graph.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dnode_off.p, node_off.data(), (g + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dlink_off.p, link_off.data(), (g + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    {
      std::vector<uint32_t> s1(g, 1u);
      HIPCHECK(hipMemcpyAsync(dlcg.p, s1.data(), (size_t)g * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
      HIPCHECK(hipStreamSynchronize(stream));
    }
    csr_valid = false;
    drop_graph();
    return hipStreamSynchronize(stream);
  }

  LinkCsrArgs csr_args() {
    LinkCsrArgs a;
    a.src = dsrc.p;
    a.tgt = dtgt.p;
    a.strength = dls.p;
    a.distance = dld.p;
    a.bias = dlb.p;
    a.m = m;
    a.n = n;
    a.row = row.p;
    a.adj = adj.p;
    a.owner = owner.p;
    a.rows_long = rows_long.p;
    a.entries = entries.p;
    a.terms = terms.p;
    a.node_graph = dnode_graph.p;
    a.link_base = dlink_off.p;
    a.seeds = dseeds_link.p;
    return a;
  }

  hipError_t build_rows() {
    const size_t c = m > 0 ? m : 1;
    for (auto* a : {&sk0, &sk1}) HIPCHECK(a->reserve(2 * c, 0, stream));
    for (auto* a : {&sv0, &sv1}) HIPCHECK(a->reserve(2 * c, 0, stream));
    HIPCHECK(shist.reserve(sort_hist_words((uint32_t)(2 * c)), 0, stream));
    HIPCHECK(row.reserve((size_t)n + 1, 0, stream));
    HIPCHECK(adj.reserve(2 * c, 0, stream));
    HIPCHECK(owner.reserve(2 * c, 0, stream));
    HIPCHECK(entries.reserve(2 * c, 0, stream));
    HIPCHECK(terms.reserve(4 * c, 0, stream));
    HIPCHECK(rows_long.reserve(n > 0 ? n : 1, 0, stream));
    SortBuffers sb;
    sb.k0 = sk0.p;
    sb.k1 = sk1.p;
    sb.v0 = sv0.p;
    sb.v1 = sv1.p;
    sb.hist = shist.p;
    sb.glob = sglob.p;
    sb.count = scount.p;
    sb.cap = (uint32_t)(2 * c);
    HIPCHECK(link_csr(csr_args(), sb, &long_count, stream));
    HIPCHECK(small_schedule(dsmall.p, (uint32_t)small.size(), dsrc.p, dtgt.p, dls.p, dld.p, dlb.p, sorder.p,
                            slevels.p, slinks.p, stream));
    csr_valid = true;
    return hipSuccess;
  }

  void timer(int phase, hipEvent_t a, hipStream_t on = nullptr) {
    hipEvent_t b = nullptr;
    if (!a || hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, on ? on : stream);
    pending.push_back({phase, {a, b}});
  }
  hipEvent_t mark(hipStream_t on = nullptr) {
    hipEvent_t a = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return nullptr;
    (void)hipEventRecord(a, on ? on : stream);
    return a;
  }

  // x, y, vx, vy and the random states back to their copies from before the
  // tick; the host's alpha and tick count are uploaded again.
  hipError_t restore() {
    const size_t bn = (size_t)n * sizeof(double);
    double* to[4] = {dx.p, dy.p, dvx.p, dvy.p};
    for (int k = 0; k < 4; ++k) HIPCHECK(hipMemcpyAsync(to[k], snap.p + (size_t)k * n, bn, hipMemcpyDeviceToDevice, stream));
    HIPCHECK(hipMemcpyAsync(dlcg.p, snap_lcg.p, (size_t)graphs * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream));
    const BatchState hs{scalars[D3F_BS_ALPHA], (unsigned long long)scalars[D3F_BS_TICK], 0ull};
    HIPCHECK(hipMemcpyAsync(dstate, &hs, sizeof(hs), hipMemcpyHostToDevice, stream));
    return hipStreamSynchronize(stream);
  }

  // A small graph's tree outgrew the exact kernel: after restore(), the flag
  // is cleared and the tick fails with its cause.
  hipError_t small_overflow() {
    HIPCHECK(hipMemsetAsync(derror, 0, sizeof(uint32_t), stream));
    HIPCHECK(hipStreamSynchronize(stream));
    note_failure(hipErrorNotSupported,
                 "batch_tick: a graph below 2000 nodes needs a quadtree deeper than the batched exact kernel "
                 "holds (about 60 levels) or more than 3 n + 128 cells; the tick did not run",
                 __FILE__, __LINE__);
    return hipErrorNotSupported;
  }

  // Copies of x, y, vx, vy and the random states from before the tick.
  hipError_t snapshot() {
    const size_t bn = (size_t)n * sizeof(double);
    const double* from[4] = {dx.p, dy.p, dvx.p, dvy.p};
    for (int k = 0; k < 4; ++k)
      HIPCHECK(hipMemcpyAsync(snap.p + (size_t)k * n, from[k], bn, hipMemcpyDeviceToDevice, stream));
    return hipMemcpyAsync(snap_lcg.p, dlcg.p, (size_t)graphs * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
  }

  // The device part of a tick. graph: the capturable form (no host
  // synchronisation; nothing the host reads back). The small graphs' exact
  // kernel runs on the side stream while the FMM (which leaves their nodes
  // alone) runs on the main one; the link sweep waits for both.
  hipError_t device_tick(bool graph) {
    hipError_t e;
    hipLaunchKernelGGL(k_batch_step, dim3(1), dim3(64), 0, stream, dstate, scalars[D3F_BS_ALPHA_TARGET],
                       scalars[D3F_BS_ALPHA_DECAY]);
    hipLaunchKernelGGL(k_batch_seeds, dim3(ceil_div(graphs, 256u)), dim3(256), 0, stream, dnode_off.p, graphs, dlcg.p,
                       dseeds_mb.p, dseeds_link.p);
    const bool timed = g_batch_timing && !graph;
    hipEvent_t ev = nullptr;
    // many-body; the small graphs' exact manyBody and link (from the tick's
    // starting state, which the FMM leaves alone for their nodes) start on
    // the side stream once the FMM's dual traversal is done, beside its field
    // passes rather than beside its sort and its latency-bound rounds. fork is
    // recorded here first, so a FMM that returns before its own record still
    // orders them after the snapshot and the last tick's integrate.
    FmmArgs a = fmm_args();
    a.graph = graph;
    if (!small.empty()) {
      a.built = fork;
      HIPCHECK(hipEventRecord(fork, stream));
    }
    hipEvent_t evf = timed ? mark() : nullptr;
    if ((e = fmm_run(fmm, tree, a, stream)) != hipSuccess) return e;
    timer(D3F_PHASE_FMM, evf);
    if (!small.empty()) {
      HIPCHECK(hipStreamWaitEvent(side, fork, 0));
      ev = timed ? mark(side) : nullptr;
      SmallExactArgs sa;
      sa.x = dx.p;
      sa.y = dy.p;
      sa.vx = dvx.p;
      sa.vy = dvy.p;
      sa.charge = dcharge.p;
      sa.src = dsrc.p;
      sa.tgt = dtgt.p;
      sa.lstrength = dls.p;
      sa.ldistance = dld.p;
      sa.lbias = dlb.p;
      sa.lrec = slinks.p;
      sa.levels = slevels.p;
      sa.alpha_dev = &dstate->alpha;
      sa.theta2 = scalars[D3F_BS_THETA2];
      sa.dmin2 = scalars[D3F_BS_DISTANCE_MIN2];
      sa.dmax2 = scalars[D3F_BS_DISTANCE_MAX2];
      sa.link_iterations = scalars[D3F_BS_LINK_ITERATIONS];
      sa.lcg = dlcg.p;
      sa.cells = scells.p;
      sa.tcells = stcells.p;
      sa.vout_x = vout_x.p;
      sa.vout_y = vout_y.p;
      sa.error = derror;
      HIPCHECK(small_exact(dsmall.p, (uint32_t)small.size(), sa, side));
      timer(D3F_PHASE_TRAVERSAL, ev, side);
      HIPCHECK(hipEventRecord(join, side));
    }
    if (!small.empty()) HIPCHECK(hipStreamWaitEvent(stream, join, 0));
    // link
    ev = timed ? mark() : nullptr;
    for (double k = 0; k < scalars[D3F_BS_LINK_ITERATIONS]; k += 1) {
      HIPCHECK(shift_pack(dx.p, dy.p, dvx.p, dvy.p, dpos.p, n, stream));
      HIPCHECK(link_sweep(csr_args(), long_count, dpos.p, dvx.p, dvy.p, scalars[D3F_BS_ALPHA], 0, stream,
                          &dstate->alpha));
    }
    // the small graphs take their exact velocities
    HIPCHECK(small_scatter(dsmall.p, (uint32_t)small.size(), vout_x.p, vout_y.p, dvx.p, dvy.p, stream));
    timer(D3F_PHASE_LINK, ev);
    // center
    ev = timed ? mark() : nullptr;
    HIPCHECK(batch_center(dx.p, dy.p, dnode_off.p, dnode_graph.p, graphs, n, dsums.p, scalars[D3F_BS_CENTER_X],
                          scalars[D3F_BS_CENTER_Y], scalars[D3F_BS_CENTER_STRENGTH], stream));
    timer(D3F_PHASE_CENTER, ev);
    // integrate
    ev = timed ? mark() : nullptr;
    HIPCHECK(integrate(dx.p, dy.p, dvx.p, dvy.p, dzero.p, dzero.p, dfixed.p, n, scalars[D3F_BS_VELOCITY_DECAY], stream));
    timer(D3F_PHASE_INTEGRATE, ev);
    return hipSuccess;
  }

  FmmArgs fmm_args() {
    FmmArgs a;
    a.x = dx.p;
    a.y = dy.p;
    a.strength = dcharge.p;
    a.vx = dvx.p;
    a.vy = dvy.p;
    a.n = n;
    const double order = scalars[D3F_BS_ORDER];
    a.order = order >= 1 && order <= (double)FMM_MAX_ORDER ? (uint32_t)order : 8;
    a.alpha = scalars[D3F_BS_ALPHA];
    a.alpha_dev = &dstate->alpha;
    a.opening = scalars[D3F_BS_OPENING];
    a.dmin2 = scalars[D3F_BS_DISTANCE_MIN2];
    a.dmax2 = scalars[D3F_BS_DISTANCE_MAX2];
    a.graphs = graphs;
    a.node_off = node_off.data();
    a.node_graph = dnode_graph.p;
    a.seeds = nullptr;
    a.seeds_dev = dseeds_mb.p;
    const double limit = scalars[D3F_BS_GRAPH_CELLS];
    a.cell_limit = limit >= 1 && limit < 4294967296.0 ? (uint32_t)limit : 0;
    a.min_graph_n = BATCH_MIN_N;  // the small graphs run the exact kernels
    return a;
  }

  // Captures one tick (buffers sized for twice the cells of the last eager
  // tick): the state snapshot, then the device part of a tick.
  hipError_t capture() {
    drop_graph();
    HIPCHECK(snap.reserve(4 * (size_t)(n > 0 ? n : 1), 0, stream));
    const uint32_t cells = 2 * fmm.params_host->level_off[FMM_LEVELS + 1] + 4096;
    HIPCHECK(fmm_reserve_capacity(fmm, tree, fmm_args(), cells, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = snapshot();
    if (e == hipSuccess) e = device_tick(true);
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return e != hipSuccess ? e : e2;
    }
    graph = g;
    HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
    graph_valid = true;
    return hipSuccess;
  }

  int32_t tick() {
    hipError_t e;
    if (!csr_valid) {
      drop_graph();
      if ((e = build_rows()) != hipSuccess) return bstatus(e);
    }
    if (state_dirty) {
      const BatchState hs{scalars[D3F_BS_ALPHA], (unsigned long long)scalars[D3F_BS_TICK], 0ull};
      if ((e = hipMemcpy(dstate, &hs, sizeof(hs), hipMemcpyHostToDevice)) != hipSuccess) return bstatus(e);
      state_dirty = false;
    }
    const bool want_graph = scalars[D3F_BS_GRAPH] != 0;
    bool done = false;
    if (want_graph && graph_valid) {
      if ((e = hipGraphLaunch(gexec, stream)) != hipSuccess) return bstatus(e);
      if ((e = hipMemcpyAsync(fmm.params_host, fmm.params, sizeof(FmmParams), hipMemcpyDeviceToHost, stream)) !=
          hipSuccess)
        return bstatus(e);
      *herror = 0;
      if (!small.empty() &&
          (e = hipMemcpyAsync(herror, derror, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess)
        return bstatus(e);
      if ((e = hipStreamSynchronize(stream)) != hipSuccess) return bstatus(e);
      ++graph_replays;
      if (!fmm.params_host->abort && !*herror) {
        done = true;
      } else {  // sizes outgrew the captured buffers, or a small graph its kernel: undo the replay
        const bool failed = *herror != 0;
        ++graph_redos;
        if ((e = restore()) != hipSuccess) return bstatus(e);
        drop_graph();
        if (failed) return bstatus(small_overflow());  // the eager tick would fail the same way
      }
    }
    if (!done) {
      // with small graphs, the copy that a failed tick goes back to
      if (!small.empty() && ((e = snap.reserve(4 * (size_t)(n > 0 ? n : 1), 0, stream)) != hipSuccess ||
                             (e = snapshot()) != hipSuccess))
        return bstatus(e);
      if ((e = device_tick(false)) != hipSuccess) return bstatus(e);
      if (!small.empty()) {
        if ((e = hipMemcpyAsync(herror, derror, sizeof(uint32_t), hipMemcpyDeviceToHost, stream)) != hipSuccess ||
            (e = hipStreamSynchronize(stream)) != hipSuccess)
          return bstatus(e);
        if (*herror) {
          if ((e = restore()) != hipSuccess) return bstatus(e);
          return bstatus(small_overflow());
        }
      }
      if (want_graph && !graph_valid && !fmm.params_host->declined && (e = capture()) != hipSuccess)
        return bstatus(e);
    }
    if (fmm.params_host->declined) return B_BAD_ARGUMENT;  // a graph with a non-finite extent
    stats.m2l += fmm.params_host->m2l;
    stats.p2p_pairs += fmm.params_host->pairs;
    stats.fmm_cells = fmm.params_host->level_off[FMM_LEVELS + 1];
    double& alpha = scalars[D3F_BS_ALPHA];
    alpha += (scalars[D3F_BS_ALPHA_TARGET] - alpha) * scalars[D3F_BS_ALPHA_DECAY];  // simulation.js:64
    scalars[D3F_BS_TICK] += 1;
    return B_OK;
  }

  static bool g_batch_timing;
};
bool Batch::g_batch_timing = false;

std::vector<Batch*> g_batches;

Batch* bget(uint32_t h) {
  if (h == 0 || h > g_batches.size()) return nullptr;
  return g_batches[h - 1];
}

}  // namespace
}  // namespace d3f

using namespace d3f;

EXPORT uint32_t batch_create(void) {
  int devices = 0;
  if (hipGetDeviceCount(&devices) != hipSuccess || devices == 0) return 0;
  Batch* b = new (std::nothrow) Batch();
  if (!b) return 0;
  if (b->init() != hipSuccess) {
    delete b;
    return 0;
  }
  for (size_t i = 0; i < g_batches.size(); ++i)
    if (!g_batches[i]) {
      g_batches[i] = b;
      return (uint32_t)i + 1;
    }
  g_batches.push_back(b);
  return (uint32_t)g_batches.size();
}

EXPORT int32_t batch_destroy(uint32_t h) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  delete b;
  g_batches[h - 1] = nullptr;
  return B_OK;
}

EXPORT int32_t batch_set_graphs(uint32_t h, uint32_t graphs, const uint32_t* node_off, const uint32_t* link_off) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  if (graphs == 0 || !node_off || !link_off || node_off[0] != 0 || link_off[0] != 0) return B_BAD_ARGUMENT;
  for (uint32_t g = 0; g < graphs; ++g)
    if (node_off[g + 1] < node_off[g] || link_off[g + 1] < link_off[g]) return B_BAD_ARGUMENT;
  return bstatus(b->set_graphs(graphs, node_off, link_off));
}

EXPORT void* batch_ptr(uint32_t h, uint32_t which) {
  Batch* b = bget(h);
  if (!b || b->graphs == 0) return nullptr;
  switch (which) {
    case D3F_B_X: return b->x.p;
    case D3F_B_Y: return b->y.p;
    case D3F_B_VX: return b->vx.p;
    case D3F_B_VY: return b->vy.p;
    case D3F_B_CHARGE: return b->charge.p;
    case D3F_B_SOURCE: b->csr_valid = false; return b->source.p;
    case D3F_B_TARGET: b->csr_valid = false; return b->target.p;
    case D3F_B_LINK_STRENGTH: b->csr_valid = false; return b->lstrength.p;
    case D3F_B_LINK_DISTANCE: b->csr_valid = false; return b->ldistance.p;
    case D3F_B_LINK_BIAS: b->csr_valid = false; return b->lbias.p;
    case D3F_B_SEED: return b->seed_ids.p;
    default: return nullptr;
  }
}

EXPORT int32_t batch_sync(uint32_t h, uint32_t direction) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  const size_t bn = (size_t)b->n * sizeof(double), bm = (size_t)b->m * sizeof(double);
  hipError_t e = hipSuccess;
  hipStream_t s = b->stream;
  if (direction == D3F_SYNC_H2D) {
    for (uint32_t g = 0; g < b->graphs; ++g)  // link endpoints stay inside their graph
      for (uint32_t k = b->link_off[g]; k < b->link_off[g + 1]; ++k)
        if (b->source.p[k] < b->node_off[g] || b->source.p[k] >= b->node_off[g + 1] ||
            b->target.p[k] < b->node_off[g] || b->target.p[k] >= b->node_off[g + 1])
          return B_BAD_ARGUMENT;
    struct { void* d; const void* p; size_t bytes; } copies[] = {
        {b->dx.p, b->x.p, bn}, {b->dy.p, b->y.p, bn}, {b->dvx.p, b->vx.p, bn}, {b->dvy.p, b->vy.p, bn},
        {b->dcharge.p, b->charge.p, bn}, {b->dls.p, b->lstrength.p, bm}, {b->dld.p, b->ldistance.p, bm},
        {b->dlb.p, b->lbias.p, bm}, {b->dsrc.p, b->source.p, (size_t)b->m * 4}, {b->dtgt.p, b->target.p, (size_t)b->m * 4}};
    for (auto& c : copies)
      if (e == hipSuccess && c.bytes) e = hipMemcpyAsync(c.d, c.p, c.bytes, hipMemcpyHostToDevice, s);
    std::vector<uint32_t> states(b->graphs);  // the random states, 32 bits each
    for (uint32_t g = 0; g < b->graphs; ++g) states[g] = (uint32_t)b->seed_ids.p[g];
    if (e == hipSuccess && b->graphs)
      e = hipMemcpyAsync(b->dlcg.p, states.data(), (size_t)b->graphs * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    b->csr_valid = false;
  } else if (direction == D3F_SYNC_D2H) {
    struct { void* p; const void* d; } copies[] = {{b->x.p, b->dx.p}, {b->y.p, b->dy.p}, {b->vx.p, b->dvx.p},
                                                   {b->vy.p, b->dvy.p}};
    for (auto& c : copies)
      if (e == hipSuccess && bn) e = hipMemcpyAsync(c.p, c.d, bn, hipMemcpyDeviceToHost, s);
    std::vector<uint32_t> states(b->graphs);
    if (e == hipSuccess && b->graphs)
      e = hipMemcpyAsync(states.data(), b->dlcg.p, (size_t)b->graphs * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    for (uint32_t g = 0; e == hipSuccess && g < b->graphs; ++g) b->seed_ids.p[g] = states[g];
  } else {
    return B_BAD_ARGUMENT;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  return bstatus(e);
}

EXPORT int32_t batch_set_scalar(uint32_t h, uint32_t which, double v) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  if (which >= D3F_BS_COUNT || which == D3F_BS_GRAPH_REPLAYS || which == D3F_BS_GRAPH_REDOS) return B_BAD_ARGUMENT;
  b->scalars[which] = v;
  if (which == D3F_BS_ALPHA || which == D3F_BS_TICK)
    b->state_dirty = true;   // read by the device each tick
  else if (which != D3F_BS_GRAPH)
    b->drop_graph();         // the captured tick holds the others as launch arguments
  return B_OK;
}

EXPORT double batch_get_scalar(uint32_t h, uint32_t which) {
  Batch* b = bget(h);
  if (!b || which >= D3F_BS_COUNT) return 0;
  if (which == D3F_BS_GRAPH_REPLAYS) return (double)b->graph_replays;
  if (which == D3F_BS_GRAPH_REDOS) return (double)b->graph_redos;
  return b->scalars[which];
}

EXPORT int32_t batch_tick(uint32_t h, uint32_t ticks) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  if (b->graphs == 0) return B_BAD_ARGUMENT;
  for (uint32_t k = 0; k < ticks; ++k) {
    const int32_t st = b->tick();
    if (st != B_OK) return st;
  }
  return B_OK;
}

// Launch-shape autotuner: coordinate descent over the FMM's result-neutral
// launch shapes (dual-traversal grid, upward-pass grid, near-field item
// blocks), each candidate timed with HIP events over `ticks` real ticks of
// the batch. Every candidate computes the same bits, so the tuning ticks are
// ordinary ticks of the simulation (one untimed tick, then `ticks` per
// candidate). chosen[3] receives the shapes kept;
// ms_out (optional, 16 floats) the per-tick time of each candidate in order.
EXPORT int32_t batch_autotune(uint32_t h, uint32_t ticks, uint32_t* chosen, float* ms_out) {
  Batch* b = bget(h);
  if (!b) return B_BAD_HANDLE;
  if (b->graphs == 0 || ticks == 0) return B_BAD_ARGUMENT;
  b->drop_graph();  // a captured tick holds the launch shapes
  static const uint32_t dual[] = {256, 512, 1024, 2048, 4096};
  static const uint32_t level[] = {4096, 8192, 16384, 32768, 65536};
  static const uint32_t extra[] = {64, 128, 256, 512};
  struct Knob { uint32_t* field; const uint32_t* values; int count; };
  Knob knobs[3] = {{&b->fmm.dual_blocks, dual, 5}, {&b->fmm.level_blocks, level, 5}, {&b->fmm.near_extra, extra, 4}};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return B_FAILED;
  // one untimed tick first: the first tick also builds the link rows and
  // grows the buffers, which would bias the first candidate
  int32_t st = b->tick();
  int slot = 0;
  for (Knob& k : knobs) {
    uint32_t best = *k.field;
    float best_ms = -1;
    for (int c = 0; c < k.count && st == B_OK; ++c) {
      *k.field = k.values[c];
      if (hipEventRecord(e0, b->stream) != hipSuccess) st = B_FAILED;
      for (uint32_t t = 0; t < ticks && st == B_OK; ++t) st = b->tick();
      float ms = 0;
      if (st == B_OK && (hipEventRecord(e1, b->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        st = B_FAILED;
      ms /= (float)ticks;
      if (ms_out && slot < 16) ms_out[slot] = ms;
      ++slot;
      if (best_ms < 0 || ms < best_ms) {
        best_ms = ms;
        best = k.values[c];
      }
    }
    *k.field = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->drop_graph();
  if (chosen) {
    chosen[0] = b->fmm.dual_blocks;
    chosen[1] = b->fmm.level_blocks;
    chosen[2] = b->fmm.near_extra;
  }
  return st;
}

EXPORT void batch_set_timing(int on) { Batch::g_batch_timing = on != 0; }

EXPORT int32_t batch_stats(uint32_t h, d3f_stats* out) {
  Batch* b = bget(h);
  if (!b || !out) return B_BAD_HANDLE;
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return bstatus(e);
  for (auto& p : b->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, p.second.first, p.second.second) == hipSuccess) {
      b->stats.ms[p.first] += ms;
      b->stats.launches[p.first] += 1;
    }
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  b->pending.clear();
  *out = b->stats;
  b->stats = d3f_stats{};
  return B_OK;
}
