/* d3fgpu: the d3-force tick engine on AMD MI355X (gfx950), as a C ABI.
 *
 * This is the drop-in boundary for the reference's core: every entry point
 * below replaces one export of the reference's Zig/wasm core
 * (src/core/root.zig), with the same handles, enums, status codes and
 * semantics. Differences forced by the boundary are marked "GPU:".
 *
 *  - Handles are 1-based u32; 0 is never valid (root.zig:29-35).
 *  - Status codes (root.zig:16-25): 0 OK, 1 GROWN, -1 BAD_HANDLE,
 *    -2 OUT_OF_MEMORY, -3 BAD_ARGUMENT, -4 WORKER_FAILED (a GPU error).
 *  - Arrays live in pinned host memory owned by the library; sim_ptr and
 *    force_ptr return their addresses (root.zig returned byte offsets into
 *    wasm linear memory). An address stays valid until the next call that
 *    returns GROWN for that simulation or force.
 *  - GPU: node arrays are mirrored in HBM. sim_sync(h, D3F_SYNC_H2D) makes
 *    the device copy current after the host wrote x, y, vx, vy, fx, fy,
 *    fixed; sim_sync(h, D3F_SYNC_D2H) brings x, y, vx, vy back and waits.
 *    Force parameter arrays are uploaded on the next force_run after any
 *    force_ptr or force_set_count call (the facade writes them only through
 *    pointers obtained right after those calls).
 *  - Every call is synchronous with respect to the random callback: jiggle
 *    draws happen inside force_run, in d3's order.
 */
#ifndef D3FGPU_H
#define D3FGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Force kinds, src/core/force.zig:8. */
enum { D3F_CENTER = 0, D3F_COLLIDE = 1, D3F_LINK = 2, D3F_MANY_BODY = 3, D3F_RADIAL = 4, D3F_X = 5, D3F_Y = 6 };
/* Scalars, force.zig:13. */
enum {
  D3F_STRENGTH = 0, D3F_CX = 1, D3F_CY = 2, D3F_THETA2 = 3, D3F_DISTANCE_MIN2 = 4, D3F_DISTANCE_MAX2 = 5,
  D3F_ITERATIONS = 6, D3F_PRECISION = 7, D3F_ORDER = 8, D3F_OPENING = 9, D3F_SEED = 10
};
/* Force arrays, force.zig:21 (f0 strength/radius, f1 target/distance, f2 bias,
 * source, target). */
enum { D3F_F0 = 0, D3F_F1 = 1, D3F_F2 = 2, D3F_SOURCE = 3, D3F_TARGET = 4 };
/* sim_ptr slots, src/core/sim.zig:14 (x, y, vx, vy, fx, fy, fixed, perm,
 * rng_buf, snap). */
enum { D3F_PX = 0, D3F_PY = 1, D3F_PVX = 2, D3F_PVY = 3, D3F_PFX = 4, D3F_PFY = 5, D3F_PFIXED = 6,
       D3F_PPERM = 7, D3F_PRNG = 8, D3F_PSNAP = 9 };
enum { D3F_SYNC_H2D = 0, D3F_SYNC_D2H = 1 };

/* The running force's random source: replaces the wasm import env.random
 * (root.zig:139-145, core.js:11-14). Called on the calling thread, inside
 * force_run, once per jiggle draw, in d3's order. */
typedef double (*d3f_random_fn)(void* ctx);
void core_set_random(d3f_random_fn fn, void* ctx);

/* root.zig:65-78. Returns a handle, or 0 when out of memory or no GPU. */
uint32_t sim_create(void);
/* root.zig:80-86. */
int32_t sim_destroy(uint32_t h);
/* root.zig:88-92: 0, or GROWN when arrays moved (refresh pointers). */
int32_t sim_set_node_count(uint32_t h, uint32_t n);
/* root.zig:94-98: address of array `which` (D3F_P*), NULL when empty. */
void* sim_ptr(uint32_t h, uint32_t which);
/* GPU: host<->HBM synchronisation of the node arrays (no reference
 * counterpart; core.js copyIn / copyOut at core.js:121-140 call it). */
int32_t sim_sync(uint32_t h, uint32_t direction);
/* root.zig:205-209: copies vx, vy (and x, y when positions != 0) into the
 * snapshot array, sim_ptr(h, D3F_PSNAP): vx, vy, x, y, n entries each.
 * GPU: the copy is taken in HBM on the simulation's stream; sim_ptr(h,
 * D3F_PSNAP) brings it to the host array. */
int32_t sim_snapshot(uint32_t h, uint32_t positions);
/* root.zig:250-254: the end of d3's tick (velocity decay, fixed nodes). */
int32_t sim_integrate(uint32_t h, double velocity_decay);
/* GPU: `ticks` whole ticks at once (no reference counterpart; it replaces the
 * facade's per-tick loop, simulation.js:63-82, of force_run for each force in
 * `forces` order with alpha = alphas[k], then sim_integrate). The result is
 * the loop's, bit for bit, including the random source's draws (a SEED =
 * random() per seeded force per tick in fast precision, as core.js:213-231
 * draws it, and every jiggle in the eager order); the caller sets each
 * force's PRECISION, ORDER and OPENING scalars first. The tick is captured
 * once as a HIP graph and replayed; a replay that meets work only the eager
 * tick does (jiggles, buffers outgrown, a declined FMM) is rolled back to its
 * start and that tick runs eagerly. Returns D3F_NOT_CAPTURED (2), having run
 * nothing, when the per-force loop must be used instead: timing on, a
 * failure injection pending, fewer than 2 nodes, a forceCollide among the
 * forces, or more than one manyBody or link force. *done receives the ticks
 * run (all of them unless a status < 0 is returned). */
#define D3F_NOT_CAPTURED 2
int32_t sim_ticks(uint32_t h, const uint32_t* forces, uint32_t nforces, uint32_t ticks, const double* alphas,
                  double velocity_decay, uint32_t* done);
/* GPU: out[0] ticks captured, out[1] ticks replayed, out[2] replays rolled
 * back and run eagerly, since sim_create (no reference counterpart). */
int32_t sim_tick_counters(uint32_t h, uint64_t* out);

/* root.zig:102-108: builds the exact quadtree on x, y (shifted = 0) or on
 * x + vx, y + vy (1). Returns GROWN when tree arrays moved. */
int32_t tree_build(uint32_t h, uint32_t shifted);
/* root.zig:113-126: 0 cell count, 1 root, 2 cells address (88-byte cells,
 * tree.zig:46-62), 3 extent address (4 f64), 4 perm address, 5 sizeof(Cell),
 * 6 points in tree. Addresses are host copies made by the call. */
uint64_t tree_info(uint32_t h, uint32_t which);
/* root.zig:130-137: cell indices in d3-quadtree visit (after = 0) or
 * visitAfter (after = 1) order into out[cell count]. */
int32_t tree_order(uint32_t h, uint32_t after, uint32_t* out);

/* root.zig:147-161: a force of `kind` (D3F_CENTER ..), or 0. */
uint32_t force_create(uint32_t h, uint32_t kind);
/* root.zig:163-170. */
int32_t force_destroy(uint32_t h, uint32_t fh);
/* root.zig:174-179: resizes the force's arrays (new float slots NaN);
 * always GROWN. */
int32_t force_set_count(uint32_t h, uint32_t fh, uint32_t count);
/* root.zig:181-186: address of array `which` (D3F_F0 ..), NULL when empty. */
void* force_ptr(uint32_t h, uint32_t fh, uint32_t which);
/* root.zig:188-194 / 196-200. */
int32_t force_set_scalar(uint32_t h, uint32_t fh, uint32_t which, double v);
double force_get_scalar(uint32_t h, uint32_t fh, uint32_t which);
/* root.zig:211-242: runs the force on the device arrays. Fast mode applies
 * when D3F_PRECISION is 1 and the node count is at least 2000
 * (root.zig:246-248). */
int32_t force_run(uint32_t h, uint32_t fh, double alpha);

/* Library introspection (no reference counterpart), for the benchmark's
 * roofline: with timing on, every phase launch is bracketed by HIP events on
 * the simulation's stream (no extra synchronisation); sim_stats waits for the
 * stream, adds the elapsed times up per phase and resets them. While timing
 * is on, the traversal also counts its algorithmic work: non-empty cell
 * visits, accepted far-field cells and leaf pair interactions, summed over
 * nodes. */
enum { D3F_PHASE_TREE = 0, D3F_PHASE_ACCUMULATE = 1, D3F_PHASE_TRAVERSAL = 2, D3F_PHASE_LINK = 3,
       D3F_PHASE_CENTER = 4, D3F_PHASE_INTEGRATE = 5, D3F_PHASE_CENTER_SUM = 6, D3F_PHASE_FMM = 7,
       D3F_PHASE_COLLIDE = 8, D3F_PHASES = 9 };
/* D3F_PHASE_CENTER is the center force on the tick's stream (including any
 * wait for its sums); D3F_PHASE_CENTER_SUM is the serial sum kernel itself,
 * which runs on a second stream overlapping the tick. */
typedef struct {
  double ms[D3F_PHASES];        /* summed kernel time per phase */
  uint64_t launches[D3F_PHASES];
  uint64_t visits, accepts, pairs;  /* Barnes-Hut work since the last call (counted while timing is on) */
  uint64_t flagged;             /* many-body nodes replayed for jiggle */
  uint64_t link_levels;         /* levels of the last exact link schedule */
  uint64_t m2l, p2p_pairs;      /* fast many-body: M2L cell pairs, P2P point pairs */
  uint64_t fmm_cells;           /* cells of the last fast tree */
  uint64_t collide_serial;      /* exact collide iterations run on the serial pass */
  uint64_t collide_status;      /* OR of the parallel collide passes' stop reasons:
                                   1 candidate list full, 2 a position left its
                                   box, 4 a zero delta (jiggle), 8 no progress,
                                   32 walk stack full */
} d3f_stats;
int32_t sim_stats(uint32_t h, d3f_stats* out);
/* FP64 direct sum for the fast mode's error measurement (no reference
 * counterpart; the reference's fmm.zig:946-971 test and
 * test/fast/error-test.js:44-58 compute the same sum on the CPU): for each
 * sample[q] < n, out[2q] and out[2q+1] receive sum over j != sample[q] of
 * d3's pair term x * strength[j] * alpha / l (coincident pairs and NaN terms
 * skipped, l < distanceMin^2 -> sqrt(distanceMin^2 * l), l >= distanceMax^2
 * skipped) on the current positions, using many-body force fh's strengths
 * and distance scalars. sample and out are host pointers. */
int32_t force_direct_sum(uint32_t h, uint32_t fh, double alpha, const uint32_t* sample, uint32_t k, double* out);
/* force_direct_sum without any device allocation (no reference counterpart;
 * for callers that cannot allocate: code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). Same arguments and results, bit for
 * bit; the sample indices and outputs are staged in a fixed device array of
 * the library (STATIC_SAMPLES = 16384 samples at most; more return
 * BAD_ARGUMENT, with core_last_error saying so), held by each device until the
 * process ends. The force's parameter arrays must already be on the device
 * (any force_run of the force, or force_direct_sum, puts them there);
 * otherwise BAD_ARGUMENT. Every call on one device shares that array: calls
 * that overlap on one GPU (several host threads) race. */
int32_t force_direct_sum_static(uint32_t h, uint32_t fh, double alpha, const uint32_t* sample, uint32_t k,
                                double* out);
void core_set_timing(int on);
/* Waits for all work the library queued on the device (benchmark brackets). */
int32_t core_synchronize(void);
/* The cause of the calling thread's last OUT_OF_MEMORY or WORKER_FAILED
 * status (no reference counterpart; the wasm core has no such errors): the
 * HIP error name, the library source line that failed and what failed there.
 * Empty before the first failure. The string is owned by the library. */
const char* core_last_error(void);
/* Failure injection for tests of the failover path (no reference
 * counterpart; it stands in for a worker that dies, pool.js:73-84): after
 * `after` more steps (force_run or sim_integrate, on any simulation) run
 * normally, the next `count` steps return `status` (WORKER_FAILED -4 or
 * OUT_OF_MEMORY -2) before any device work, and core_last_error names the
 * injection. Returns BAD_ARGUMENT for any other status. */
int32_t core_fail_next(int32_t status, uint32_t after, uint32_t count);

/* ---- Batched fast mode (no reference counterpart) ----
 * Many independent simulations ("graphs") in one node arena and one link
 * arena, ticked together on one GPU: per graph, a tick is the facade's tick
 * in fast precision with forceManyBody, forceLink and forceCenter in that
 * order, then integrate (simulation.js:49-89, root.zig:211-254 for one
 * simulation). Graph g owns nodes [node_off[g], node_off[g+1]) and links
 * [link_off[g], link_off[g+1]); link endpoints are arena node indices inside
 * their graph (checked at batch_sync). As in the reference (root.zig:246-248),
 * a graph of at least 2000 nodes runs fast mode (FMM, adjacency-sweep link)
 * and a smaller one the exact kernels (Barnes-Hut with D3F_BS_THETA2, the
 * Gauss-Seidel link), bit-identical to the reference's exact tick. Each graph
 * has its own random source, d3's lcg (lcg.js) with state D3F_B_SEED
 * (1 at the start, as every d3 simulation): per tick the manyBody and link
 * seeds are drawn from it, and a small graph's jiggle draws, in the facade's
 * order (core.js:213-231). A graph whose extent is not finite makes
 * batch_tick return BAD_ARGUMENT; a graph below 2000 nodes whose quadtree is
 * deeper than about 60 levels makes it return WORKER_FAILED. Status codes are
 * the simulation's. */
enum { D3F_B_X = 0, D3F_B_Y = 1, D3F_B_VX = 2, D3F_B_VY = 3, D3F_B_CHARGE = 4, D3F_B_SOURCE = 5, D3F_B_TARGET = 6,
       D3F_B_LINK_STRENGTH = 7, D3F_B_LINK_DISTANCE = 8, D3F_B_LINK_BIAS = 9, D3F_B_SEED = 10 };
/* Batch scalars; VELOCITY_DECAY is the multiplier (1 - d3's velocityDecay,
 * default 0.6), DISTANCE_MIN2 / MAX2 the squared distances, TICK counts the
 * ticks run. GRAPH = 1 replays one captured HIP graph per tick (captured
 * after an eager tick, with buffers sized for twice the cells that tick had;
 * a replay whose sizes do not fit is undone and the tick redone eagerly, and
 * the graph recaptured); results are bit-identical to the eager ticks.
 * GRAPH_CELLS, when nonzero, caps the cells a replay may hold (0: the
 * buffers' capacity), so that replays take the redo path (for tests).
 * GRAPH_REPLAYS and GRAPH_REDOS are read-only counts of the replays launched
 * and of those redone eagerly. */
enum { D3F_BS_ALPHA = 0, D3F_BS_ALPHA_DECAY = 1, D3F_BS_ALPHA_TARGET = 2, D3F_BS_VELOCITY_DECAY = 3,
       D3F_BS_DISTANCE_MIN2 = 4, D3F_BS_DISTANCE_MAX2 = 5, D3F_BS_ORDER = 6, D3F_BS_OPENING = 7, D3F_BS_CENTER_X = 8,
       D3F_BS_CENTER_Y = 9, D3F_BS_CENTER_STRENGTH = 10, D3F_BS_LINK_ITERATIONS = 11, D3F_BS_TICK = 12,
       D3F_BS_GRAPH = 13, D3F_BS_GRAPH_CELLS = 14, D3F_BS_GRAPH_REPLAYS = 15, D3F_BS_GRAPH_REDOS = 16,
       D3F_BS_THETA2 = 17, D3F_BS_COUNT = 18 };
uint32_t batch_create(void);
int32_t batch_destroy(uint32_t b);
/* node_off, link_off: graphs + 1 host entries each, starting at 0. */
int32_t batch_set_graphs(uint32_t b, uint32_t graphs, const uint32_t* node_off, const uint32_t* link_off);
/* Pinned host array D3F_B_*: f64 per node (X..CHARGE), u32 per link
 * (SOURCE, TARGET), f64 per link (LINK_*), u64 per graph (SEED: the random
 * state, low 32 bits used; batch_sync D2H brings the current states back,
 * H2D sets them). */
void* batch_ptr(uint32_t b, uint32_t which);
int32_t batch_sync(uint32_t b, uint32_t direction);
int32_t batch_set_scalar(uint32_t b, uint32_t which, double v);
double batch_get_scalar(uint32_t b, uint32_t which);
int32_t batch_tick(uint32_t b, uint32_t ticks);
/* Autotunes the FMM's launch shapes (dual-traversal grid, upward-pass
 * grid, blocks sharing near-field items; no result depends on them) by timing
 * `ticks` real ticks per candidate (after one untimed tick), one shape at a
 * time; keeps the fastest.
 * chosen[3] receives the shapes kept; ms (optional, 16 floats) the per-tick
 * time of each candidate. The tuning ticks advance the batch. No reference
 * counterpart. */
int32_t batch_autotune(uint32_t b, uint32_t ticks, uint32_t* chosen, float* ms);
void batch_set_timing(int on);
int32_t batch_stats(uint32_t b, d3f_stats* out);

/* ---- Learned warm start (no reference counterpart) ----
 * A message-passing network predicting a layout from node features: 8
 * features -> 64, 4 layers h' = relu(h Ws + mean_nbr(h) Wn + b), -> 2
 * positions; BF16 weights and activations with FP32 accumulation on the
 * matrix cores, forward and backward. gnn_param_count() FP32 parameters in
 * the layout W0[16][64] b0[64] {Ws[64][64] Wn[64][64] b[64]} x 4 Wo[64][16]
 * bo[16] (padding rows / columns zero). Status codes are the simulation's. */
uint32_t gnn_create(uint64_t seed);
int32_t gnn_destroy(uint32_t h);
uint32_t gnn_param_count(void);
float* gnn_params(uint32_t h); /* pinned host copy of the parameters */
int32_t gnn_sync_params(uint32_t h, uint32_t direction); /* H2D also resets the optimizer */
/* row[n + 1], col[row[n]]: symmetric adjacency rows; features: n x 8 FP32. */
int32_t gnn_set_graph(uint32_t h, uint32_t n, const uint32_t* row, const uint32_t* col, const float* features);
int32_t gnn_forward(uint32_t h, float* positions); /* n x 2 */
/* The graph from its m links alone, on the GPU: symmetric adjacency rows
 * without self-links or repeats, and the 8 input features (log(1 + degree),
 * 1 / sqrt(1 + degree), six smooth random-walk modes from seeded subspace
 * iteration); features, when not null, receives them (n x 8 FP32). */
int32_t gnn_set_links(uint32_t h, uint32_t n, uint32_t m, const uint32_t* source, const uint32_t* target,
                      uint64_t seed, float* features);
/* Predicted layout of the graph set by gnn_set_links (n x 2 FP64): the
 * network's output centred, scaled to the size minimising d3's energy over
 * uniform scalings, plus a seeded jitter of +-jitter times the mean link
 * length. */
int32_t gnn_predict(uint32_t h, double jitter, uint64_t seed, double* positions);
/* Stress pairs for training: |p_i - p_j| should be d. */
int32_t gnn_set_pairs(uint32_t h, uint32_t pairs, const uint32_t* i, const uint32_t* j, const float* d);
/* Gradient of sum(dp * positions) into grad (gnn_param_count() floats). */
int32_t gnn_gradient(uint32_t h, const float* dp, float* grad);
/* One Adam step on the stress; *loss receives the stress before the step. */
int32_t gnn_train_step(uint32_t h, float lr, float* loss);

#ifdef __cplusplus
}
#endif

#endif
