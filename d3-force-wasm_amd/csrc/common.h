// Shared definitions of the d3fgpu HIP kernels and host runtime (gfx950).
//
// Floating point: every kernel is compiled with -ffp-contract=off and without
// fast-math, so each FP64 +, -, *, / and sqrt rounds exactly once, as in the
// reference's wasm core; this file also pins contraction off per translation
// unit.
#pragma once
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace d3f {

constexpr uint32_t NONE = 0xffffffffu;
constexpr uint32_t MAX_DEPTH = 2200;  // tree.zig:33
constexpr int KEY_LEVELS = 32;        // tree.zig:30

// The reference's 88-byte quadtree cell (tree.zig:46-62): children 0..3,
// the cell's points perm[first .. first+count), d3's visit bounds, and the
// per-force summary (x, y, value, r).
struct Cell {
  uint32_t child[4];
  uint32_t first, count;
  double x0, y0, x1, y1;
  double x, y, value, r;
};
static_assert(sizeof(Cell) == 88, "reference cell layout");

// The traversal's view of a cell: centre of charge, charge, the opening
// threshold w*w/theta2, the number of cells in its subtree (cells are numbered
// in d3's visit preorder, so a cell's first child is the next cell and its
// subtree ends `skip` cells on), its leaf chain, and flags (bit 0 leaf, bit 1
// "a truthy leaf is reachable"). A visit reads the first 48 bytes.
struct TCell {
  double qx, qy, value, w2t;
  uint32_t skip, first, count, head;
  uint32_t flags, pad[3];
};
static_assert(sizeof(TCell) == 64, "traversal cell layout");
constexpr uint32_t TC_LEAF = 1, TC_REACH = 2;

// Device-resident scalars of the last tree build.
struct TreeMeta {
  double extent[4];
  uint32_t n;       // nodes
  uint32_t m;       // points in the tree (no NaN coordinate)
  uint32_t len;     // cells
  uint32_t depth;   // deepest cell
  uint32_t mixed;   // key groups holding distinct positions (sub-2^-32 splits)
  uint32_t valid;   // 1 when the tree has a root
  uint32_t overflow;  // a deep walk exceeded its stack
  uint32_t ticket;  // block counter of the box reduction
  uint32_t long_group;  // a run of equal high key halves longer than the in-place sort takes
  uint32_t ncut;      // subtrees of at most ACC_SUB cells whose parent's subtree is larger (or the root's)
  uint32_t ntop;      // cells whose subtree exceeds ACC_SUB cells
  uint32_t topdepth;  // deepest of those cells
  uint32_t stop;      // capturable build: a tree its buffers or kernels cannot finish (see tree_build_exact)
};

// Subtrees of at most this many cells are accumulated by one workgroup in LDS.
// Their list entries are (root, size | depth << CUT_SIZE_BITS).
constexpr uint32_t ACC_SUB = 1024;
constexpr uint32_t CUT_SIZE_BITS = 11, CUT_SIZE_MASK = (1u << CUT_SIZE_BITS) - 1;

// A depth-first frame of the below-key-resolution build (tree.zig:64-76).
struct Frame {
  uint32_t lo, hi, depth, parent, slot, pad;
  double x0, y0, x1, y1;
};
constexpr uint32_t DEEP_STACK = 3 * MAX_DEPTH + 64;

// Notes the first failure since the last status the library returned: the
// HIP error, what failed and its source line (reported by core_last_error).
void note_failure(hipError_t e, const char* what, const char* file, int line);
// Publishes the noted failure, or names e alone when none was noted; clears
// the note on success.
void publish_failure(hipError_t e);

#define HIPCHECK(expr)                                         \
  do {                                                         \
    hipError_t e_ = (expr);                                    \
    if (e_ != hipSuccess) {                                    \
      ::d3f::note_failure(e_, #expr, __FILE__, __LINE__);      \
      return e_;                                               \
    }                                                          \
  } while (0)

// Returns err after noting `why` as its cause.
#define HIPFAIL(err, why)                                      \
  do {                                                         \
    ::d3f::note_failure((err), (why), __FILE__, __LINE__);     \
    return (err);                                              \
  } while (0)

__host__ __device__ inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// Fast mode's counter-based jiggle (hash.zig:10-25): the splitmix64
// finaliser of (seed, i, j, salt) -> 53 uniform bits -> (u - 0.5) * 1e-6.
__host__ __device__ inline double hash_jiggle(uint64_t seed, uint32_t i, uint32_t j, uint32_t salt) {
  uint64_t h = seed ^ ((uint64_t)i << 32 | j) ^ ((uint64_t)salt * 0x9E3779B97F4A7C15ull);
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h = h ^ (h >> 31);
  const double u = (double)(h >> 11) * 0x1.0p-53;
  return (u - 0.5) * 1e-6;
}

// ---- tree.hip ----
struct TreeBuffers {
  uint32_t cap = 0;        // nodes the point buffers hold
  uint32_t cell_cap = 0;   // cells the cell buffers hold
  uint64_t *keys = nullptr, *keys_tmp = nullptr, *keys_all = nullptr;
  uint32_t *perm = nullptr, *perm_tmp = nullptr, *rest = nullptr;
  uint32_t *flags = nullptr, *offs = nullptr;        // compaction
  uint32_t *count = nullptr, *base = nullptr;        // cells per key group, prefix
  int32_t *dleft = nullptr;                          // divergence depth with the previous group
  uint32_t *gend = nullptr, *leafdepth = nullptr, *mixed = nullptr, *mixed_list = nullptr;
  uint32_t *hist = nullptr;                          // radix histograms
  uint32_t *partials = nullptr;                      // scan partials
  void* box_partials = nullptr;
  Frame* deep_stack = nullptr;
  Cell* cells = nullptr;
  TCell* tcells = nullptr;
  uint32_t* cdepth = nullptr;                        // depth of each cell
  uint32_t* cskip = nullptr;                         // cells in each cell's subtree (itself included)
  uint2* cut = nullptr;                              // subtree roots (TreeMeta ncut), see ACC_SUB
  uint32_t* top = nullptr;                           // cells above them (TreeMeta ntop)
  uint32_t *cell_s = nullptr, *cell_d = nullptr;     // each cell's key group and depth
  uint32_t* rs_glob = nullptr;                       // radix digit totals, 8 passes
  uint32_t *dbins = nullptr, *dstart = nullptr, *dorder = nullptr;  // cells by depth
  TreeMeta* meta = nullptr;       // device
  TreeMeta* meta_host = nullptr;  // pinned mirror
};

// The streams every simulation on the current device runs on: its main
// stream, a side stream for work that overlaps the tick (forceCenter's sums)
// and an auxiliary one for work that overlaps a force's own launches (the
// FMM's multipoles). Created once, back to back, so each gets a hardware
// queue of its own; streams created and destroyed with each simulation could
// come to share one queue, and their work would then run one after the other.
hipError_t shared_streams(hipStream_t* main, hipStream_t* side, hipStream_t* aux);

hipError_t tree_reserve(TreeBuffers& t, uint32_t n);
hipError_t tree_reserve_cells(TreeBuffers& t, uint32_t cells);
void tree_free(TreeBuffers& t);
// Builds the exact quadtree over (x, y)[0..n) on `stream`; on return
// meta_host holds the tree's scalars (the call synchronises once, after the
// cell count is known, to size the cell buffers).
// graph: the capturable form. Nothing is read back; launches are sized by n
// and cell_cap, and meta->stop is set, with no cell written, for a tree this
// form does not build: no finite point, a run of equal high key halves too
// long for the in-place sort, points distinct below the key resolution, or
// more cells than cell_cap. The kernels that use the tree then do nothing.
hipError_t tree_build_exact(TreeBuffers& t, const double* x, const double* y, uint32_t n, hipStream_t stream,
                            bool graph = false);
// Stable LSD radix sort of (keys, vals)[0 .. *count) on the 8-bit digits
// [first_digit, first_digit + digits); n_max bounds *count. The result is in
// keys / vals (after an odd number of passes it is copied back from the
// ping-pong buffers).
hipError_t sort_digits(uint32_t* hist, uint32_t* glob, uint64_t* keys, uint32_t* vals, uint64_t* keys_tmp,
                       uint32_t* vals_tmp, uint32_t n_max, const uint32_t* count, int first_digit, int digits,
                       hipStream_t stream);
size_t sort_hist_words(uint32_t n_max);  // hist size sort_digits needs
// Exclusive scan of in[0 .. *count_dev) (n_max bounds the count; partials
// holds ceil(n_max / 4096) + 1 words); the total goes to *total_out.
hipError_t scan_u32(const uint32_t* in, uint32_t* out, uint32_t n_max, const uint32_t* count_dev, uint32_t* partials,
                    uint32_t* total_out, hipStream_t stream);

// ---- collide.hip ----
constexpr uint32_t COLLIDE_K = 64;        // candidates per node
constexpr uint32_t COLLIDE_STACK = 160;   // walk stack of the candidate pass
// Status bits of a parallel pass; any of them sends the iteration to the
// serial kernel.
constexpr uint32_t COLLIDE_FULL = 1;      // more candidates than COLLIDE_K
constexpr uint32_t COLLIDE_MOVED = 2;     // a position left its box
constexpr uint32_t COLLIDE_JIGGLE = 4;    // an overlapping pair with a zero delta
constexpr uint32_t COLLIDE_STALLED = 8;   // a wait made no progress for a second
constexpr uint32_t COLLIDE_FULL_STACK = 32;  // a walk deeper than COLLIDE_STACK allows
constexpr uint32_t COLLIDE_NEED_MOVED = 0x80;  // serial continuation: a push left its box
struct alignas(64) CRec {
  double x0, y0, x1, y1, r;
  uint32_t child[4];
  uint32_t head;  // a leaf's chain head, NONE for an internal cell
  uint32_t pad;
};
static_assert(sizeof(CRec) == 64, "collide walk record");
struct CollideArgs {
  Cell* cells;
  CRec* rec;                  // 64-byte walk records of the cells
  const TreeMeta* meta;
  const uint32_t* perm;
  const double *px, *py;      // tree positions x + vx, y + vy at the iteration's start
  const double *x, *y;
  double *vx, *vy;
  const double* radii;
  double strength;
  double scale;               // box half-width in units of |strength| * largest radius
  uint32_t n;
  uint32_t m;                 // points in the tree (perm[0 .. m))
  uint32_t key_shift;         // pair keys are j << key_shift | i
  uint32_t *ccount, *coff;    // per node: candidates, offset of its pairs
  uint32_t *cand, *cleaf, *ctop;  // per candidate: node, leaf cell, first undecided ancestor (or NONE)
  uint64_t* pkeys;
  uint32_t* pvals;
  uint32_t *run_start, *rank, *indeg, *done;
  uint32_t* fin;              // per node: 1 once its pass is done (or it has no candidates)
  uint32_t *status, *ticket;
  uint32_t* stop;             // lowest node parked for a jiggle (NONE: none)
};
struct CollideSerialState {
  uint32_t node, sp, inside, need, has_jig, pad;
  double xi, yi;
  double jig[2];
};
// Cell radii for collide (collide.zig:104-130) into t.cells[].r.
hipError_t collide_radii(TreeBuffers& t, const double* radii, hipStream_t stream);
// The walk records of cells[0 .. len) (after collide_radii).
hipError_t collide_records(const CollideArgs& a, uint32_t len, hipStream_t stream);
hipError_t collide_candidates(const CollideArgs& a, hipStream_t stream);
hipError_t collide_pairs(const CollideArgs& a, hipStream_t stream);
hipError_t collide_ranks(const CollideArgs& a, uint32_t pairs, hipStream_t stream);
hipError_t collide_apply(const CollideArgs& a, hipStream_t stream);
// The reference's pass on one thread, over every node (list == nullptr,
// count == n) or over the listed ones, checking the boxes.
hipError_t collide_serial(const CollideArgs& a, uint32_t* stack, CollideSerialState* st, const uint32_t* list,
                          uint32_t count, hipStream_t stream);
// The nodes with fin == 0 in index order into list; their count to *count_out.
hipError_t collide_unfinished(const CollideArgs& a, uint32_t* flag, uint32_t* pos, uint32_t* list,
                              const uint32_t* n_dev, uint32_t* partials, uint32_t* count_out, hipStream_t stream);

// ---- many_body.hip ----
// Cells grouped by depth: t.dorder[t.dstart[d] .. t.dstart[d + 1]) holds the
// cells of depth d (after a tree build, whose meta_host gives len and depth).
hipError_t tree_depth_order(TreeBuffers& t, hipStream_t stream);
// visitAfter charge sums (many_body.zig:101-135) into tcells (qx, qy, value).
// graph: after a capturable tree build (sizes from cell_cap, nothing read back).
hipError_t mb_accumulate(TreeBuffers& t, const double* x, const double* y, const double* strengths, double theta2,
                         hipStream_t stream, bool graph = false);
// Copies the accumulated charges of every cell from tcells into the 88-byte
// cells' x, y, value (what the tree export shows after a many-body run).
hipError_t mb_publish_charges(TreeBuffers& t, hipStream_t stream);
enum TravMode { TRAV_RECORD = 0, TRAV_COUNT = 1, TRAV_REPLAY = 2 };
struct TravArgs {
  const TCell* tc;
  const uint32_t* order;   // node per lane slot
  uint32_t count;          // lane slots
  const uint32_t* perm;
  const double *x, *y, *s;
  double *vx, *vy;
  double dmin2, dmax2, alpha;
  uint32_t* flagged;       // RECORD: appended jiggle nodes
  uint32_t* n_flagged;
  uint32_t* draws_count;   // COUNT: draws per slot
  const uint32_t* draw_off;  // REPLAY: first draw per slot
  const double* draws;     // REPLAY: the random() values
  uint32_t len;            // cells of the tree
  unsigned long long* counters;  // visits, accepts, pairs (summed)
  // Captured ticks: count and len are read from the tree's meta (m, len; no
  // walk when stop is set) and alpha from alpha_dev. count then bounds the grid.
  const TreeMeta* meta = nullptr;
  const double* alpha_dev = nullptr;
};
hipError_t mb_traverse(const TravArgs& a, int mode, hipStream_t stream);
// Nodes with a NaN coordinate (tree.rest): their traversal reaches every
// truthy leaf through truthy cells, and any reached leaf makes v NaN.
hipError_t mb_nan_nodes(const TreeBuffers& t, uint32_t n, double* vx, double* vy, hipStream_t stream);
// The same after a capturable build (the count n - m read on the device).
hipError_t mb_nan_nodes_graph(const TreeBuffers& t, uint32_t n, double* vx, double* vy, hipStream_t stream);

// ---- batch_exact.hip ----
// One small graph of a batch (fewer nodes than the fast-mode crossover),
// ticked with the exact kernels by one workgroup: its node and link ranges
// in the arenas, its cell range in the small-graph cell arrays (cell_cap
// cells) and its graph index (random state). The link schedule: the graph's
// m links in level order at rec0 of the record array, level L's links at
// [levels[lev0 + L], levels[lev0 + L + 1]) of that range, nlev levels (the
// levels array holds m + 1 entries per graph from lev0); links of one level
// share no node, and every link's level is above those of the earlier links
// at its ends, so a level at a time gives the bits of the serial order.
struct SmallJob {
  uint32_t node0, n, link0, m, cell0, cell_cap, graph, rec0, lev0, nlev, pad[2];
};
static_assert(sizeof(SmallJob) == 48, "small-graph job");
// One link of a small graph in the level schedule: its ends (local node
// indices) and parameters.
struct SmallLink {
  uint16_t a, b;
  uint32_t pad;
  double distance, strength, bias;
};
static_assert(sizeof(SmallLink) == 32, "small-graph link record");
struct SmallExactArgs {
  const double *x, *y, *vx, *vy, *charge;
  const uint32_t *src, *tgt;
  const double *lstrength, *ldistance, *lbias;
  const SmallLink* lrec;     // the link schedule (small_schedule)
  const uint32_t* levels;
  const double* alpha_dev;   // this tick's alpha
  double theta2, dmin2, dmax2, link_iterations;
  uint32_t* lcg;             // per graph: d3's lcg state (lcg.js)
  void* cells;               // small_cell_bytes() x cells
  TCell* tcells;
  double *vout_x, *vout_y;   // the small graphs' velocities after manyBody and link
  uint32_t* error;           // set nonzero when a graph's tree outgrows the kernel
};
size_t small_cell_bytes();
// manyBody + link, exact, for every job; reads x, y, vx, vy, writes vout.
hipError_t small_exact(const SmallJob* jobs, uint32_t count, const SmallExactArgs& a, hipStream_t stream);
// The link schedule of every job from the link arrays (src, tgt global node
// indices): the levels, then the records in level order; writes each job's
// nlev. order: scratch, one u32 per scheduled link.
hipError_t small_schedule(SmallJob* jobs, uint32_t count, const uint32_t* src, const uint32_t* tgt,
                          const double* lstrength, const double* ldistance, const double* lbias, uint32_t* order,
                          uint32_t* levels, SmallLink* lrec, hipStream_t stream);
// vx, vy of every job's nodes from vout.
hipError_t small_scatter(const SmallJob* jobs, uint32_t count, const double* vout_x, const double* vout_y, double* vx,
                         double* vy, hipStream_t stream);

// ---- link.hip ----
// A node's position and velocity side by side, so an edge reads each endpoint
// from one 32-byte record (the link pass's working copy).
struct alignas(32) LinkNode {
  double x, y, vx, vy;
};
struct LinkLevelArgs {
  const uint32_t *src, *tgt;
  const double *strength, *distance, *bias;
  const double *x, *y;
  double *vx, *vy;
  LinkNode* nodes;  // n records
  uint32_t n;
  double alpha;
  uint32_t* zero_flag;
  const double* alpha_dev = nullptr;  // replaces alpha when set (captured ticks)
};
// Copies x, y, vx, vy into the records.
hipError_t link_pack(const LinkLevelArgs& a, hipStream_t stream);
hipError_t link_level(const LinkLevelArgs& a, uint32_t lo, uint32_t hi, hipStream_t stream);
// Levels 1 .. levels, level l being edges off[l] .. off[l + 1] (host offsets):
// large levels one launch each, runs of small ones in one workgroup.
hipError_t link_levels(const LinkLevelArgs& a, const uint32_t* off, uint32_t levels, hipStream_t stream);
// Copies the records' velocities back unless a level raised the zero flag
// (then vx, vy keep their values from before the pass).
hipError_t link_unpack(const LinkLevelArgs& a, hipStream_t stream);
struct LinkSerialState {
  uint32_t edge;    // next edge; m when done
  uint32_t need;    // bit 0: x needs a draw, bit 1: y needs a draw
  double jig[2];    // drawn jiggles for `edge` (x then y)
  uint32_t has_jig;
};
hipError_t link_serial(const uint32_t* src, const uint32_t* tgt, const double* strength, const double* distance,
                       const double* bias, uint32_t m, const double* x, const double* y, double* vx, double* vy,
                       double alpha, LinkSerialState* state, hipStream_t stream);

// ---- shared scratch of the fast-mode kernels ----
struct SortBuffers {
  uint64_t *k0 = nullptr, *k1 = nullptr;  // keys and their ping-pong partner
  uint32_t *v0 = nullptr, *v1 = nullptr;  // values likewise
  uint32_t *hist = nullptr, *glob = nullptr, *count = nullptr;
  uint32_t cap = 0;
};

// ---- fast.hip ----
struct GridParams {
  double x0, y0, inv_h;
  uint32_t gx, gy, cells, m, sort_n, active;
};
struct GridBuffers {
  double *px = nullptr, *py = nullptr;  // x + vx, y + vy
  uint32_t* start = nullptr;
  double4* cand = nullptr;              // binned nodes: x, y, radius, node bits
  void* partials = nullptr;
  GridParams* params = nullptr;
};
size_t grid_partials_bytes();
// One adjacency entry in row order: the other node, the edge (with bit 31
// set on the target side), and the edge's strength, distance and bias
// factor (b on the target side, 1 - b on the source side).
struct LinkEntry {
  uint32_t j, e;
  double strength, distance, bf;
};
static_assert(sizeof(LinkEntry) == 32, "link entry layout");
struct LinkCsrArgs {
  const uint32_t *src, *tgt;
  const double *strength, *distance, *bias;
  uint32_t m, n;
  uint32_t *row, *adj, *owner, *rows_long;  // row[n + 1], adj[2m], owner[2m], rows_long[n]
  LinkEntry* entries;                       // [2m]
  double* terms;                            // [4m] scratch of a sweep
  // A batch of graphs (null for one simulation): each node's graph, each
  // graph's first link and jiggle seed.
  const uint32_t* node_graph = nullptr;
  const uint32_t* link_base = nullptr;
  const uint64_t* seeds = nullptr;
};
// Adjacency rows of the links (link_fast.zig:105-130) and their packed
// entries; *long_count (host) receives the number of rows longer than the
// per-thread limit. Synchronises.
hipError_t link_csr(const LinkCsrArgs& a, const SortBuffers& sb, uint32_t* long_count, hipStream_t stream);
// One Jacobi sweep over the rows on pos = (x + vx, y + vy) (link_fast.zig:60-99).
// alpha_dev, when given, replaces alpha (read on the device, for captured graphs).
// seed_dev, when given, replaces seed the same way.
hipError_t link_sweep(const LinkCsrArgs& a, uint32_t long_count, const double2* pos, double* vx, double* vy,
                      double alpha, uint64_t seed, hipStream_t stream, const double* alpha_dev = nullptr,
                      const uint64_t* seed_dev = nullptr);
// One iteration of the uniform-grid collide (grid.zig:221-285).
hipError_t grid_collide_iteration(const double* x, const double* y, double* vx, double* vy, const double* radii,
                                  uint32_t n, double strength, uint64_t seed, const GridBuffers& g,
                                  const SortBuffers& sb, hipStream_t stream);

// ---- fmm.hip ----
constexpr uint32_t FMM_LEAF = 32, FMM_LEVELS = 16;  // fmm.zig:43-47
// Cells of at most this many points get their multipole straight from their
// points (one wave each); the larger ones (at most m / FMM_DIRECT per depth)
// by M2M from their children.
constexpr uint32_t FMM_DIRECT = 2048;
constexpr uint32_t FMM_MAX_ORDER = 24;
constexpr uint32_t FMM_MAX_STEPS = 2 * FMM_LEVELS + 2;  // dual-tree rounds

// A cell of the fast tree (fmm.zig:68-78): level order, children
// contiguous; centre and radius in the unit frame u = (x - ox) / w.
struct FCell {
  double cx, cy, r;
  uint32_t first, count, child0, nchild, depth, graph;
};
static_assert(sizeof(FCell) == 48, "fast cell layout");

// One graph of a fast many-body run (the runtime's simulation is a forest of
// one): its box and the reference's fast extent and unit frame
// (tree.zig:357-359, fmm.zig:253-270), its scaled distance limits, its
// jiggle seed, its root cell and its points' range in Morton order.
struct FmmGraph {
  double box[4], extent[4];
  uint32_t node_base, node_end;      // the graph's nodes [node_base, node_end)
  double scale;                      // 65536 / side (tree.zig:360)
  double ox, oy, w, inv_w, dmin_s, dmax_s;
  uint64_t seed;
  uint32_t m, ok, ring, root, point_off, pad;
};

// Device-resident scalars of one fast many-body run.
constexpr uint32_t FRONT_STRIDE = 32;  // words per frontier count (128 bytes)
struct FmmParams {
  uint32_t n, m_total, roots, declined;    // nodes, points in trees, graphs with a tree, non-finite extents
  uint32_t cell_overflow;
  uint32_t level_n[FMM_LEVELS + 2];        // cells per depth
  uint32_t level_off[FMM_LEVELS + 3];      // first cell of each depth
  // dual-tree frontier sizes, one a cache line (round s at s * FRONT_STRIDE):
  // a round's blocks read its size while they add to the next round's
  alignas(128) uint32_t front[(FMM_MAX_STEPS + 2) * FRONT_STRIDE];
  // list lengths, each on a cache line of its own: every tile of a dual-tree
  // round reserves its output with one atomic on each of front[step + 1],
  // m2l and p2p
  alignas(128) uint32_t m2l;
  alignas(128) uint32_t p2p;
  alignas(128) uint32_t list_overflow;
  uint32_t long_segs;                      // target segments ordered by the long-segment sort
  unsigned long long pairs;                // point pairs of the P2P list
  // near-field work items, partial-sum slots, combines (taken by atomics)
  alignas(128) uint32_t near_items;
  alignas(128) uint32_t near_slots;
  alignas(128) uint32_t near_combines;
  // read by every kernel, apart from the counters above
  alignas(128) uint32_t abort;             // a device-side size check failed: later kernels skip the tick
  uint32_t ncg;                            // cells, or 0 when aborted (the count the per-cell kernels use)
  alignas(128) uint32_t big_n[FMM_LEVELS + 1];  // cells of more than FMM_DIRECT points per depth
  uint32_t mid_n;                          // cells of the upward pass's wave-per-cell list
};

struct FmmBuffers {
  uint32_t cell_cap = 0, point_cap = 0, front_cap = 0, m2l_cap = 0, p2p_cap = 0, cellarr_cap = 0;
  uint32_t graph_cap = 0, chunk_cap = 0;
  FmmGraph* graphs = nullptr;              // per graph
  uint64_t* seeds = nullptr;
  uint32_t* node_off_dev = nullptr;
  uint32_t *chunk_lo = nullptr, *chunk_hi = nullptr, *chunk_first = nullptr;
  void* box_parts = nullptr;
  std::vector<uint32_t> node_off_host;     // the forest the chunk arrays describe
  uint32_t chunks = 0;
  size_t exp_cap = 0;                      // complex coefficients mexp and lexp hold
  FCell* cells = nullptr;
  double* bbox = nullptr;                  // 4 per cell
  uint32_t* parent = nullptr;              // per cell (cell_cap): its parent, NONE for a root
  uint16_t* fl = nullptr;                  // per point: first depth it starts a cell | (leaf depth + 1) << 8
  uint32_t blk_cap = 0;                    // blocks of the cell build the per-block arrays hold
  uint32_t *blk_cnt = nullptr, *blk_fb = nullptr, *blk_off = nullptr, *blk_nb = nullptr;  // depth-major
  uint32_t big_cap = 0;                    // per depth
  uint32_t* big = nullptr;                 // cells of more than FMM_DIRECT points, per depth
  uint32_t mid_cap = 0;
  uint32_t* mid = nullptr;                 // cells the upward pass gives a wave each
  float *mexp = nullptr, *lexp = nullptr;  // 2(p+1) per cell
  double *px = nullptr, *py = nullptr, *ux = nullptr, *uy = nullptr;
  float* ps = nullptr;
  uint2 *front_a = nullptr, *front_b = nullptr;
  uint32_t long_cap = 0, slot_cap = 0, item_cap = 0;
  // launch shapes (no result depends on them): the dual traversal's grid, the
  // grid of the per-cell upward pass (4 waves a block), the blocks sharing
  // near-field work items
  uint32_t dual_blocks = 1024, level_blocks = 16384, near_extra = 256;
  // the last eager tree's deepest level + 2: the depths the capturable form's
  // loops cover (a deeper tree aborts the replay)
  uint32_t build_levels = FMM_LEVELS;
  uint32_t graph_levels = FMM_LEVELS;
  uint32_t* long_list = nullptr;           // targets whose segments the long-segment sort orders
  uint4* near_items = nullptr;             // (leaf, target batch, chunk or ALL, slot)
  uint4* near_combines = nullptr;          // (leaf, target batch, first slot, chunks)
  float2* near_partial = nullptr;          // slot_cap x 64 per-target partial sums
  uint64_t *m2l = nullptr, *p2p = nullptr;         // target << nbits | source, as emitted
  uint32_t *m2l_src = nullptr, *p2p_src = nullptr;  // sources grouped by target
  uint32_t *m2l_off = nullptr, *m2l_cnt = nullptr, *p2p_off = nullptr, *p2p_cnt = nullptr, *cursor = nullptr;
  uint32_t* scan_partials = nullptr;
  FmmParams* params = nullptr;       // device
  FmmParams* params_host = nullptr;  // pinned
  hipEvent_t ev_geom = nullptr, ev_mp = nullptr;  // the upward pass's fork to and join from an auxiliary stream
};
void fmm_free(FmmBuffers& f);
struct FmmArgs {
  const double *x, *y, *strength;
  double *vx, *vy;
  uint32_t n, order;
  double alpha, opening, dmin2, dmax2;
  // The forest: graphs g own nodes [node_off[g], node_off[g + 1]) (host
  // array of graphs + 1 entries); node_graph (device, n entries) maps a node
  // to its graph and may be null for one graph; seeds (host) holds each
  // graph's jiggle seed.
  uint32_t graphs = 1;
  const uint32_t* node_off = nullptr;
  const uint32_t* node_graph = nullptr;
  const uint64_t* seeds = nullptr;
  // Capturable form: no host synchronisation, launches sized by the buffers'
  // capacities (reserved beforehand by fmm_reserve_capacity) and by counts
  // read on the device; alpha read from alpha_dev. The caller reads
  // params_host after the tick and redoes it eagerly when abort is set.
  bool graph = false;
  const double* alpha_dev = nullptr;
  const uint64_t* seeds_dev = nullptr;  // device seeds, in place of the host ones
  // Capturable form only: when nonzero, the size guard treats this many cells
  // as the capacity (a tick with more cells aborts and is redone eagerly).
  uint32_t cell_limit = 0;
  // Graphs with fewer nodes than this get no tree: their nodes are left as
  // they are (the batch ticks them with the exact kernels).
  uint32_t min_graph_n = 0;
  // Eager form only: when set, the multipoles run on this stream beside the
  // dual traversal (forked after the geometry, joined before M2L).
  hipStream_t aux = nullptr;
  // when set, recorded on the stream once the dual traversal is done (a
  // caller's own work on nodes the FMM does not write can start there; the
  // FMM may return before recording it, so the caller records it first)
  hipEvent_t built = nullptr;
};
// Fast many-body (fmm.zig:247-354) over every graph of the forest at once.
// With one graph, returns hipErrorNotSupported when the FMM declines (a
// non-finite extent) and the caller then runs the exact kernel, as the
// reference does; with more, declined graphs are counted in
// params_host->declined and left unchanged.
hipError_t fmm_run(FmmBuffers& f, TreeBuffers& t, const FmmArgs& a, hipStream_t stream);
// Sizes every FMM buffer for the capturable form: cells for `cells` cells,
// the lists, near-field items and expansions for what those cells allow.
hipError_t fmm_reserve_capacity(FmmBuffers& f, TreeBuffers& t, const FmmArgs& a, uint32_t cells, hipStream_t stream);
// out[2q], out[2q + 1] = FP64 direct-sum many-body velocity increment of node
// sample[q] (device pointers).
hipError_t direct_sum(const double* x, const double* y, const double* strength, uint32_t n, const uint32_t* sample,
                      uint32_t k, double alpha, double dmin2, double dmax2, double* out, hipStream_t stream);

// ---- simple.hip ----
// Workspace of the chunked center sums (per axis and chunk of 32 nodes).
struct CenterWork {
  uint32_t cap = 0;                 // chunks per axis
  double* csum = nullptr;           // 2 cap: chunk sums in any order (the binade guesses)
  double* start = nullptr;          // 2 cap: exclusive scan of csum
  long long* ct = nullptr;          // 2 cap x 6: chunk transforms in ulp units
  uint32_t* meta = nullptr;         // 2 cap: binade, sign, valid
  uint32_t* need = nullptr;         // 2 cap: chunk predicted to need its values
  uint32_t* fidx = nullptr;         // 2 cap: exclusive scan of need
  uint32_t* segflag = nullptr;      // 2 cap: a segment starts at the chunk
  uint32_t* segidx = nullptr;       // 2 cap: exclusive scan of segflag
  double* segtmp = nullptr;         // 2 cap x 8: segment records at their first chunk
  double* seglist = nullptr;        // 2 cap x 8: segment records in order
  uint32_t* round_f = nullptr;      // 2 (cap / 128 + 2): staged index at each staging round
  uint32_t* segrange = nullptr;     // 4: first segment and end of each axis
  double* staged = nullptr;         // 2 cap x 32: values of needing chunks, in order
  uint32_t* partials = nullptr;
  uint32_t* count = nullptr;
};
void center_work_free(CenterWork& w);
// sums[0..1] = d3's serial sums of x and y, bit for bit.
hipError_t center_raw_sum(const double* x, const double* y, uint32_t n, double* sums, CenterWork& w,
                          hipStream_t stream);
// x -= (sums[0] / n - cx) * strength, y likewise.
hipError_t center_apply(double* x, double* y, uint32_t n, const double* sums, double cx, double cy, double strength,
                        hipStream_t stream);
hipError_t integrate(double* x, double* y, double* vx, double* vy, const double* fx, const double* fy,
                     const uint8_t* fixed, uint32_t n, double decay, hipStream_t stream);
hipError_t axis_force(const double* pos, double* v, const double* strength, const double* target, uint32_t n,
                      double alpha, hipStream_t stream, const double* alpha_dev = nullptr);
hipError_t radial_force(const double* x, const double* y, double* vx, double* vy, const double* strength,
                        const double* radius, uint32_t n, double cx, double cy, double alpha, hipStream_t stream,
                        const double* alpha_dev = nullptr);
hipError_t shift_positions(const double* x, const double* y, const double* vx, const double* vy, double* px,
                           double* py, uint32_t n, hipStream_t stream);
// out[0..n) = vx, out[n..2n) = vy and, when x is not null, out[2n..3n) = x,
// out[3n..4n) = y (sim_snapshot).
hipError_t snapshot_copy(const double* vx, const double* vy, const double* x, const double* y, uint32_t n,
                         double* out, hipStream_t stream);

// ---- captured ticks (simple.hip) ----
// A simulation's tick captured once as a HIP graph and replayed: what
// changes between ticks (alpha, the fast forces' seeds) is read on the
// device from TickParams[tick], the tick's position in the uploaded run.
constexpr int TICK_SEEDS = 8;        // seeded forces a captured tick holds
constexpr uint32_t TICK_RUN = 256;   // ticks one upload of TickParams covers
struct TickParams {
  double alpha;
  uint64_t seed[TICK_SEEDS];
};
struct TickState {
  double alpha;                 // the running tick's
  uint64_t seed[TICK_SEEDS];
  uint32_t tick;                // index of the running tick in the TickParams run
  uint32_t first_fail;          // the first tick whose replay could not finish it (NONE: none)
  uint32_t status;              // that tick's flags
  uint32_t pad;
};
// Flags a replay raises (any non-zero word) for work only the eager tick does.
struct TickNotes {
  const uint32_t* flag[8];
  uint32_t bit[8];
  uint32_t count;
};
// Start of a replay: the tick's alpha and seeds, and `nzero` words cleared.
hipError_t tick_begin(TickState* st, const TickParams* params, uint32_t* zero, uint32_t nzero, hipStream_t stream);
// The tick's starting x, y, vx, vy, kept until a replay fails: snap holds
// the state at the start of the first failing tick.
hipError_t tick_snapshot(const TickState* st, const double* x, const double* y, const double* vx, const double* vy,
                         uint32_t n, double* snap, hipStream_t stream);
// End of a replay: the flags noted, the tick counted.
hipError_t tick_end(TickState* st, const TickNotes& notes, hipStream_t stream);
// pos[i] = (x + vx, y + vy): the shifted positions as pairs, so the fast link
// sweep gathers an endpoint from one place.
hipError_t shift_pack(const double* x, const double* y, const double* vx, const double* vy, double2* pos, uint32_t n,
                      hipStream_t stream);
// A batch of graphs: sums[2g], sums[2g + 1] = d3's serial sums of graph g's
// x and y (nodes node_off[g] .. node_off[g + 1] - 1 in order), then
// x -= (sx / n_g - cx) * strength per node (simple.zig:14-31 per graph).
hipError_t batch_center(double* x, double* y, const uint32_t* node_off, const uint32_t* node_graph, uint32_t graphs,
                        uint32_t n, double* sums, double cx, double cy, double strength, hipStream_t stream);

}  // namespace d3f
