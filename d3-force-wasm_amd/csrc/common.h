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

// The traversal's 64-byte view of a cell, one scalar load per visit:
// centre of charge, charge, the opening threshold w*w/theta2, children, the
// leaf chain and flags (bit 0 leaf, bit 1 "a truthy leaf is reachable").
struct TCell {
  double qx, qy, value, w2t;
  uint32_t child[4];
  uint32_t first, count, head, flags;
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
};

// A depth-first frame of the below-key-resolution build (tree.zig:64-76).
struct Frame {
  uint32_t lo, hi, depth, parent, slot, pad;
  double x0, y0, x1, y1;
};
constexpr uint32_t DEEP_STACK = 3 * MAX_DEPTH + 64;

#define HIPCHECK(expr)                          \
  do {                                          \
    hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return e_;            \
  } while (0)

__host__ __device__ inline uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

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
  uint32_t *cell_s = nullptr, *cell_d = nullptr;     // each cell's key group and depth
  uint32_t* rs_glob = nullptr;                       // radix digit totals, 8 passes
  uint32_t *dbins = nullptr, *dstart = nullptr, *dorder = nullptr;  // cells by depth
  TreeMeta* meta = nullptr;       // device
  TreeMeta* meta_host = nullptr;  // pinned mirror
};

hipError_t tree_reserve(TreeBuffers& t, uint32_t n);
hipError_t tree_reserve_cells(TreeBuffers& t, uint32_t cells);
void tree_free(TreeBuffers& t);
// Builds the exact quadtree over (x, y)[0..n) on `stream`; on return
// meta_host holds the tree's scalars (the call synchronises once, after the
// cell count is known, to size the cell buffers).
hipError_t tree_build_exact(TreeBuffers& t, const double* x, const double* y, uint32_t n, hipStream_t stream);

// ---- many_body.hip ----
// visitAfter charge sums (many_body.zig:101-135) into cells and tcells.
hipError_t mb_accumulate(TreeBuffers& t, const double* x, const double* y, const double* strengths, double theta2,
                         hipStream_t stream);
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
  uint32_t stack_cap;      // entries per wave
  unsigned long long* counters;  // visits, accepts, pairs (summed)
};
hipError_t mb_traverse(const TravArgs& a, int mode, hipStream_t stream);
// Nodes with a NaN coordinate (tree.rest): their traversal reaches every
// truthy leaf through truthy cells, and any reached leaf makes v NaN.
hipError_t mb_nan_nodes(const TreeBuffers& t, uint32_t n, double* vx, double* vy, hipStream_t stream);

// ---- link.hip ----
struct LinkLevelArgs {
  const uint32_t *src, *tgt;
  const double *strength, *distance, *bias;
  const double *x, *y;
  double *vx, *vy;
  double alpha;
  uint32_t* zero_flag;
};
hipError_t link_level(const LinkLevelArgs& a, uint32_t lo, uint32_t hi, hipStream_t stream);
struct LinkSerialState {
  uint32_t edge;    // next edge; m when done
  uint32_t need;    // bit 0: x needs a draw, bit 1: y needs a draw
  double jig[2];    // drawn jiggles for `edge` (x then y)
  uint32_t has_jig;
};
hipError_t link_serial(const uint32_t* src, const uint32_t* tgt, const double* strength, const double* distance,
                       const double* bias, uint32_t m, const double* x, const double* y, double* vx, double* vy,
                       double alpha, LinkSerialState* state, hipStream_t stream);

// ---- simple.hip ----
// sums[0..1] = d3's serial sums of x and y (one workgroup).
hipError_t center_raw_sum(const double* x, const double* y, uint32_t n, double* sums, hipStream_t stream);
// x -= (sums[0] / n - cx) * strength, y likewise.
hipError_t center_apply(double* x, double* y, uint32_t n, const double* sums, double cx, double cy, double strength,
                        hipStream_t stream);
hipError_t integrate(double* x, double* y, double* vx, double* vy, const double* fx, const double* fy,
                     const uint8_t* fixed, uint32_t n, double decay, hipStream_t stream);
hipError_t axis_force(const double* pos, double* v, const double* strength, const double* target, uint32_t n,
                      double alpha, hipStream_t stream);
hipError_t radial_force(const double* x, const double* y, double* vx, double* vy, const double* strength,
                        const double* radius, uint32_t n, double cx, double cy, double alpha, hipStream_t stream);
hipError_t shift_positions(const double* x, const double* y, const double* vx, const double* vy, double* px,
                           double* py, uint32_t n, hipStream_t stream);

}  // namespace d3f
