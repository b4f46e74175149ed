// hsrq_host.h -- host-side internals shared by the plan (hsrq_api.cpp) and
// the request server (hsrq_server.cpp): the context, error plumbing, the
// per-problem workspace layout and descriptor, the item-table builders of
// one sweep step, and the kernel launchers (hsrq_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <vector>

#include "../../include/hsrq.h"
#include "hsrq_internal.h"

struct hsrq_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int tile = 0;          // 0: choose per problem
  int exact = 0;
  int precision = HSRQ_PREC_FP64;
  int lookahead = -1;    // k > 0: bulk panel pass on a second stream, k blocks per
                         // CU; 0: off; -1: automatic
  int cus = 0;           // compute units of the device
  int pack = 0;          // shifts per wavefront in the diagonal sweep: 1, 4 (0: auto)
  int lanes = 0;         // concurrent lanes of a plan (0: auto)
  void* post = nullptr;  // descriptors and items of the products after the solve
  size_t post_cap = 0;
};

namespace hsrq {

hipError_t& last_hip_error();
// Frees the context's buffer of the products after the solve (hsrq_post.cpp).
void release_post(hsrq_context* ctx);
// FP64 residuals of the problems' eigenvectors (host res, sum of m entries;
// hsrq_post.cpp).  Synchronous.
int residuals_run(hsrq_context* ctx, int count, const hsrq_problem* probs, double* res);
// Device address of the allocation-free entry point's scratch array on the
// current device (hsrq_noalloc.hip); once per device.
void noalloc_setup(int device);

#define HIP_TRY(expr)                                                 \
  do {                                                                \
    hipError_t e_ = (expr);                                           \
    if (e_ != hipSuccess) {                                           \
      ::hsrq::last_hip_error() = e_;                                  \
      return e_ == hipErrorOutOfMemory ? HSRQ_ENOMEM : HSRQ_EHIP;     \
    }                                                                 \
  } while (0)

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// Shifts in one launch from which the automatic layout packs 4 shifts per
// wavefront (beyond 16 workgroups of 4 wavefronts per CU at 1 per wavefront).
constexpr int kPackedMinShifts = 4096;
// Shifts from which a plan of few problems cuts them into two lanes.
constexpr int kSplitMinShifts = 1 << 30;

int validate_problem(const hsrq_problem& q);
// Tile rows for a problem of order n: explicit request, context default or
// the size rule; rounded to a multiple of 64 in [64, 256].
int choose_tile(const hsrq_context* ctx, int tlsz, int n);
// Diagonal-sweep layout: 1 or 4 shifts per wavefront.
int choose_pack(int requested, int b, int total_shifts);
// Workgroups per CU of a panel pass that runs beside the diagonal sweeps of
// this layout and tile, automatic choice: the most that leave room for a
// sweep workgroup for one shift per wavefront, when the lane's shifts take at
// most one sweep workgroup per CU (cus CUs); 0 (no lookahead) otherwise.
int lookahead_blocks_per_cu(int pack, int b, int shifts, int cus);

// Planning (hsrq_sched.cpp).  A contiguous range of one problem's shifts.
// Shifts are independent given H, so a problem can be cut into pieces that
// run in different lanes; every shift's arithmetic is the same in any piece
// (results are bit-identical).
struct Piece {
  int prob, k0, m;
};
// Tile of the automatic policy for a problem of order n with `shifts`
// shifts among `count`, on a device of `cus` compute units.
int auto_tile(int n, long long shifts, int count, int cus);
// Lanes of a plan: groups of pieces and the tile of each group.
void assign_lanes(const hsrq_context* ctx, int count, const hsrq_problem* in, int tlsz,
                  std::vector<std::vector<Piece>>& groups, std::vector<int>& tiles);
// Start step of each problem of a lane of `steps` steps.
std::vector<int> step_offsets(const std::vector<ProbDev>& probs, int steps);

// Device workspace of one problem (bytes, 256-aligned pieces).
// Workspace bytes of one problem; eq: with the FP32 panel pass's operands.
size_t prob_workspace_bytes(int n, int m, int b, bool eq = false);
// Fills the descriptor of problem q with tile b, start vector `its` (or per
// shift its_k, device, m entries) and the workspace at `ws`
// (prob_workspace_bytes(q.n, q.m, b) bytes).  eq: the problem runs the FP32
// panel pass (its equilibration exponents are set up).
void prob_describe(ProbDev& P, const hsrq_problem& q, int b, int its, int exact, char* ws,
                   const int* its_k = nullptr, bool eq = false);

// Item tables.  `i` is the problem's index in the descriptor array.  The
// builders append to any container with push_back (std::vector, or the
// fixed-capacity host buffers of the allocation-free entry point).
bool norm_block_needed_host(const ProbDev& P, int rb, int cb);
template <class GV, class RV>
void add_norm_items(const ProbDev& P, int i, GV& part, RV& fin) {
  if (P.m == 0) return;
  for (int t = 0; t < P.T; ++t) fin.push_back(RowItem{i, t, 0, 0});
  for (int rb = 0; rb < P.nrb; ++rb)
    for (int cb = 0; cb < P.nrb; ++cb)
      if (norm_block_needed_host(P, rb, cb)) part.push_back(GemmItem{i, 0, rb, cb, 1, 0});
}
template <class RV>
void add_init_items(const ProbDev& P, int i, RV& init) {
  for (int rb = 0; rb < P.nrb; ++rb)
    for (int k0 = 0; k0 < P.m; k0 += kShiftsPerK1) init.push_back(RowItem{i, rb, k0, 0});
}
// Items of tile column J of problem i (J = T-1 ... 0): the diagonal sweep
// (k1) and the panel pass of the rows above tile J-1 (gb; tile J-1's own
// rows are updated at the start of its sweep).  The panel pass goes out in
// row-block pairs, XCD-aware: blocks are dealt to the 8 XCDs round robin, so
// pairs 8t .. 8t+7 go out interleaved and every XCD sweeps all column blocks
// of one pair from its own L2.
template <class KV, class GV>
void add_step_items(const ProbDev& P, int i, int J, int pack, KV& k1, GV& gb) {
  if (P.m == 0 || J < 0 || J >= P.T) return;
  const int k1step = pack == 4 ? 16 : kShiftsPerK1;
  for (int k0 = 0; k0 < P.m; k0 += k1step) k1.push_back(K1Item{i, J, k0, 0});
  if (J == 0) return;
  const int rbn = tile_first_rowblock(P, J - 1);
  const int ncb = (2 * P.m + kGemmCols - 1) / kGemmCols;
  const int npairs = (rbn + kGemmRowBlocks - 1) / kGemmRowBlocks;
  for (int t = 0; t < npairs; t += 8) {
    const int te = npairs < t + 8 ? npairs : t + 8;
    for (int cb = 0; cb < ncb; ++cb)
      for (int q = t; q < te; ++q) {
        const int rb = q * kGemmRowBlocks;
        const int nrbk = rbn - rb < kGemmRowBlocks ? rbn - rb : kGemmRowBlocks;
        gb.push_back(GemmItem{i, J, rb, cb, nrbk, 0});
      }
  }
}
template <class SV>
void add_bt_items(const ProbDev& P, int i, SV& bt) {
  for (int k = 0; k < P.m; ++k) bt.push_back(ShiftItem{i, k, 0, 0});
}
struct StepItems {
  std::vector<K1Item> k1;          // diagonal sweep
  std::vector<GemmItem> gb;        // panel pass of the rows above the next diagonal
                                   // tile (that tile's own rows: in its sweep)
  void clear() {
    k1.clear();
    gb.clear();
  }
};
inline void add_step_items(const ProbDev& P, int i, int J, int pack, StepItems& out) {
  add_step_items(P, i, J, pack, out.k1, out.gb);
}

// Launchers (hsrq_kernels.hip).  Zero counts launch nothing.
hipError_t launch_norms(const ProbDev*, const GemmItem*, int, const RowItem*, int, hipStream_t);
// Equilibration exponents of H for the FP32 panel pass (ProbDev::ccol, rrow)
// of problem d_prob (device descriptor) of order n with nrb row blocks.
hipError_t launch_equilibrate(const ProbDev* d_prob, int n, int nrb, hipStream_t);
hipError_t launch_init(const ProbDev*, const RowItem*, int, hipStream_t);
hipError_t launch_diag(int rpl, const ProbDev*, const K1Item*, int, hipStream_t);
hipError_t launch_diag16(int rpl, const ProbDev*, const K1Item*, int, hipStream_t);
// Panel pass; max_blocks > 0 caps the grid (blocks loop over the items).
hipError_t launch_update(const ProbDev*, const GemmItem*, int, hipStream_t, int max_blocks = 0);
hipError_t launch_update32(const ProbDev*, const GemmItem*, int, hipStream_t,
                           int max_blocks = 0);
hipError_t launch_backtransform(const ProbDev*, const ShiftItem*, int, hipStream_t);
hipError_t launch_mfma_selftest(int*, hipStream_t);
// ||H(i,:)||_1 per row into rs (n); work: 64 n doubles.
hipError_t launch_rowsums(int n, const double* H, int64_t ldH, double* work, double* rs,
                          hipStream_t st);
// Diagonal sweep of tile size b in the given layout.
hipError_t launch_sweep(int pack, int b, const ProbDev*, const K1Item*, int, hipStream_t);
// Whether launch_step serves plans of this layout and tile size.
bool step_fusable(int pack, int b);
// One launch: the one-shift sweeps k1 of a step beside the panel pass gb of
// the step before (capped at max_blocks blocks that loop over its items).
hipError_t launch_step(int b, const ProbDev*, const K1Item* k1, int nk1, const GemmItem* gb,
                       int ngb, int max_blocks, hipStream_t);

}  // namespace hsrq
