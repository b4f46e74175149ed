// hsrq_api.cpp -- host side of libhsrq: planner, launch schedule, restart
// logic, C-ABI (hsrq.h) and the drop-in tiled_inverse_iteration
// (reference inverse_iteration.h:19-22, inverse_iteration.c:128-335).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/hsrq.h"
#include "../../include/inverse_iteration.h"
#include "hsrq_internal.h"

namespace hsrq {
hipError_t launch_norms(const ProbDev*, const GemmItem*, int, const RowItem*, int, hipStream_t);
bool norm_block_needed_host(const ProbDev& P, int rb, int cb);
hipError_t launch_init(const ProbDev*, const RowItem*, int, hipStream_t);
hipError_t launch_diag(int, const ProbDev*, const K1Item*, int, hipStream_t);
hipError_t launch_update(const ProbDev*, const GemmItem*, int, hipStream_t);
hipError_t launch_params(const ProbDev*, const K1Item*, int, hipStream_t);
hipError_t launch_diag16(int, const ProbDev*, const K1Item*, int, hipStream_t);
hipError_t launch_backtransform(const ProbDev*, const ShiftItem*, int, hipStream_t);
hipError_t launch_mfma_selftest(int*, hipStream_t);
hipError_t launch_residual(int n, const double* H, int64_t ldH, int m, const double* wr,
                           const double* X, int64_t ldX, double* work, double* res,
                           hipStream_t st);
}  // namespace hsrq

using namespace hsrq;

#define HIP_TRY(expr)                         \
  do {                                        \
    hipError_t e_ = (expr);                   \
    if (e_ != hipSuccess) {                   \
      last_hip_error() = e_;                  \
      return e_ == hipErrorOutOfMemory ? HSRQ_ENOMEM : HSRQ_EHIP; \
    }                                         \
  } while (0)

static hipError_t& last_hip_error() {
  static thread_local hipError_t e = hipSuccess;
  return e;
}

struct hsrq_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int tile = 0;          // 0: choose per problem
  int exact = 0;
  int precision = HSRQ_PREC_FP64;
  int lookahead = 0;     // 1: bulk panel pass on a second stream
  int pack = 0;          // shifts per wavefront in the diagonal sweep: 1, 4 (0: auto)
};

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

int choose_tile(const hsrq_context* ctx, int tlsz, int n) {
  int t = tlsz > 0 ? tlsz : ctx->tile;
  if (t <= 0) t = n >= 4096 ? 128 : 64;
  t = ((t + kRowBlock - 1) / kRowBlock) * kRowBlock;
  return std::min(256, std::max(64, t));
}

}  // namespace

// A plan: device problem descriptors, workspace and the launch schedule of a
// set of problems sharing one tile size.
struct hsrq_plan {
  hsrq_context* ctx = nullptr;
  int b = 64;
  int count = 0;
  std::vector<ProbDev> probs;         // host copy
  std::vector<int> status_off;        // prefix of m
  int total_m = 0;
  int steps = 0;
  // device allocations
  void* d_work = nullptr;
  size_t work_bytes = 0;
  ProbDev* d_probs = nullptr;
  void* d_items = nullptr;
  // item tables (device offsets)
  const RowItem* d_norm_items = nullptr;
  int n_norm_items = 0;
  const GemmItem* d_part_items = nullptr;
  int n_part_items = 0;
  const RowItem* d_init_items = nullptr;
  int n_init_items = 0;
  const ShiftItem* d_bt_items = nullptr;
  int n_bt_items = 0;
  std::vector<const K1Item*> k1_ptr, pr_ptr;
  std::vector<int> k1_cnt, pr_cnt;
  int pack = 1;   // shifts per wavefront of the diagonal sweep
  // panel passes: rows of the next diagonal tile ("a", on the main stream,
  // it gates the next diagonal sweep) and all rows above it ("b", on the
  // side stream, overlapping the next diagonal sweep)
  std::vector<const GemmItem*> ga_ptr, gb_ptr;
  std::vector<int> ga_cnt, gb_cnt;
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::vector<hipEvent_t> ev_p, ev_b;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

namespace {

void plan_free(hsrq_plan* p) {
  if (!p) return;
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (auto e : p->ev_p) (void)hipEventDestroy(e);
  for (auto e : p->ev_b) (void)hipEventDestroy(e);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->ev_join) (void)hipEventDestroy(p->ev_join);
  if (p->side) (void)hipStreamDestroy(p->side);
  if (p->d_work) (void)hipFree(p->d_work);
  if (p->d_probs) (void)hipFree(p->d_probs);
  if (p->d_items) (void)hipFree(p->d_items);
  delete p;
}

int validate_problem(const hsrq_problem& q) {
  if (q.n <= 0 || q.m < 0 || q.ldH < q.n || q.ldvr < q.n || !q.H) return HSRQ_EINVAL;
  if (q.m > 0 && (!q.wr || !q.vr)) return HSRQ_EINVAL;
  return HSRQ_OK;
}

// Builds problem descriptors, workspace and item tables.  its: start vector.
int plan_build(hsrq_context* ctx, int count, const hsrq_problem* in, int tlsz, int its,
               hsrq_plan** out) {
  *out = nullptr;
  if (!ctx || count < 0 || (count > 0 && !in)) return HSRQ_EINVAL;
  for (int i = 0; i < count; ++i) {
    const int rc = validate_problem(in[i]);
    if (rc) return rc;
  }
  HIP_TRY(hipSetDevice(ctx->device));
  hsrq_plan* p = new (std::nothrow) hsrq_plan();
  if (!p) return HSRQ_ENOMEM;
  p->ctx = ctx;
  int nmax = 0;
  for (int i = 0; i < count; ++i) nmax = std::max(nmax, in[i].n);
  p->b = choose_tile(ctx, tlsz, nmax);
  const int b = p->b;

  // --- descriptors and workspace layout ---
  std::vector<size_t> off;   // per problem: offset table
  size_t bytes = 0;
  p->probs.resize(count);
  p->status_off.resize(count + 1, 0);
  struct Lay {
    size_t rt0, rt1, c, s, bm0, bm1, aux0, aux1, ynb, rowinf, diagn, part, scal, e, sy, sz, st;
  };
  std::vector<Lay> lay(count);
  for (int i = 0; i < count; ++i) {
    const hsrq_problem& q = in[i];
    ProbDev& P = p->probs[i];
    std::memset(&P, 0, sizeof(P));
    P.H = q.H;
    P.ldH = q.ldH;
    P.n = q.n;
    P.m = q.m;
    P.b = b;
    P.T = (q.n + b - 1) / b;
    P.b0 = q.n - (P.T - 1) * b;
    P.nb0blk = (P.b0 + kRowBlock - 1) / kRowBlock;
    P.nrb = P.nb0blk + (P.T - 1) * (b / kRowBlock);
    P.wr = q.wr;
    P.Y = q.vr;
    P.ldY = q.ldvr;
    P.its = its;
    P.force_exact = ctx->exact;
    const size_t nm = (size_t)q.n * q.m * sizeof(double);
    const size_t Tm = (size_t)P.T * q.m;
    Lay& L = lay[i];
    L.rt0 = bytes; bytes += align_up(nm);
    L.rt1 = bytes; bytes += align_up(nm);
    L.c = bytes; bytes += align_up(nm);
    L.s = bytes; bytes += align_up(nm);
    L.bm0 = bytes; bytes += align_up((size_t)b * 2 * q.m * sizeof(double));
    L.bm1 = bytes; bytes += align_up((size_t)b * 2 * q.m * sizeof(double));
    L.aux0 = bytes; bytes += align_up((size_t)4 * q.m * sizeof(double));
    L.aux1 = bytes; bytes += align_up((size_t)4 * q.m * sizeof(double));
    L.ynb = bytes; bytes += align_up((size_t)P.nrb * q.m * sizeof(double));
    L.rowinf = bytes; bytes += align_up((size_t)P.T * sizeof(double));
    L.diagn = bytes; bytes += align_up((size_t)P.T * sizeof(double));
    L.part = bytes; bytes += align_up((size_t)P.nrb * q.n * sizeof(double));
    L.scal = bytes; bytes += align_up(4 * sizeof(double));
    L.e = bytes; bytes += align_up(Tm * sizeof(int));
    L.sy = bytes; bytes += align_up(Tm * sizeof(int));
    L.sz = bytes; bytes += align_up(Tm * sizeof(int));
    L.st = bytes; bytes += align_up((size_t)std::max(q.m, 1) * sizeof(int));
    p->status_off[i + 1] = p->status_off[i] + q.m;
  }
  p->total_m = p->status_off[count];
  p->count = count;
  p->work_bytes = bytes;
  hipError_t e = hipMalloc(&p->d_work, std::max<size_t>(bytes, kAlign));
  if (e != hipSuccess) {
    last_hip_error() = e;
    plan_free(p);
    return HSRQ_ENOMEM;
  }
  char* base = (char*)p->d_work;
  for (int i = 0; i < count; ++i) {
    ProbDev& P = p->probs[i];
    const Lay& L = lay[i];
    P.Rt[0] = (double*)(base + L.rt0);
    P.Rt[1] = (double*)(base + L.rt1);
    P.C = (double*)(base + L.c);
    P.S = (double*)(base + L.s);
    P.Bm[0] = (double*)(base + L.bm0);
    P.Bm[1] = (double*)(base + L.bm1);
    P.aux[0] = (double*)(base + L.aux0);
    P.aux[1] = (double*)(base + L.aux1);
    P.Ynb = (double*)(base + L.ynb);
    P.rowinf = (double*)(base + L.rowinf);
    P.diagn = (double*)(base + L.diagn);
    P.part = (double*)(base + L.part);
    P.scal = (double*)(base + L.scal);
    P.E = (int*)(base + L.e);
    P.SY = (int*)(base + L.sy);
    P.SZ = (int*)(base + L.sz);
    P.status = (int*)(base + L.st);
  }

  // --- item tables ---
  std::vector<RowItem> norm_items, init_items;
  std::vector<GemmItem> part_items;
  std::vector<ShiftItem> bt_items;
  int steps = 0;
  for (int i = 0; i < count; ++i) {
    const ProbDev& P = p->probs[i];
    if (P.m == 0) continue;
    steps = std::max(steps, P.T);
    for (int t = 0; t < P.T; ++t) norm_items.push_back({i, t, 0, 0});
    for (int rb = 0; rb < P.nrb; ++rb)
      for (int cb = 0; cb < P.nrb; ++cb)
        if (norm_block_needed_host(P, rb, cb)) part_items.push_back({i, 0, rb, cb, 1, 0});
    for (int rb = 0; rb < P.nrb; ++rb)
      for (int k0 = 0; k0 < P.m; k0 += kShiftsPerK1) init_items.push_back({i, rb, k0, 0});
    for (int k = 0; k < P.m; ++k) bt_items.push_back({i, k, 0, 0});
  }
  p->steps = steps;
  std::vector<std::vector<K1Item>> k1(steps), pr(steps);
  // packed sweep (4 shifts per wavefront, 16 per workgroup) unless asked
  // otherwise: it issues ~4x fewer instructions per shift step; one shift
  // per wavefront keeps the shortest latency per step for single problems.
  p->pack = ctx->pack ? ctx->pack : 4;
  if (b % 64 != 0 || (b / 16 & (b / 16 - 1)) != 0) p->pack = 1;   // 16 lanes x 2^k rows
  const int k1step = p->pack == 4 ? 16 : kShiftsPerK1;
  std::vector<std::vector<GemmItem>> ga(steps), gb(steps);
  for (int s = 0; s < steps; ++s) {
    for (int i = 0; i < count; ++i) {
      const ProbDev& P = p->probs[i];
      if (P.m == 0) continue;
      const int J = P.T - 1 - s;
      if (J < 0) continue;
      for (int k0 = 0; k0 < P.m; k0 += k1step) k1[s].push_back({i, J, k0, 0});
      for (int k0 = 0; k0 < P.m; k0 += kShiftsPerK1) pr[s].push_back({i, J, k0, 0});
      if (J >= 1) {
        // rows of the next diagonal tile (part a) and all rows above (part
        // b), in pairs of row blocks; XCD-aware order: blocks are dealt to
        // the 8 XCDs round robin, so the pairs 8t..8t+7 go out interleaved
        // and every XCD sweeps all column blocks of one pair from its L2.
        const int nrbJ = tile_first_rowblock(P, J);
        const int rbn = tile_first_rowblock(P, J - 1);
        const int ncb = (2 * P.m + kGemmCols - 1) / kGemmCols;
        auto emit = [&](std::vector<GemmItem>& dst, int rs, int re) {
          std::vector<std::pair<int, int>> pairs;
          for (int rb = rs; rb < re; rb += kGemmRowBlocks)
            pairs.push_back({rb, std::min(kGemmRowBlocks, re - rb)});
          for (size_t t = 0; t < pairs.size(); t += 8) {
            const size_t te = std::min(pairs.size(), t + 8);
            for (int cb = 0; cb < ncb; ++cb)
              for (size_t q = t; q < te; ++q)
                dst.push_back({i, J, pairs[q].first, cb, pairs[q].second, 0});
          }
        };
        emit(ga[s], rbn, nrbJ);
        emit(gb[s], 0, rbn);
      }
    }
  }
  size_t nbytes = 0;
  nbytes += align_up(norm_items.size() * sizeof(RowItem));
  nbytes += align_up(part_items.size() * sizeof(GemmItem));
  nbytes += align_up(init_items.size() * sizeof(RowItem));
  nbytes += align_up(bt_items.size() * sizeof(ShiftItem));
  for (int s = 0; s < steps; ++s) {
    nbytes += align_up(k1[s].size() * sizeof(K1Item));
    nbytes += align_up(pr[s].size() * sizeof(K1Item));
    nbytes += align_up(ga[s].size() * sizeof(GemmItem));
    nbytes += align_up(gb[s].size() * sizeof(GemmItem));
  }
  std::vector<char> host(std::max<size_t>(nbytes, kAlign));
  e = hipMalloc(&p->d_items, std::max<size_t>(nbytes, kAlign));
  if (e == hipSuccess) e = hipMalloc(&p->d_probs, std::max<size_t>(count, 1) * sizeof(ProbDev));
  if (e != hipSuccess) {
    last_hip_error() = e;
    plan_free(p);
    return HSRQ_ENOMEM;
  }
  size_t pos = 0;
  auto put = [&](const void* src, size_t sz) -> const void* {
    if (sz) std::memcpy(host.data() + pos, src, sz);
    const void* dptr = (const char*)p->d_items + pos;
    pos += align_up(sz);
    return dptr;
  };
  p->d_norm_items = (const RowItem*)put(norm_items.data(), norm_items.size() * sizeof(RowItem));
  p->n_norm_items = (int)norm_items.size();
  p->d_part_items = (const GemmItem*)put(part_items.data(), part_items.size() * sizeof(GemmItem));
  p->n_part_items = (int)part_items.size();
  p->d_init_items = (const RowItem*)put(init_items.data(), init_items.size() * sizeof(RowItem));
  p->n_init_items = (int)init_items.size();
  p->d_bt_items = (const ShiftItem*)put(bt_items.data(), bt_items.size() * sizeof(ShiftItem));
  p->n_bt_items = (int)bt_items.size();
  p->k1_ptr.resize(steps);
  p->k1_cnt.resize(steps);
  p->pr_ptr.resize(steps);
  p->pr_cnt.resize(steps);
  p->ga_ptr.resize(steps);
  p->ga_cnt.resize(steps);
  p->gb_ptr.resize(steps);
  p->gb_cnt.resize(steps);
  for (int s = 0; s < steps; ++s) {
    p->k1_ptr[s] = (const K1Item*)put(k1[s].data(), k1[s].size() * sizeof(K1Item));
    p->k1_cnt[s] = (int)k1[s].size();
    p->pr_ptr[s] = (const K1Item*)put(pr[s].data(), pr[s].size() * sizeof(K1Item));
    p->pr_cnt[s] = (int)pr[s].size();
    p->ga_ptr[s] = (const GemmItem*)put(ga[s].data(), ga[s].size() * sizeof(GemmItem));
    p->ga_cnt[s] = (int)ga[s].size();
    p->gb_ptr[s] = (const GemmItem*)put(gb[s].data(), gb[s].size() * sizeof(GemmItem));
    p->gb_cnt[s] = (int)gb[s].size();
  }
  e = hipMemcpy(p->d_items, host.data(), nbytes, hipMemcpyHostToDevice);
  if (e == hipSuccess && count > 0)
    e = hipMemcpy(p->d_probs, p->probs.data(), count * sizeof(ProbDev), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    last_hip_error() = e;
    plan_free(p);
    return HSRQ_EHIP;
  }
  e = hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming);
  p->ev_p.assign(steps, nullptr);
  p->ev_b.assign(steps, nullptr);
  for (int s = 0; s < steps && e == hipSuccess; ++s) {
    e = hipEventCreateWithFlags(&p->ev_p[s], hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ev_b[s], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    last_hip_error() = e;
    plan_free(p);
    return HSRQ_EHIP;
  }
  *out = p;
  return HSRQ_OK;
}

// Launch schedule.  Main stream: norms, init, then per step s the diagonal
// sweeps, the scaling exponents (after the previous step's bulk panel pass)
// and the panel pass of the rows of the next diagonal tile; side stream: the
// panel pass of all rows above, which overlaps the next diagonal sweep.
// With `prof`, every launch is bracketed by timing events on its stream and
// kernel_ms/launches accumulate per kind (0 norms, 1 init, 2 diag,
// 3 update, 4 back-transform, 5 params).
struct Prof {
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
};

int plan_enqueue_impl(hsrq_plan* p, Prof* prof) {
  hipStream_t st = p->ctx->stream;
  hipStream_t sd = p->ctx->lookahead ? p->side : st;
  const int rpl = p->b <= 64 ? 1 : (p->b <= 128 ? 2 : 4);
  auto run = [&](int kind, hipStream_t s, auto&& fn) -> int {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (prof) {
      HIP_TRY(hipEventCreate(&e0));
      HIP_TRY(hipEventCreate(&e1));
      HIP_TRY(hipEventRecord(e0, s));
    }
    HIP_TRY(fn(s));
    if (prof) {
      HIP_TRY(hipEventRecord(e1, s));
      prof->ev.push_back(e0);
      prof->ev.push_back(e1);
      prof->kind.push_back(kind);
    }
    return HSRQ_OK;
  };
  int rc;
  rc = run(0, st, [&](hipStream_t s) {
    return launch_norms(p->d_probs, p->d_part_items, p->n_part_items, p->d_norm_items,
                        p->n_norm_items, s);
  });
  if (rc) return rc;
  rc = run(1, st, [&](hipStream_t s) {
    return launch_init(p->d_probs, p->d_init_items, p->n_init_items, s);
  });
  if (rc) return rc;
  HIP_TRY(hipEventRecord(p->ev_fork, st));
  HIP_TRY(hipStreamWaitEvent(sd, p->ev_fork, 0));
  int last_b = -1;
  for (int s = 0; s < p->steps; ++s) {
    rc = run(2, st, [&](hipStream_t x) {
      return p->pack == 4 ? launch_diag16(p->b / 16, p->d_probs, p->k1_ptr[s], p->k1_cnt[s], x)
                          : launch_diag(rpl, p->d_probs, p->k1_ptr[s], p->k1_cnt[s], x);
    });
    if (rc) return rc;
    if (p->ga_cnt[s] + p->gb_cnt[s] == 0) continue;
    if (last_b >= 0) HIP_TRY(hipStreamWaitEvent(st, p->ev_b[last_b], 0));
    rc = run(5, st, [&](hipStream_t x) {
      return launch_params(p->d_probs, p->pr_ptr[s], p->pr_cnt[s], x);
    });
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev_p[s], st));
    rc = run(3, st, [&](hipStream_t x) {
      return launch_update(p->d_probs, p->ga_ptr[s], p->ga_cnt[s], x);
    });
    if (rc) return rc;
    HIP_TRY(hipStreamWaitEvent(sd, p->ev_p[s], 0));
    rc = run(3, sd, [&](hipStream_t x) {
      return launch_update(p->d_probs, p->gb_ptr[s], p->gb_cnt[s], x);
    });
    if (rc) return rc;
    HIP_TRY(hipEventRecord(p->ev_b[s], sd));
    last_b = s;
  }
  // join the side stream (also when it had no work: graph capture needs it)
  HIP_TRY(hipEventRecord(p->ev_join, sd));
  HIP_TRY(hipStreamWaitEvent(st, p->ev_join, 0));
  return run(4, st, [&](hipStream_t s) {
    return launch_backtransform(p->d_probs, p->d_bt_items, p->n_bt_items, s);
  });
}

int plan_enqueue(hsrq_plan* p) { return plan_enqueue_impl(p, nullptr); }

int plan_enqueue_profiled(hsrq_plan* p, double* kernel_ms, int* launches) {
  Prof prof;
  int rc = plan_enqueue_impl(p, &prof);
  if (!rc) {
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    HIP_TRY(hipStreamSynchronize(p->side));
    for (int k = 0; k < 6; ++k) {
      kernel_ms[k] = 0.0;
      launches[k] = 0;
    }
    for (size_t i = 0; i < prof.kind.size(); ++i) {
      float ms = 0.f;
      HIP_TRY(hipEventElapsedTime(&ms, prof.ev[2 * i], prof.ev[2 * i + 1]));
      kernel_ms[prof.kind[i]] += ms;
      launches[prof.kind[i]] += 1;
    }
  }
  for (auto e : prof.ev) (void)hipEventDestroy(e);
  return rc;
}

int plan_read_status(hsrq_plan* p, int* status) {
  for (int i = 0; i < p->count; ++i) {
    const ProbDev& P = p->probs[i];
    if (P.m == 0) continue;
    HIP_TRY(hipMemcpyAsync(status + p->status_off[i], P.status, P.m * sizeof(int),
                           hipMemcpyDeviceToHost, p->ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  return HSRQ_OK;
}

// Solves one batch with restarts: unconverged shifts of each problem are
// moved behind the converged ones (stable, wr and vr together) and solved
// again with the its-th restart vector, at most n times (dlaein semantics;
// the reference's intent at inverse_iteration.c:160-328).
int solve_with_restarts(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                        int* status_out) {
  hsrq_plan* p = nullptr;
  int rc = plan_build(ctx, count, probs, tlsz, -1, &p);
  if (rc) return rc;
  rc = plan_enqueue(p);
  std::vector<int> st(std::max(p->total_m, 1));
  if (!rc) rc = plan_read_status(p, st.data());
  const int b = p->b;
  std::vector<int> off(p->status_off);
  plan_free(p);
  if (rc) return rc;

  for (int i = 0; i < count; ++i) {
    hsrq_problem q = probs[i];
    int* sti = st.data() + off[i];
    int first = 0;
    for (int its = 0; its < q.n; ++its) {
      const int cnt = q.m - first;
      int nc = 0;
      for (int r = first; r < q.m; ++r) nc += (sti[r] & HSRQ_CONVERGED) ? 1 : 0;
      if (nc == cnt) break;
      // stable partition of [first, m): converged first
      std::vector<int> order;
      order.reserve(cnt);
      for (int r = first; r < q.m; ++r)
        if (sti[r] & HSRQ_CONVERGED) order.push_back(r);
      for (int r = first; r < q.m; ++r)
        if (!(sti[r] & HSRQ_CONVERGED)) order.push_back(r);
      std::vector<double> wr_h(q.m);
      HIP_TRY(hipMemcpy(wr_h.data(), q.wr, q.m * sizeof(double), hipMemcpyDeviceToHost));
      double* tmp = nullptr;
      HIP_TRY(hipMalloc(&tmp, (size_t)q.n * cnt * sizeof(double)));
      std::vector<double> wr_new(wr_h);
      std::vector<int> st_new(sti, sti + q.m);
      for (int t = 0; t < cnt; ++t) {
        const int src = order[t];
        wr_new[first + t] = wr_h[src];
        st_new[first + t] = sti[src];
        HIP_TRY(hipMemcpyAsync(tmp + (size_t)t * q.n, q.vr + (size_t)src * q.ldvr,
                               q.n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
      }
      HIP_TRY(hipMemcpy2DAsync(q.vr + (size_t)first * q.ldvr, q.ldvr * sizeof(double), tmp,
                               q.n * sizeof(double), q.n * sizeof(double), cnt,
                               hipMemcpyDeviceToDevice, ctx->stream));
      HIP_TRY(hipMemcpyAsync(q.wr, wr_new.data(), q.m * sizeof(double), hipMemcpyHostToDevice,
                             ctx->stream));
      HIP_TRY(hipStreamSynchronize(ctx->stream));
      (void)hipFree(tmp);
      std::copy(st_new.begin(), st_new.end(), sti);
      first += nc;
      // re-solve the unconverged tail with the its-th restart vector
      hsrq_problem tail = q;
      tail.m = q.m - first;
      tail.wr = q.wr + first;
      tail.vr = q.vr + (size_t)first * q.ldvr;
      hsrq_plan* pt = nullptr;
      rc = plan_build(ctx, 1, &tail, b, its, &pt);
      if (rc) return rc;
      rc = plan_enqueue(pt);
      if (!rc) rc = plan_read_status(pt, sti + first);
      plan_free(pt);
      if (rc) return rc;
    }
  }
  if (status_out) std::copy(st.begin(), st.begin() + off[count], status_out);
  return HSRQ_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------
extern "C" {

int hsrq_context_create(hsrq_context** ctx, int device) {
  if (!ctx) return HSRQ_EINVAL;
  *ctx = nullptr;
  int dev = device;
  if (dev < 0) HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipSetDevice(dev));
  hsrq_context* c = new (std::nothrow) hsrq_context();
  if (!c) return HSRQ_ENOMEM;
  c->device = dev;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    last_hip_error() = e;
    delete c;
    return HSRQ_EHIP;
  }
  *ctx = c;
  return HSRQ_OK;
}

void hsrq_context_destroy(hsrq_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

int hsrq_context_set_tile(hsrq_context* ctx, int tlsz) {
  if (!ctx || tlsz < 0 || tlsz > 256) return HSRQ_EINVAL;
  ctx->tile = tlsz;
  return HSRQ_OK;
}

int hsrq_context_set_exact_norms(hsrq_context* ctx, int on) {
  if (!ctx) return HSRQ_EINVAL;
  ctx->exact = on ? 1 : 0;
  return HSRQ_OK;
}

int hsrq_context_set_packing(hsrq_context* ctx, int shifts_per_wave) {
  if (!ctx || (shifts_per_wave != 0 && shifts_per_wave != 1 && shifts_per_wave != 4))
    return HSRQ_EINVAL;
  ctx->pack = shifts_per_wave;
  return HSRQ_OK;
}

int hsrq_context_set_lookahead(hsrq_context* ctx, int on) {
  if (!ctx) return HSRQ_EINVAL;
  ctx->lookahead = on ? 1 : 0;
  return HSRQ_OK;
}

int hsrq_context_set_precision(hsrq_context* ctx, int prec) {
  if (!ctx) return HSRQ_EINVAL;
  if (prec == HSRQ_PREC_FP64) {
    ctx->precision = prec;
    return HSRQ_OK;
  }
  return prec == HSRQ_PREC_FP32 ? HSRQ_EUNSUPPORTED : HSRQ_EINVAL;
}

void* hsrq_context_stream(hsrq_context* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

const char* hsrq_strerror(int code) {
  switch (code) {
    case HSRQ_OK: return "success";
    case HSRQ_EINVAL: return "invalid argument";
    case HSRQ_EHIP: return hipGetErrorString(last_hip_error());
    case HSRQ_ENOMEM: return "device allocation failed";
    case HSRQ_EUNSUPPORTED: return "unsupported size or mode";
    case HSRQ_ESELFTEST: return "device self-test failed (MFMA operand layout)";
    case HSRQ_EBUSY: return "capacity exceeded";
    case HSRQ_ENOTFOUND: return "unknown request";
    default: return "unknown error";
  }
}

int hsrq_selftest(hsrq_context* ctx) {
  if (!ctx) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  int* d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(int)));
  int h = 0;
  hipError_t e = hipMemcpy(d, &h, sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_mfma_selftest(d, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) {
    last_hip_error() = e;
    return HSRQ_EHIP;
  }
  return h == 0 ? HSRQ_OK : HSRQ_ESELFTEST;
}

int hsrq_solve_device(hsrq_context* ctx, int n, const double* dH, int ldH, int m, double* dwr,
                      double* dvr, int ldvr, int tlsz, int* status) {
  if (!ctx) return HSRQ_EINVAL;
  hsrq_problem q = {n, dH, ldH, m, dwr, dvr, ldvr};
  if (validate_problem(q)) return HSRQ_EINVAL;
  if (m == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  return solve_with_restarts(ctx, 1, &q, tlsz, status);
}

int hsrq_solve_batch_device(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                            int* status) {
  if (!ctx || count < 0 || (count > 0 && !probs)) return HSRQ_EINVAL;
  if (count == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  return solve_with_restarts(ctx, count, probs, tlsz, status);
}

int hsrq_plan_create(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                     hsrq_plan** plan) {
  if (!plan) return HSRQ_EINVAL;
  return plan_build(ctx, count, probs, tlsz, -1, plan);
}

int hsrq_plan_execute(hsrq_plan* plan) {
  if (!plan) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(plan->ctx->device));
  return plan_enqueue(plan);
}

int hsrq_plan_execute_graph(hsrq_plan* plan) {
  if (!plan) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(plan->ctx->device));
  hipStream_t st = plan->ctx->stream;
  if (!plan->graph_exec) {
    HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = plan_enqueue(plan);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIP_TRY(e);
    plan->graph = g;
    HIP_TRY(hipGraphInstantiate(&plan->graph_exec, g, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(plan->graph_exec, st));
  return HSRQ_OK;
}

int hsrq_plan_execute_profiled(hsrq_plan* plan, double* kernel_ms, int* launches) {
  if (!plan || !kernel_ms || !launches) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(plan->ctx->device));
  return plan_enqueue_profiled(plan, kernel_ms, launches);
}

int hsrq_plan_sync(hsrq_plan* plan) {
  if (!plan) return HSRQ_EINVAL;
  HIP_TRY(hipStreamSynchronize(plan->ctx->stream));
  return HSRQ_OK;
}

int hsrq_plan_status(hsrq_plan* plan, int* status) {
  if (!plan || !status) return HSRQ_EINVAL;
  return plan_read_status(plan, status);
}

size_t hsrq_plan_workspace_bytes(const hsrq_plan* plan) { return plan ? plan->work_bytes : 0; }

void hsrq_plan_destroy(hsrq_plan* plan) { plan_free(plan); }

int hsrq_residuals_device(hsrq_context* ctx, int n, const double* dH, int ldH, int m,
                          const double* dwr, const double* dvr, int ldvr, double* res) {
  if (!ctx || n <= 0 || m < 0 || ldH < n || ldvr < n || !dH || (m > 0 && (!dwr || !dvr || !res)))
    return HSRQ_EINVAL;
  if (m == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  double* work = nullptr;
  double* dres = nullptr;
  HIP_TRY(hipMalloc(&work, ((size_t)n * m + n + 4) * sizeof(double)));
  hipError_t e = hipMalloc(&dres, (size_t)m * sizeof(double));
  if (e == hipSuccess)
    e = launch_residual(n, dH, ldH, m, dwr, dvr, ldvr, work, dres, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(res, dres, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(work);
  if (dres) (void)hipFree(dres);
  if (e != hipSuccess) {
    last_hip_error() = e;
    return HSRQ_EHIP;
  }
  return HSRQ_OK;
}

// Drop-in entry point (reference inverse_iteration.c:128-335): host
// buffers in and out, synchronous, on the current device.
void tiled_inverse_iteration(int n, const double* H, int ldH, int nrhs, double* wr, int tlsz,
                             int rhs_tlsz, double* vr, int ldvr) {
  if (n <= 0 || !H || ldH < n || nrhs < 0 || tlsz <= 0 || tlsz % 4 != 0 || rhs_tlsz <= 0 ||
      (nrhs > 0 && (!wr || !vr || ldvr < n))) {
    std::fprintf(stderr, "tiled_inverse_iteration: invalid argument\n");
    return;
  }
  if (nrhs == 0) return;
  static std::mutex mu;
  static hsrq_context* ctx = nullptr;
  std::lock_guard<std::mutex> lock(mu);
  int rc = HSRQ_OK;
  if (!ctx) rc = hsrq_context_create(&ctx, -1);
  double *dH = nullptr, *dwr = nullptr, *dvr = nullptr;
  if (!rc) rc = hipMalloc(&dH, (size_t)ldH * n * sizeof(double)) == hipSuccess ? 0 : HSRQ_ENOMEM;
  if (!rc) rc = hipMalloc(&dwr, (size_t)nrhs * sizeof(double)) == hipSuccess ? 0 : HSRQ_ENOMEM;
  if (!rc) rc = hipMalloc(&dvr, (size_t)n * nrhs * sizeof(double)) == hipSuccess ? 0 : HSRQ_ENOMEM;
  if (!rc && hipMemcpy(dH, H, (size_t)ldH * n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = HSRQ_EHIP;
  if (!rc && hipMemcpy(dwr, wr, (size_t)nrhs * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = HSRQ_EHIP;
  if (!rc) rc = hsrq_solve_device(ctx, n, dH, ldH, nrhs, dwr, dvr, n, tlsz, nullptr);
  if (!rc && hipMemcpy2D(vr, (size_t)ldvr * sizeof(double), dvr, (size_t)n * sizeof(double),
                         (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToHost) != hipSuccess)
    rc = HSRQ_EHIP;
  if (!rc && hipMemcpy(wr, dwr, (size_t)nrhs * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = HSRQ_EHIP;
  if (dH) (void)hipFree(dH);
  if (dwr) (void)hipFree(dwr);
  if (dvr) (void)hipFree(dvr);
  if (rc) std::fprintf(stderr, "tiled_inverse_iteration: %s\n", hsrq_strerror(rc));
}

}  // extern "C"
