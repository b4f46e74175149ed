// hsrq_api.cpp -- host runtime of libhsrq: contexts, plans (the device
// schedule of a set of problems), restarts, the device validator, the C-ABI
// (hsrq.h) and the drop-in tiled_inverse_iteration (reference
// inverse_iteration.h:19-22, inverse_iteration.c:128-335).
//
// A plan splits its problems into lanes.  A lane is a set of problems that
// share one tile size and one diagonal-sweep layout; its launch sequence
// runs on its own stream, so the latency-bound sweeps of one lane overlap
// the MFMA panel passes of another.  The lanes fork from and join into the
// context stream, so the whole plan captures into one HIP graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hsrq.h"
#include "../../include/inverse_iteration.h"
#include "hsrq_host.h"

using namespace hsrq;

namespace {

// Problems of one plan with the same tile and sweep layout, on one stream.
struct Lane {
  int b = 64;
  int pack = 1;
  int prec = HSRQ_PREC_FP64;     // panel-pass precision
  std::vector<ProbDev> probs;    // host copy of the descriptors
  ProbDev* d_probs = nullptr;
  void* d_items = nullptr;
  const GemmItem* norm_part = nullptr;
  int n_norm_part = 0;
  const RowItem* norm_fin = nullptr;
  int n_norm_fin = 0;
  const RowItem* init = nullptr;
  int n_init = 0;
  const ShiftItem* bt = nullptr;
  int n_bt = 0;
  int steps = 0;
  std::vector<const K1Item*> k1;
  std::vector<int> n_k1;
  std::vector<const GemmItem*> gb;   // panel pass of the rows above the next tile
  std::vector<int> n_gb;
  hipStream_t stream = nullptr;   // owned unless it is the context stream
  bool own_stream = false;
  hipStream_t side = nullptr;     // lookahead stream (context option)
  int side_blocks = 0;            // grid cap of the panel passes on `side`
  bool fuse = false;              // with `side`: one k_step launch per step instead
  int cus = 1;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr, ev_done = nullptr;
  std::vector<hipEvent_t> ev_p, ev_b;
  bool in_arena = false;          // d_probs, d_items carved from a caller's buffer
};

// Bump allocator over a buffer the caller keeps (plans built repeatedly on
// the same memory: the restart rounds).
struct DevArena {
  char* base = nullptr;
  size_t cap = 0, used = 0;
  void* take(size_t bytes) {
    const size_t b = align_up(std::max<size_t>(bytes, 1));
    if (used + b > cap) return nullptr;
    void* p = base + used;
    used += b;
    return p;
  }
};

void lane_free(Lane* L) {
  if (!L) return;
  for (auto e : L->ev_p) (void)hipEventDestroy(e);
  for (auto e : L->ev_b) (void)hipEventDestroy(e);
  if (L->ev_fork) (void)hipEventDestroy(L->ev_fork);
  if (L->ev_join) (void)hipEventDestroy(L->ev_join);
  if (L->ev_done) (void)hipEventDestroy(L->ev_done);
  if (L->side) (void)hipStreamDestroy(L->side);
  if (L->own_stream && L->stream) (void)hipStreamDestroy(L->stream);
  if (!L->in_arena) {
    if (L->d_probs) (void)hipFree(L->d_probs);
    if (L->d_items) (void)hipFree(L->d_items);
  }
  delete L;
}

}  // namespace

struct PieceRef {
  int lane, local, prob, k0;
};

struct hsrq_plan {
  hsrq_context* ctx = nullptr;
  int count = 0;
  std::vector<Lane*> lanes;
  std::vector<PieceRef> pieces;         // where each problem's shift ranges run
  std::vector<int> tile_of;             // plan-order problem -> tile
  std::vector<int> status_off;          // prefix of m in plan order
  int total_m = 0;
  void* d_work = nullptr;
  size_t work_bytes = 0;
  bool own_work = true;                 // false: d_work is in a caller's arena
  hipEvent_t ev_fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
};

namespace {

void plan_free(hsrq_plan* p) {
  if (!p) return;
  if (p->graph_exec) (void)hipGraphExecDestroy(p->graph_exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  for (Lane* L : p->lanes) lane_free(L);
  if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
  if (p->own_work && p->d_work) (void)hipFree(p->d_work);
  delete p;
}

// Workgroups per CU of the capped panel pass beside a diagonal sweep
// (lookahead_blocks_per_cu).  Measured, n = 8000, m = 800, tile 128 (MI355X,
// 128-VGPR sweep and panel pass): 7.9 ms without lookahead, 6.61 / 6.31 /
// 7.33 ms with 2 / 3 / 4 blocks per CU.

// Item tables of one lane, uploaded once.  `single`: the plan has one lane,
// which then may take a second stream for the panel pass of the rows above
// (with several lanes the sweeps of one lane already overlap the panel
// passes of another; the combination is not used).
int lane_build(hsrq_context* ctx, Lane* L, int total_lane_shifts, bool single, DevArena* ar) {
  L->pack = choose_pack(ctx->pack, L->b, total_lane_shifts);
  std::vector<GemmItem> part;
  std::vector<RowItem> fin, init;
  std::vector<ShiftItem> bt;
  int steps = 0;
  const int cnt = (int)L->probs.size();
  for (int i = 0; i < cnt; ++i) {
    const ProbDev& P = L->probs[i];
    if (P.m == 0) continue;
    steps = std::max(steps, P.T);
    add_norm_items(P, i, part, fin);
    add_init_items(P, i, init);
    add_bt_items(P, i, bt);
  }
  L->steps = steps;
  // Problem i runs its T_i steps from step off[i] on: step s works on its tile
  // column T_i - 1 - (s - off[i]).  The offsets level the work per step (all
  // problems starting together would crowd the first steps and leave the last
  // ones to the largest problems).
  std::vector<int> off = step_offsets(L->probs, steps);
  std::vector<StepItems> st(steps);
  for (int s = 0; s < steps; ++s)
    for (int i = 0; i < cnt; ++i)
      add_step_items(L->probs[i], i, L->probs[i].T - 1 - (s - off[i]), L->pack, st[s]);

  size_t nbytes = align_up(part.size() * sizeof(GemmItem)) + align_up(fin.size() * sizeof(RowItem)) +
                  align_up(init.size() * sizeof(RowItem)) + align_up(bt.size() * sizeof(ShiftItem));
  for (const auto& x : st)
    nbytes += align_up(x.k1.size() * sizeof(K1Item)) +
              align_up(x.gb.size() * sizeof(GemmItem));
  nbytes = std::max(nbytes, kAlign);
  std::vector<char> host(nbytes);
  if (ar) {
    L->in_arena = true;
    L->d_items = ar->take(nbytes);
    L->d_probs = (ProbDev*)ar->take(std::max<size_t>(cnt, 1) * sizeof(ProbDev));
    if (!L->d_items || !L->d_probs) return HSRQ_ENOMEM;
  } else {
    HIP_TRY(hipMalloc(&L->d_items, nbytes));
    HIP_TRY(hipMalloc(&L->d_probs, std::max<size_t>(cnt, 1) * sizeof(ProbDev)));
  }
  size_t pos = 0;
  auto put = [&](const void* src, size_t sz) -> const void* {
    if (sz) std::memcpy(host.data() + pos, src, sz);
    const void* d = (const char*)L->d_items + pos;
    pos += align_up(sz);
    return d;
  };
  L->norm_part = (const GemmItem*)put(part.data(), part.size() * sizeof(GemmItem));
  L->n_norm_part = (int)part.size();
  L->norm_fin = (const RowItem*)put(fin.data(), fin.size() * sizeof(RowItem));
  L->n_norm_fin = (int)fin.size();
  L->init = (const RowItem*)put(init.data(), init.size() * sizeof(RowItem));
  L->n_init = (int)init.size();
  L->bt = (const ShiftItem*)put(bt.data(), bt.size() * sizeof(ShiftItem));
  L->n_bt = (int)bt.size();
  for (const auto& x : st) {
    L->k1.push_back((const K1Item*)put(x.k1.data(), x.k1.size() * sizeof(K1Item)));
    L->n_k1.push_back((int)x.k1.size());
    L->gb.push_back((const GemmItem*)put(x.gb.data(), x.gb.size() * sizeof(GemmItem)));
    L->n_gb.push_back((int)x.gb.size());
  }
  HIP_TRY(hipMemcpy(L->d_items, host.data(), nbytes, hipMemcpyHostToDevice));
  if (cnt > 0)
    HIP_TRY(hipMemcpy(L->d_probs, L->probs.data(), cnt * sizeof(ProbDev), hipMemcpyHostToDevice));
  if (L->own_stream) HIP_TRY(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
  const int la = !single             ? 0
                 : ctx->lookahead < 0 ? lookahead_blocks_per_cu(L->pack, L->b, total_lane_shifts,
                                                                ctx->cus)
                                      : ctx->lookahead;
  if (la > 0) {   // 0: off (asked for, or the automatic choice for packed sweeps)
    HIP_TRY(hipStreamCreateWithFlags(&L->side, hipStreamNonBlocking));
    L->side_blocks = la * std::max(ctx->cus, 1);
    L->fuse = L->prec == HSRQ_PREC_FP64 && step_fusable(L->pack, L->b);
    L->cus = std::max(ctx->cus, 1);
  }
  HIP_TRY(hipEventCreateWithFlags(&L->ev_fork, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&L->ev_join, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&L->ev_done, hipEventDisableTiming));
  L->ev_p.assign(steps, nullptr);
  L->ev_b.assign(steps, nullptr);
  for (int s = 0; s < steps; ++s) {
    HIP_TRY(hipEventCreateWithFlags(&L->ev_p[s], hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&L->ev_b[s], hipEventDisableTiming));
  }
  return HSRQ_OK;
}

// Builds descriptors, one workspace allocation and the lanes.  its: start
// vector (-1: eps3 * 1; >= 0: the its-th restart vector; -2: given in vr),
// or per shift in its_k_of[problem] (device arrays; null entries: its).  ar:
// build workspace and item tables in this arena instead of allocating them.
int plan_build(hsrq_context* ctx, int count, const hsrq_problem* in, int tlsz, int its,
               hsrq_plan** out, const int* const* its_k_of = nullptr, DevArena* ar = nullptr) {
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
  p->count = count;
  p->status_off.assign(count + 1, 0);
  for (int i = 0; i < count; ++i) p->status_off[i + 1] = p->status_off[i] + in[i].m;
  p->total_m = p->status_off[count];

  std::vector<std::vector<Piece>> groups;
  std::vector<int> tiles;
  assign_lanes(ctx, count, in, tlsz, groups, tiles);
  p->tile_of.assign(count, 64);
  size_t bytes = 0;
  std::vector<std::vector<size_t>> ws_off(groups.size());
  for (size_t l = 0; l < groups.size(); ++l)
    for (size_t j = 0; j < groups[l].size(); ++j) {
      const Piece& pc = groups[l][j];
      p->pieces.push_back({(int)l, (int)j, pc.prob, pc.k0});
      p->tile_of[pc.prob] = tiles[l];
      ws_off[l].push_back(bytes);
      bytes += prob_workspace_bytes(in[pc.prob].n, pc.m, tiles[l],
                                    ctx->precision != HSRQ_PREC_FP64);
    }
  p->work_bytes = bytes;
  hipError_t e = hipSuccess;
  if (ar) {
    p->own_work = false;
    p->d_work = ar->take(std::max(bytes, kAlign));
    if (!p->d_work) e = hipErrorOutOfMemory;
  } else {
    e = hipMalloc(&p->d_work, std::max(bytes, kAlign));
  }
  if (e != hipSuccess) {
    last_hip_error() = e;
    plan_free(p);
    return HSRQ_ENOMEM;
  }
  int rc = HSRQ_OK;
  for (size_t l = 0; l < groups.size() && !rc; ++l) {
    Lane* L = new (std::nothrow) Lane();
    if (!L) {
      rc = HSRQ_ENOMEM;
      break;
    }
    p->lanes.push_back(L);
    L->b = tiles[l];
    L->prec = ctx->precision;
    L->own_stream = l > 0;
    L->stream = l == 0 ? ctx->stream : nullptr;
    int shifts = 0;
    for (size_t j = 0; j < groups[l].size(); ++j) {
      const Piece& pc = groups[l][j];
      hsrq_problem q = in[pc.prob];
      q.m = pc.m;
      q.wr = q.wr + pc.k0;
      q.vr = q.vr + (size_t)pc.k0 * q.ldvr;
      ProbDev P;
      const int* itk = its_k_of && its_k_of[pc.prob] ? its_k_of[pc.prob] + pc.k0 : nullptr;
      prob_describe(P, q, L->b, its, ctx->exact, (char*)p->d_work + ws_off[l][j], itk,
                    L->prec != HSRQ_PREC_FP64);
      L->probs.push_back(P);
      shifts += pc.m;
    }
    rc = lane_build(ctx, L, shifts, groups.size() == 1, ar);
  }
  if (!rc) {
    e = hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming);
    if (e != hipSuccess) {
      last_hip_error() = e;
      rc = HSRQ_EHIP;
    }
  }
  if (rc) {
    plan_free(p);
    return rc;
  }
  *out = p;
  return HSRQ_OK;
}

// Per-launch timing (kinds: 0 norms, 1 init, 2 diag, 3 update,
// 4 back-transform; 5 reserved).
struct Prof {
  std::vector<hipEvent_t> ev;
  std::vector<int> kind;
};

// Launch sequence of one lane on its stream.  Per step: the diagonal sweeps
// (each starting with the panel pass of its own tile's rows), then the panel
// pass of the rows above the next diagonal tile (on the lookahead stream when
// enabled, where it overlaps the next sweep: its grid is capped so that every
// CU keeps room for a sweep workgroup, and the sweeps run at raised wave
// priority).  Both kernels derive their scale exponents themselves.
int lane_enqueue(Lane* L, Prof* prof) {
  hipStream_t st = L->stream;
  hipStream_t sd = L->side ? L->side : st;
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
  int rc = run(0, st, [&](hipStream_t s) {
    return launch_norms(L->d_probs, L->norm_part, L->n_norm_part, L->norm_fin, L->n_norm_fin, s);
  });
  if (rc) return rc;
  rc = run(1, st, [&](hipStream_t s) { return launch_init(L->d_probs, L->init, L->n_init, s); });
  if (rc) return rc;
  if (L->prec != HSRQ_PREC_FP64)   // equilibration exponents of the FP32 panel pass
    for (size_t i = 0; i < L->probs.size(); ++i)
      if (L->probs[i].m > 0)
        HIP_TRY(launch_equilibrate(L->d_probs + i, L->probs[i].n, L->probs[i].nrb, st));
  if (L->fuse && !prof) {
    // Launch s: the sweeps of step s beside the panel pass of step s - 1, on
    // the lane's stream.  The sweep of step s starts from its tile's rows as
    // the pass of step s - 2 left them and overwrites the coefficient
    // buffers that pass read; the pass of step s - 1 reads what the sweep of
    // step s - 1 wrote: both are in launch s - 1.  (The profiled form below
    // keeps the two kernels apart, to time them.)
    for (int s = 0; s <= L->steps; ++s) {
      const bool sw = s < L->steps, pa = s >= 1 && L->n_gb[s - 1] > 0;
      if (!sw && !pa) continue;
      // the pass takes every workgroup slot the sweeps leave (4 per CU)
      const int nk1 = sw ? L->n_k1[s] : 0;
      const int cap = std::max(L->side_blocks, 4 * L->cus - (nk1 + 7) / 8 * 8);
      HIP_TRY(launch_step(L->b, L->d_probs, sw ? L->k1[s] : nullptr, nk1,
                          pa ? L->gb[s - 1] : nullptr, pa ? L->n_gb[s - 1] : 0, cap, st));
    }
    return run(4, st, [&](hipStream_t s) { return launch_backtransform(L->d_probs, L->bt, L->n_bt, s); });
  }
  if (L->side) {
    HIP_TRY(hipEventRecord(L->ev_fork, st));
    HIP_TRY(hipStreamWaitEvent(sd, L->ev_fork, 0));
  }
  auto upd = L->prec == HSRQ_PREC_FP64 ? launch_update : launch_update32;
  for (int s = 0; s < L->steps; ++s) {
    // The sweep of step s starts from its tile's rows as the panel pass of
    // step s-2 left them, and overwrites the coefficient buffers that pass
    // read; the panel pass of step s-1 (rows above the tile) may still run.
    if (L->side && s >= 2 && L->n_gb[s - 2] > 0)
      HIP_TRY(hipStreamWaitEvent(st, L->ev_b[s - 2], 0));
    rc = run(2, st, [&](hipStream_t x) {
      return launch_sweep(L->pack, L->b, L->d_probs, L->k1[s], L->n_k1[s], x);
    });
    if (rc) return rc;
    if (L->n_gb[s] == 0) continue;
    if (L->side) {   // rows above: beside the next sweep
      HIP_TRY(hipEventRecord(L->ev_p[s], st));
      HIP_TRY(hipStreamWaitEvent(sd, L->ev_p[s], 0));
    }
    rc = run(3, sd, [&](hipStream_t x) {
      return upd(L->d_probs, L->gb[s], L->n_gb[s], x, L->side_blocks);
    });
    if (rc) return rc;
    if (L->side) HIP_TRY(hipEventRecord(L->ev_b[s], sd));
  }
  if (L->side) {   // join (also without work: graph capture needs it)
    HIP_TRY(hipEventRecord(L->ev_join, sd));
    HIP_TRY(hipStreamWaitEvent(st, L->ev_join, 0));
  }
  return run(4, st, [&](hipStream_t s) { return launch_backtransform(L->d_probs, L->bt, L->n_bt, s); });
}

int plan_enqueue_impl(hsrq_plan* p, Prof* prof) {
  hipStream_t main = p->ctx->stream;
  HIP_TRY(hipEventRecord(p->ev_fork, main));
  for (Lane* L : p->lanes)
    if (L->stream != main) HIP_TRY(hipStreamWaitEvent(L->stream, p->ev_fork, 0));
  for (Lane* L : p->lanes) {
    const int rc = lane_enqueue(L, prof);
    if (rc) return rc;
  }
  for (Lane* L : p->lanes)
    if (L->stream != main) {
      HIP_TRY(hipEventRecord(L->ev_done, L->stream));
      HIP_TRY(hipStreamWaitEvent(main, L->ev_done, 0));
    }
  return HSRQ_OK;
}

int plan_enqueue(hsrq_plan* p) { return plan_enqueue_impl(p, nullptr); }

int plan_enqueue_profiled(hsrq_plan* p, double* kernel_ms, int* launches) {
  Prof prof;
  int rc = plan_enqueue_impl(p, &prof);
  if (!rc) {
    HIP_TRY(hipStreamSynchronize(p->ctx->stream));
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
  for (const PieceRef& r : p->pieces) {
    const ProbDev& P = p->lanes[r.lane]->probs[r.local];
    if (P.m == 0) continue;
    HIP_TRY(hipMemcpyAsync(status + p->status_off[r.prob] + r.k0, P.status, P.m * sizeof(int),
                           hipMemcpyDeviceToHost, p->ctx->stream));
  }
  HIP_TRY(hipStreamSynchronize(p->ctx->stream));
  return HSRQ_OK;
}

// Restarts of the shifts a first solve left unconverged (dlaein semantics,
// the reference's intent at inverse_iteration.c:160-328, quirks Q2/Q3):
// round its = 0, 1, ... solves the still unconverged shifts again from the
// its-th restart vector, at most n rounds; a shift keeps the vector of the
// round it converged in (or of round n-1), and each round moves the shifts
// that converged in it behind the ones before, stable.
//
// The rounds of one shift are independent of each other given H and the
// shift, so they are evaluated speculatively: a round of this engine solves,
// for every pending shift of every problem, the next K start vectors at once
// (the shift repeated K times with per-shift start vectors, one plan over all
// problems), and each shift takes the first of them that converges.  K is
// the budget of solves per round over the pending shifts, so few pending
// shifts advance many start vectors per round.  Results, order and status
// are those of the one-round-at-a-time loop (each shift's arithmetic does not
// depend on what shares its launch).  Memory: one scratch allocation for the
// whole restart phase and one arena for the plans, grown when a round needs
// more; no allocation per round otherwise.
int restart_rounds(hsrq_context* ctx, int count, const hsrq_problem* probs,
                   const std::vector<int>& tile_of, const std::vector<int>& off,
                   std::vector<int>& st) {
  struct Tail {
    int prob = 0, first = 0, cnt = 0, its0 = 0;
    std::vector<double> wr;          // the tail's shifts (tail order)
    std::vector<int> pend;           // tail indices still unresolved
    std::vector<int> its_pass;       // round of convergence (INT_MAX: never)
    std::vector<int> stat;           // status of the kept solve
    double* dbuf = nullptr;          // kept vectors, n x cnt (scratch)
    double* dwr = nullptr;           // this round's shifts (scratch)
    int* dits = nullptr;             // this round's start vectors (scratch)
    double* dy = nullptr;            // this round's vectors, n x mr (scratch)
    int K = 0, mr = 0;
  };
  std::vector<Tail> tails;
  hipStream_t stream = ctx->stream;
  // stable partition converged-first of each problem (converged columns move
  // forward in place; the tail columns are rewritten at the end)
  for (int i = 0; i < count; ++i) {
    const hsrq_problem& q = probs[i];
    int* sti = st.data() + off[i];
    int nc = 0;
    for (int k = 0; k < q.m; ++k) nc += (sti[k] & HSRQ_CONVERGED) ? 1 : 0;
    if (nc == q.m) continue;
    std::vector<double> wr_h(q.m);
    HIP_TRY(hipMemcpy(wr_h.data(), q.wr, q.m * sizeof(double), hipMemcpyDeviceToHost));
    Tail t;
    t.prob = i;
    t.first = nc;
    t.cnt = q.m - nc;
    std::vector<int> st_new(q.m);
    std::vector<double> wr_new(q.m);
    int pc = 0;
    for (int k = 0; k < q.m; ++k)
      if (sti[k] & HSRQ_CONVERGED) {
        if (pc != k)
          HIP_TRY(hipMemcpyAsync(q.vr + (size_t)pc * q.ldvr, q.vr + (size_t)k * q.ldvr,
                                 q.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
        st_new[pc] = sti[k];
        wr_new[pc++] = wr_h[k];
      }
    for (int k = 0; k < q.m; ++k)
      if (!(sti[k] & HSRQ_CONVERGED)) t.wr.push_back(wr_h[k]);
    std::copy(st_new.begin(), st_new.begin() + nc, sti);
    std::copy(wr_new.begin(), wr_new.begin() + nc, wr_h.begin());
    std::copy(t.wr.begin(), t.wr.end(), wr_h.begin() + nc);
    HIP_TRY(hipMemcpyAsync(q.wr, wr_h.data(), q.m * sizeof(double), hipMemcpyHostToDevice,
                           stream));
    HIP_TRY(hipStreamSynchronize(stream));
    t.its_pass.assign(t.cnt, INT_MAX);
    t.stat.assign(t.cnt, 0);
    for (int j = 0; j < t.cnt; ++j) t.pend.push_back(j);
    tails.push_back(std::move(t));
  }
  if (tails.empty()) return HSRQ_OK;

  // solves per round: about 2 GiB of workspace and vectors (~40 n bytes per
  // solve), 1024 .. 16384
  int nmax = 0;
  size_t pend_total = 0;
  for (const Tail& t : tails) {
    nmax = std::max(nmax, probs[t.prob].n);
    pend_total += t.cnt;
  }
  const size_t budget = std::min<size_t>(16384, std::max<size_t>(1024, ((size_t)2 << 30) / (40 * (size_t)nmax)));
  // scratch: kept vectors, and the round buffers sized for `budget` solves
  size_t sc = 0;
  for (const Tail& t : tails) sc += align_up((size_t)probs[t.prob].n * t.cnt * sizeof(double));
  const size_t round_bytes = std::max(budget, pend_total) * (size_t)nmax * sizeof(double) +
                             tails.size() * 2 * kAlign +
                             std::max(budget, pend_total) * (sizeof(double) + sizeof(int));
  char* scratch = nullptr;
  HIP_TRY(hipMalloc(&scratch, sc + align_up(round_bytes) + 2 * align_up(pend_total * 8) + kAlign));
  DevArena ar;
  size_t pos = 0;
  for (Tail& t : tails) {
    t.dbuf = (double*)(scratch + pos);
    pos += align_up((size_t)probs[t.prob].n * t.cnt * sizeof(double));
  }
  char* rb = scratch + pos;
  const size_t rb_cap = align_up(round_bytes) + 2 * align_up(pend_total * 8) + kAlign;
  // one round; returns 1 when nothing is pending
  auto round = [&]() -> int {
    size_t npend = 0;
    for (const Tail& t : tails) npend += t.pend.size();
    if (npend == 0) return 1;
    const int K = (int)std::max<size_t>(1, budget / npend);
    size_t total = 0;
    for (Tail& t : tails) {
      t.K = t.pend.empty() ? 0 : std::min(K, probs[t.prob].n - t.its0);
      t.mr = t.K * (int)t.pend.size();
      total += t.mr;
    }
    // per tail K copies of its pending shifts, start vectors its0 .. its0+K-1
    std::vector<double> h_wr(total);
    std::vector<int> h_its(total);
    double* dwr_all = (double*)rb;
    int* dits_all = (int*)(rb + align_up(total * sizeof(double)));
    char* pw = rb + align_up(total * sizeof(double)) + align_up(total * sizeof(int));
    std::vector<hsrq_problem> vq;
    std::vector<const int*> itsk;
    std::vector<int> vtile;
    size_t p2 = 0;
    for (Tail& t : tails) {
      if (t.mr == 0) continue;
      const hsrq_problem& q = probs[t.prob];
      t.dwr = dwr_all + p2;
      t.dits = dits_all + p2;
      t.dy = (double*)pw;
      pw += align_up((size_t)q.n * t.mr * sizeof(double));
      const size_t np = t.pend.size();
      for (int k = 0; k < t.K; ++k)
        for (size_t a = 0; a < np; ++a) {
          h_wr[p2 + k * np + a] = t.wr[t.pend[a]];
          h_its[p2 + k * np + a] = t.its0 + k;
        }
      p2 += t.mr;
      vq.push_back({q.n, q.H, q.ldH, t.mr, t.dwr, t.dy, q.n});
      itsk.push_back(t.dits);
      vtile.push_back(tile_of[t.prob]);
    }
    if ((size_t)(pw - rb) > rb_cap) return HSRQ_ENOMEM;   // (the round fits by construction)
    // a plan's problems share one tile (assign_lanes), so the rounds' do too
    const int tl = vtile[0];
    for (int x : vtile)
      if (x != tl) return HSRQ_EUNSUPPORTED;
    HIP_TRY(hipMemcpyAsync(dwr_all, h_wr.data(), total * sizeof(double), hipMemcpyHostToDevice,
                           stream));
    HIP_TRY(hipMemcpyAsync(dits_all, h_its.data(), total * sizeof(int), hipMemcpyHostToDevice,
                           stream));
    hsrq_plan* pl = nullptr;
    int rc2;
    for (;;) {
      ar.used = 0;
      rc2 = ar.base ? plan_build(ctx, (int)vq.size(), vq.data(), tl, -1, &pl, itsk.data(), &ar)
                    : HSRQ_ENOMEM;
      if (rc2 != HSRQ_ENOMEM) break;
      // grow the arena, build again
      const size_t want = std::max<size_t>(2 * ar.cap, (size_t)64 << 20);
      if (ar.base) (void)hipFree(ar.base);
      ar.base = nullptr;
      ar.cap = 0;
      HIP_TRY(hipMalloc(&ar.base, want));
      ar.cap = want;
    }
    if (rc2) return rc2;
    rc2 = plan_enqueue(pl);
    std::vector<int> sv(std::max<size_t>(total, 1));
    if (!rc2) rc2 = plan_read_status(pl, sv.data());
    plan_free(pl);
    if (rc2) return rc2;
    // each pending shift keeps the first of its K solves that converged, or
    // the solve of round n-1
    size_t base = 0;
    for (Tail& t : tails) {
      if (t.mr == 0) continue;
      const hsrq_problem& q = probs[t.prob];
      const int np = (int)t.pend.size();
      const bool last = t.its0 + t.K >= q.n;
      std::vector<int> keep;
      for (int a = 0; a < np; ++a) {
        const int j = t.pend[a];
        int got = -1;
        for (int k = 0; k < t.K && got < 0; ++k)
          if (sv[base + (size_t)k * np + a] & HSRQ_CONVERGED) got = k;
        if (got < 0 && !last) {
          keep.push_back(j);
          continue;
        }
        const int k = got >= 0 ? got : t.K - 1;
        t.its_pass[j] = got >= 0 ? t.its0 + k : INT_MAX;
        t.stat[j] = sv[base + (size_t)k * np + a];
        HIP_TRY(hipMemcpyAsync(t.dbuf + (size_t)j * q.n, t.dy + ((size_t)k * np + a) * q.n,
                               q.n * sizeof(double), hipMemcpyDeviceToDevice, stream));
      }
      t.pend = keep;
      t.its0 += t.K;
      base += t.mr;
    }
    HIP_TRY(hipStreamSynchronize(stream));
    return 0;
  };
  int rc = HSRQ_OK;
  for (;;) {
    const int r = round();
    if (r == 1) break;
    if (r) {
      rc = r;
      break;
    }
  }
  // final order of each tail: by the round of convergence, stable
  for (Tail& t : tails) {
    if (rc) break;
    const hsrq_problem& q = probs[t.prob];
    std::vector<int> order(t.cnt);
    for (int j = 0; j < t.cnt; ++j) order[j] = j;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return t.its_pass[a] < t.its_pass[b]; });
    std::vector<double> wr_t(t.cnt);
    int* sti = st.data() + off[t.prob];
    for (int r = 0; r < t.cnt && !rc; ++r) {
      const int j = order[r];
      wr_t[r] = t.wr[j];
      sti[t.first + r] = t.stat[j];
      if (hipMemcpyAsync(q.vr + (size_t)(t.first + r) * q.ldvr, t.dbuf + (size_t)j * q.n,
                         q.n * sizeof(double), hipMemcpyDeviceToDevice, stream) != hipSuccess)
        rc = HSRQ_EHIP;
    }
    if (!rc && hipMemcpyAsync(q.wr + t.first, wr_t.data(), t.cnt * sizeof(double),
                              hipMemcpyHostToDevice, stream) != hipSuccess)
      rc = HSRQ_EHIP;
    if (!rc && hipStreamSynchronize(stream) != hipSuccess) rc = HSRQ_EHIP;
  }
  (void)hipStreamSynchronize(stream);
  if (ar.base) (void)hipFree(ar.base);
  (void)hipFree(scratch);
  return rc;
}

// Solves a set of problems: one plan for all of them, then the restarts of
// the shifts that did not converge (restart_rounds).
int solve_with_restarts(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                        int* status_out) {
  hsrq_plan* p = nullptr;
  int rc = plan_build(ctx, count, probs, tlsz, -1, &p);
  if (rc) return rc;
  rc = plan_enqueue(p);
  std::vector<int> st(std::max(p->total_m, 1));
  if (!rc) rc = plan_read_status(p, st.data());
  std::vector<int> off(p->status_off);
  std::vector<int> tile_of(p->tile_of);
  plan_free(p);
  if (rc) return rc;
  rc = restart_rounds(ctx, count, probs, tile_of, off, st);
  if (rc) return rc;
  if (status_out) std::copy(st.begin(), st.begin() + off[count], status_out);
  return HSRQ_OK;
}

// Reduced-precision solve (HSRQ_PREC_FP32): one pass with the FP32 panel
// pass for every shift, then the FP64 residual of every vector; the shifts
// whose vector did not converge or whose residual exceeds 10 n eps are solved
// again on the FP64 path (with its restarts) and scattered back.  The final
// order is the FP64 path's: converged first (stable), unchanged when every
// shift converged.  HSRQ_PREC_FP32_RAW: the FP32 pass alone, no restarts.
int solve_reduced(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                  int* status_out) {
  hsrq_plan* p = nullptr;
  int rc = plan_build(ctx, count, probs, tlsz, -1, &p);
  if (rc) return rc;
  rc = plan_enqueue(p);
  std::vector<int> st(std::max(p->total_m, 1));
  if (!rc) rc = plan_read_status(p, st.data());
  std::vector<int> off(p->status_off);
  std::vector<int> tile_of(p->tile_of);
  plan_free(p);
  if (rc) return rc;
  if (ctx->precision == HSRQ_PREC_FP32) {
    hipStream_t stream = ctx->stream;
    for (int i = 0; i < count; ++i) {
      const hsrq_problem& q = probs[i];
      const int n = q.n, m = q.m;
      if (m == 0) continue;
      int* sti = st.data() + off[i];
      // FP64 residuals of the FP32 vectors
      std::vector<double> res(m);
      rc = residuals_run(ctx, 1, &q, res.data());
      if (rc) return rc;
      hipError_t e = hipSuccess;
      std::vector<int> bad;
      for (int k = 0; k < m; ++k)
        if (!(sti[k] & HSRQ_CONVERGED) || !(res[k] <= 10.0 * n)) bad.push_back(k);
      if (bad.empty()) continue;
      // gather, solve on the FP64 path, scatter back
      const int nb = (int)bad.size();
      std::vector<double> wr_h(m);
      HIP_TRY(hipMemcpy(wr_h.data(), q.wr, m * sizeof(double), hipMemcpyDeviceToHost));
      double *twr = nullptr, *tvr = nullptr;
      HIP_TRY(hipMalloc(&twr, (size_t)nb * sizeof(double)));
      e = hipMalloc(&tvr, (size_t)n * nb * sizeof(double));
      std::vector<double> wsub(nb);
      for (int j = 0; j < nb; ++j) wsub[j] = wr_h[bad[j]];
      if (e == hipSuccess) e = hipMemcpy(twr, wsub.data(), nb * sizeof(double), hipMemcpyHostToDevice);
      std::vector<int> stb(nb, 0);
      if (e == hipSuccess) {
        const int prec = ctx->precision;
        ctx->precision = HSRQ_PREC_FP64;
        hsrq_problem sub = {n, q.H, q.ldH, nb, twr, tvr, n};
        rc = solve_with_restarts(ctx, 1, &sub, tile_of[i], stb.data());
        ctx->precision = prec;
      } else {
        last_hip_error() = e;
        rc = HSRQ_EHIP;
      }
      if (!rc) {
        e = hipMemcpy(wsub.data(), twr, nb * sizeof(double), hipMemcpyDeviceToHost);
        for (int j = 0; j < nb && e == hipSuccess; ++j) {
          e = hipMemcpyAsync(q.vr + (size_t)bad[j] * q.ldvr, tvr + (size_t)j * n,
                             n * sizeof(double), hipMemcpyDeviceToDevice, stream);
          wr_h[bad[j]] = wsub[j];
          sti[bad[j]] = stb[j];
        }
        if (e == hipSuccess)
          e = hipMemcpyAsync(q.wr, wr_h.data(), m * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) {
          last_hip_error() = e;
          rc = HSRQ_EHIP;
        }
      }
      (void)hipFree(twr);
      if (tvr) (void)hipFree(tvr);
      if (rc) return rc;
      // converged first, stable (a no-op when every shift converged)
      std::vector<int> order;
      for (int k = 0; k < m; ++k)
        if (sti[k] & HSRQ_CONVERGED) order.push_back(k);
      if ((int)order.size() < m) {
        for (int k = 0; k < m; ++k)
          if (!(sti[k] & HSRQ_CONVERGED)) order.push_back(k);
        double* tmp = nullptr;
        HIP_TRY(hipMalloc(&tmp, (size_t)n * m * sizeof(double)));
        std::vector<double> wr2(m);
        std::vector<int> st2(m);
        e = hipSuccess;
        for (int t = 0; t < m && e == hipSuccess; ++t) {
          wr2[t] = wr_h[order[t]];
          st2[t] = sti[order[t]];
          e = hipMemcpyAsync(tmp + (size_t)t * n, q.vr + (size_t)order[t] * q.ldvr,
                             n * sizeof(double), hipMemcpyDeviceToDevice, stream);
        }
        if (e == hipSuccess)
          e = hipMemcpy2DAsync(q.vr, q.ldvr * sizeof(double), tmp, n * sizeof(double),
                               n * sizeof(double), m, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess)
          e = hipMemcpyAsync(q.wr, wr2.data(), m * sizeof(double), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) {
          last_hip_error() = e;
          return HSRQ_EHIP;
        }
        std::copy(st2.begin(), st2.end(), sti);
      }
    }
  }
  if (status_out) std::copy(st.begin(), st.begin() + off[count], status_out);
  return HSRQ_OK;
}

int solve_any(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz, int* status) {
  if (ctx->precision == HSRQ_PREC_FP64)
    return solve_with_restarts(ctx, count, probs, tlsz, status);
  return solve_reduced(ctx, count, probs, tlsz, status);
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
  hipError_t e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    last_hip_error() = e;
    delete c;
    return HSRQ_EHIP;
  }
  noalloc_setup(dev);
  *ctx = c;
  return HSRQ_OK;
}

void hsrq_context_destroy(hsrq_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  release_post(ctx);
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

int hsrq_context_set_lanes(hsrq_context* ctx, int lanes) {
  if (!ctx || lanes < 0 || lanes > 8) return HSRQ_EINVAL;
  ctx->lanes = lanes;
  return HSRQ_OK;
}

int hsrq_context_set_lookahead(hsrq_context* ctx, int blocks_per_cu) {
  if (!ctx || blocks_per_cu < -1 || blocks_per_cu > 8) return HSRQ_EINVAL;
  ctx->lookahead = blocks_per_cu;
  return HSRQ_OK;
}

int hsrq_context_set_precision(hsrq_context* ctx, int prec) {
  if (!ctx || (prec != HSRQ_PREC_FP64 && prec != HSRQ_PREC_FP32 && prec != HSRQ_PREC_FP32_RAW))
    return HSRQ_EINVAL;
  ctx->precision = prec;
  return HSRQ_OK;
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
    case HSRQ_ERANGE: return "result outside the floating-point range";
    case HSRQ_ESCRATCH:
      return "problem exceeds the allocation-free scratch (HSRQ_NOALLOC_SCRATCH_BYTES, "
             "HSRQ_NOALLOC_MAX_SHIFTS)";
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
  return solve_any(ctx, 1, &q, tlsz, status);
}

int hsrq_solve_batch_device(hsrq_context* ctx, int count, const hsrq_problem* probs, int tlsz,
                            int* status) {
  if (!ctx || count < 0 || (count > 0 && !probs)) return HSRQ_EINVAL;
  if (count == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  return solve_any(ctx, count, probs, tlsz, status);
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

int hsrq_plan_lanes(const hsrq_plan* plan, int max_lanes, int* tile, int* packing, int* problems) {
  if (!plan || max_lanes < 0) return HSRQ_EINVAL;
  const int nl = (int)plan->lanes.size();
  for (int l = 0; l < std::min(nl, max_lanes); ++l) {
    if (tile) tile[l] = plan->lanes[l]->b;
    if (packing) packing[l] = plan->lanes[l]->pack;
    if (problems) problems[l] = (int)plan->lanes[l]->probs.size();
  }
  return nl;
}

void hsrq_plan_destroy(hsrq_plan* plan) { plan_free(plan); }

int hsrq_autotune(hsrq_context* ctx, int count, const hsrq_problem* probs, int* tile,
                  int* packing, int* lanes, double* ms) {
  if (!ctx || count <= 0 || !probs) return HSRQ_EINVAL;
  for (int i = 0; i < count; ++i)
    if (validate_problem(probs[i])) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  const int t0 = ctx->tile, p0 = ctx->pack, l0 = ctx->lanes;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(hipEventCreate(&e0));
  hipError_t he = hipEventCreate(&e1);
  if (he != hipSuccess) {
    (void)hipEventDestroy(e0);
    last_hip_error() = he;
    return HSRQ_EHIP;
  }
  double best = INFINITY;
  int bt = 0, bp = 0, bl = 0, rc = HSRQ_OK;
  static const int kTiles[] = {64, 128, 256};
  static const int kPacks[] = {1, 4};
  static const int kLanes[] = {1, 2, 3, 4};
  for (int t : kTiles)
    for (int pk : kPacks)
      for (int ln : kLanes) {
        if (rc) break;
        if (choose_pack(pk, t, 1 << 30) != pk) continue;   // layout unsupported at this tile
        ctx->tile = t;
        ctx->pack = pk;
        ctx->lanes = ln;
        hsrq_plan* p = nullptr;
        rc = plan_build(ctx, count, probs, 0, -1, &p);
        if (rc == HSRQ_ENOMEM) {   // this candidate does not fit: skip it
          rc = HSRQ_OK;
          continue;
        }
        // warm (captures the graph), then one timed replay
        if (!rc) rc = hsrq_plan_execute_graph(p);
        if (!rc && (he = hipEventRecord(e0, ctx->stream)) != hipSuccess) rc = HSRQ_EHIP;
        if (!rc) rc = hsrq_plan_execute_graph(p);
        if (!rc && (he = hipEventRecord(e1, ctx->stream)) != hipSuccess) rc = HSRQ_EHIP;
        if (!rc && (he = hipEventSynchronize(e1)) != hipSuccess) rc = HSRQ_EHIP;
        float el = 0.f;
        if (!rc && (he = hipEventElapsedTime(&el, e0, e1)) != hipSuccess) rc = HSRQ_EHIP;
        if (rc == HSRQ_EHIP) last_hip_error() = he;
        plan_free(p);
        if (!rc && el < best) {
          best = el;
          bt = t;
          bp = pk;
          bl = ln;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc || !(best < INFINITY)) {
    ctx->tile = t0;
    ctx->pack = p0;
    ctx->lanes = l0;
    return rc ? rc : HSRQ_ENOMEM;
  }
  ctx->tile = bt;
  ctx->pack = bp;
  ctx->lanes = bl;
  if (tile) *tile = bt;
  if (packing) *packing = bp;
  if (lanes) *lanes = bl;
  if (ms) *ms = best;
  return HSRQ_OK;
}

// Drop-in entry point (reference inverse_iteration.c:128-335): host
// buffers in and out, synchronous, on the current device.  The device copies
// of H, wr and vr are kept between calls (grown when a call needs more), so a
// repeated call pays for the copies and the solve only.  H goes through two
// pinned staging buffers kept with them: the host copies the Hessenberg part
// of a chunk of columns (rows 0 .. j+1 of column j) into one while the other
// is in flight to the device, so the host reads half of H and the DMA runs
// from pinned memory (n = 8000: about 10 ms for H instead of ~65 ms through
// pitched pageable copies).  The kernels read nothing below the subdiagonal,
// where the device copy holds whatever the staging held.  Nothing past
// element (n-1) ldH + n-1 of the caller's H is read, as in the reference.
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
  static double *dH = nullptr, *dwr = nullptr, *dvr = nullptr;
  static size_t capH = 0, capw = 0, capv = 0;
  constexpr size_t kStageBytes = (size_t)32 << 20;
  static char* stage[2] = {nullptr, nullptr};
  static hipEvent_t staged[2] = {nullptr, nullptr};
  std::lock_guard<std::mutex> lock(mu);
  int rc = HSRQ_OK;
  if (!ctx) rc = hsrq_context_create(&ctx, -1);
  auto grow = [&](double*& p, size_t& cap, size_t need) {
    if (rc || need <= cap) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, need * sizeof(double)) != hipSuccess) {
      p = nullptr;
      rc = HSRQ_ENOMEM;
      return;
    }
    cap = need;
  };
  grow(dH, capH, (size_t)n * n);
  grow(dwr, capw, (size_t)nrhs);
  grow(dvr, capv, (size_t)n * nrhs);
  for (int b = 0; b < 2 && !rc; ++b) {
    if (!stage[b] && hipHostMalloc((void**)&stage[b], kStageBytes, 0) != hipSuccess) {
      stage[b] = nullptr;
      rc = HSRQ_ENOMEM;
    }
    if (!rc && !staged[b] && hipEventCreateWithFlags(&staged[b], hipEventDisableTiming) != hipSuccess)
      rc = HSRQ_EHIP;
  }
  if (!rc) {
    hipStream_t st = static_cast<hipStream_t>(hsrq_context_stream(ctx));
    const int per = (int)std::max<size_t>(1, kStageBytes / ((size_t)n * sizeof(double)));
    for (int j0 = 0, c = 0; j0 < n && !rc; j0 += per, ++c) {
      const int w = std::min(per, n - j0), b = c & 1;
      if (c >= 2 && hipEventSynchronize(staged[b]) != hipSuccess) {
        rc = HSRQ_EHIP;
        break;
      }
      double* dst = reinterpret_cast<double*>(stage[b]);
      auto copy_cols = [&](int ja, int jb) {
        for (int j = ja; j < jb; ++j)
          std::memcpy(dst + (size_t)(j - j0) * n, H + (size_t)j * ldH,
                      (size_t)std::min(n, j + 2) * sizeof(double));
      };
      // large chunks: four host threads (one memcpy stream does not reach
      // the host's copy bandwidth)
      const size_t chunk_bytes = (size_t)w * std::min(n, j0 + w) * sizeof(double);
      if (chunk_bytes >= ((size_t)4 << 20) && w >= 4) {
        std::thread th[3];
        for (int t = 0; t < 3; ++t)
          th[t] = std::thread(copy_cols, j0 + w * t / 4, j0 + w * (t + 1) / 4);
        copy_cols(j0 + w * 3 / 4, j0 + w);
        for (auto& x : th) x.join();
      } else {
        copy_cols(j0, j0 + w);
      }
      if (hipMemcpyAsync(dH + (size_t)j0 * n, dst, (size_t)w * n * sizeof(double),
                         hipMemcpyHostToDevice, st) != hipSuccess ||
          hipEventRecord(staged[b], st) != hipSuccess)
        rc = HSRQ_EHIP;
    }
  }
  if (!rc && hipMemcpy(dwr, wr, (size_t)nrhs * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = HSRQ_EHIP;
  std::vector<int> status(nrhs, 0);
  if (!rc) rc = hsrq_solve_device(ctx, n, dH, n, nrhs, dwr, dvr, n, tlsz, status.data());
  if (!rc && hipMemcpy2D(vr, (size_t)ldvr * sizeof(double), dvr, (size_t)n * sizeof(double),
                         (size_t)n * sizeof(double), nrhs, hipMemcpyDeviceToHost) != hipSuccess)
    rc = HSRQ_EHIP;
  if (!rc && hipMemcpy(wr, dwr, (size_t)nrhs * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = HSRQ_EHIP;
  if (rc) {
    std::fprintf(stderr, "tiled_inverse_iteration: %s\n", hsrq_strerror(rc));
    return;
  }
  // the reference reports each unconverged vector on stdout
  // (shifted_Hessenberg_solve.c:232-233)
  for (int k = 0; k < nrhs; ++k)
    if (!(status[k] & HSRQ_CONVERGED)) std::printf("not converged\n");
}

}  // extern "C"
