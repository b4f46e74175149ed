// hsrq_server.cpp -- continuous-batching request server.
//
// Requests (one Hessenberg problem and its shifts each) arrive at any time.
// A worker thread runs the device schedule in ticks.  One tick advances
// every active request by one tile column of its own sweep:
//   norms + init of newly admitted requests,
//   diagonal sweeps (one launch per tile size),
//   scaling exponents, the panel passes, back-transforms of requests that
//   reached tile column 0.
// Every kernel reads item tables, so requests of different n, m and
// progress share each launch.  Requests join at tick boundaries as soon as
// their HBM footprint fits the server's arena (first come, first served),
// and leave when their last tick completes, so the work of requests that
// arrived at different times overlaps on the device.  A request's arithmetic
// does not depend on which other requests share a launch: its results are
// bit-identical alone and in any mix.
//
// Every result is validated online: in the tick of a request's
// back-transform, the residual validator (hsrq_gemm.hip; the reference's
// validate_real_eigenvectors, validation.c:79-171) computes ||(H - mu I) x||
// / (||H||_oo ||x|| eps) of its vectors on the device, and the status words
// returned carry HSRQ_VALIDATED where it is at most 10 n.
//
// H, wr, X and the per-(tile, shift) scale exponents stay resident in the
// arena until the client fetches the result.  Unconverged shifts restart
// in the same slot with the dlaein vectors (converged first, as the
// reference's intent at inverse_iteration.c:75-125, :160-328).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/hsrq.h"
#include "hsrq_gemm.h"
#include "hsrq_host.h"

using namespace hsrq;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (host bookkeeping only).
class Arena {
 public:
  void init(char* base, size_t bytes) {
    base_ = base;
    free_.clear();
    free_[0] = bytes;
    size_ = bytes;
  }
  // offset or SIZE_MAX
  size_t alloc(size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes) {
        const size_t off = it->first, len = it->second;
        free_.erase(it);
        if (len > bytes) free_[off + bytes] = len - bytes;
        used_ += bytes;
        peak_ = std::max(peak_, used_);
        return off;
      }
    return SIZE_MAX;
  }
  void release(size_t off, size_t bytes) {
    bytes = align_up(std::max<size_t>(bytes, 1));
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char* ptr(size_t off) const { return base_ + off; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t size() const { return size_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
};

enum class St { Queued, Copying, Ready, Active, Finishing, Done, Failed };

struct Req {
  int64_t id = 0;
  hsrq_request in{};
  int b = 64;
  size_t footprint = 0, off = SIZE_MAX;
  // device copies (host requests) or the caller's buffers (device requests)
  const double* dH = nullptr;
  int ldH = 0;
  double* dwr = nullptr;
  double* dvr = nullptr;
  int ldvr = 0;
  char* ws = nullptr;
  hipEvent_t ev_copy = nullptr;
  St st = St::Queued;
  int slot = -1;
  int J = 0;
  int its = -1;          // start vector of the current solve
  int first = 0;         // shifts [0, first) converged and final
  std::vector<double> wr;      // host copy, final order
  std::vector<int> status;     // host copy, final order
  int* st_pinned = nullptr;    // status of the last solve (pinned)
  double* res_pinned = nullptr;  // its residuals (pinned)
  std::vector<double> res;     // host copy, final order
  bool fresh = false;          // needs norms + init in the next tick
  bool waited = false;         // a client waits for (or fetched) the result
  int rc = HSRQ_OK;
  Clock::time_point t_submit, t_done;
};

}  // namespace

struct hsrq_server {
  hsrq_context* ctx = nullptr;
  int device = 0;
  int tile = 0;
  int max_active = 1024;
  hipStream_t stream = nullptr, copy = nullptr;
  void* d_arena = nullptr;
  Arena arena;
  ProbDev* d_probs = nullptr;           // max_active descriptors
  GemmDesc* d_gdesc = nullptr;          // max_active validator descriptors
  std::vector<int> free_slots;
  // tick buffers, double buffered
  char* h_items[2] = {nullptr, nullptr};
  void* d_items[2] = {nullptr, nullptr};
  size_t cap_items[2] = {0, 0};
  ProbDev* h_probs[2] = {nullptr, nullptr};
  GemmDesc* h_gdesc[2] = {nullptr, nullptr};
  hipEvent_t ev_tick[2] = {nullptr, nullptr};
  bool tick_pending[2] = {false, false};
  std::vector<Req*> finishing[2];
  // requests
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<Req*> queue;
  std::vector<Req*> live;               // admitted, not yet fetched
  std::map<int64_t, Req*> by_id;
  int64_t next_id = 1;
  bool stop = false;
  std::thread worker;
  // statistics
  int64_t submitted = 0, completed = 0, eigvecs = 0, ticks = 0, validated = 0;
  double worst_residual = 0.0;   // max residual / (10 n) of completed requests
  std::vector<double> lat_ms;
  Clock::time_point t_first, t_last;
  bool any = false;
  int worker_rc = HSRQ_OK;
};

namespace {

// Workspace of the request's solve, then of its validation.
size_t footprint(const hsrq_request& r, int b) {
  size_t f = prob_workspace_bytes(r.n, r.m, b) + align_up(resid_workspace_doubles(r.n, r.m) * 8);
  if (r.location == HSRQ_ON_HOST)
    f += align_up((size_t)r.n * r.n * sizeof(double)) + align_up((size_t)r.m * sizeof(double)) +
         align_up((size_t)r.n * r.m * sizeof(double));
  return f;
}

int auto_tile_server(const hsrq_server* s, int n) {
  return choose_tile(nullptr, s->tile > 0 ? s->tile : 128, n);
}

// (Re)describes request r in its slot for the solve of shifts [first, m).
void describe(hsrq_server* s, Req* r, ProbDev& P) {
  hsrq_problem q;
  q.n = r->in.n;
  q.H = r->dH;
  q.ldH = r->ldH;
  q.m = r->in.m - r->first;
  q.wr = r->dwr + r->first;
  q.vr = r->dvr + (size_t)r->first * r->ldvr;
  q.ldvr = r->ldvr;
  prob_describe(P, q, r->b, r->its, s->ctx->exact, r->ws);
}

// Admission and copies (worker thread, lock held).
int admit(hsrq_server* s) {
  while (!s->stop && !s->queue.empty() && !s->free_slots.empty()) {
    Req* r = s->queue.front();
    const size_t off = s->arena.alloc(r->footprint);
    if (off == SIZE_MAX) break;   // first come, first served
    s->queue.pop_front();
    r->off = off;
    r->slot = s->free_slots.back();
    s->free_slots.pop_back();
    char* p = s->arena.ptr(off);
    const int n = r->in.n, m = r->in.m;
    if (r->in.location == HSRQ_ON_HOST) {
      double* H = (double*)p;
      p += align_up((size_t)n * n * sizeof(double));
      r->dwr = (double*)p;
      p += align_up((size_t)m * sizeof(double));
      r->dvr = (double*)p;
      p += align_up((size_t)n * m * sizeof(double));
      r->dH = H;
      r->ldH = n;
      r->ldvr = n;
      HIP_TRY(hipMemcpy2DAsync(H, n * sizeof(double), r->in.H, (size_t)r->in.ldH * sizeof(double),
                               n * sizeof(double), n, hipMemcpyHostToDevice, s->copy));
      if (m > 0)
        HIP_TRY(hipMemcpyAsync(r->dwr, r->in.wr, m * sizeof(double), hipMemcpyHostToDevice, s->copy));
      HIP_TRY(hipEventRecord(r->ev_copy, s->copy));
      r->st = St::Copying;
    } else {
      r->dH = r->in.H;
      r->ldH = r->in.ldH;
      r->dwr = const_cast<double*>(r->in.wr);
      r->dvr = r->in.vr;
      r->ldvr = r->in.ldvr;
      r->st = St::Ready;
    }
    r->ws = p;
    s->live.push_back(r);
  }
  for (Req* r : s->live)
    if (r->st == St::Copying) {
      const hipError_t e = hipEventQuery(r->ev_copy);
      if (e == hipSuccess)
        r->st = St::Ready;
      else if (e != hipErrorNotReady)
        HIP_TRY(e);
    }
  return HSRQ_OK;
}

void finish(hsrq_server* s, Req* r) {
  r->st = St::Done;
  r->t_done = Clock::now();
  s->completed += 1;
  const double bound = 10.0 * r->in.n;
  for (int k = 0; k < r->in.m; ++k) {
    s->eigvecs += (r->status[k] & HSRQ_CONVERGED) ? 1 : 0;
    if (r->res[k] <= bound) {
      r->status[k] |= HSRQ_VALIDATED;
      s->validated += 1;
    }
    if (r->status[k] & HSRQ_CONVERGED)
      s->worst_residual = std::max(s->worst_residual, std::isnan(r->res[k]) ? INFINITY
                                                                            : r->res[k] / bound);
  }
  s->lat_ms.push_back(std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count());
  s->t_last = r->t_done;
  // a client may already wait for r (requests without shifts finish in
  // enqueue_tick, outside any tick's retirement)
  s->cv_done.notify_all();
}

// After the tick in which r's back-transform ran: merge the status, restart
// the unconverged shifts or finish (worker thread, lock held).
int complete(hsrq_server* s, Req* r) {
  const int m = r->in.m, cnt = m - r->first;
  for (int t = 0; t < cnt; ++t) {
    r->status[r->first + t] = r->st_pinned[t];
    r->res[r->first + t] = r->res_pinned[t];
  }
  int nc = 0;
  for (int t = r->first; t < m; ++t) nc += (r->status[t] & HSRQ_CONVERGED) ? 1 : 0;
  const int its_next = r->its + 1;
  if (nc == cnt || its_next >= r->in.n) {
    finish(s, r);
    return HSRQ_OK;
  }
  // stable partition of [first, m), converged first; vr columns move through
  // the idle workspace (its first n*m doubles: the cross-over column buffer)
  std::vector<int> order;
  for (int t = r->first; t < m; ++t)
    if (r->status[t] & HSRQ_CONVERGED) order.push_back(t);
  for (int t = r->first; t < m; ++t)
    if (!(r->status[t] & HSRQ_CONVERGED)) order.push_back(t);
  std::vector<double> wr_new(r->wr), res_new(r->res);
  std::vector<int> st_new(r->status);
  double* tmp = (double*)r->ws;
  const int n = r->in.n;
  for (int t = 0; t < cnt; ++t) {
    const int src = order[t];
    wr_new[r->first + t] = r->wr[src];
    st_new[r->first + t] = r->status[src];
    res_new[r->first + t] = r->res[src];
    HIP_TRY(hipMemcpyAsync(tmp + (size_t)t * n, r->dvr + (size_t)src * r->ldvr, n * sizeof(double),
                           hipMemcpyDeviceToDevice, s->stream));
  }
  HIP_TRY(hipMemcpy2DAsync(r->dvr + (size_t)r->first * r->ldvr, r->ldvr * sizeof(double), tmp,
                           n * sizeof(double), n * sizeof(double), cnt, hipMemcpyDeviceToDevice,
                           s->stream));
  r->wr = wr_new;
  r->status = st_new;
  r->res = res_new;
  HIP_TRY(hipMemcpyAsync(r->dwr, r->wr.data(), m * sizeof(double), hipMemcpyHostToDevice, s->stream));
  HIP_TRY(hipStreamSynchronize(s->stream));
  r->first += nc;
  r->its = its_next;
  r->st = St::Ready;
  return HSRQ_OK;
}

// Builds and enqueues one tick.  Returns 1 if work was enqueued.
int enqueue_tick(hsrq_server* s, int par, int& rc) {
  rc = HSRQ_OK;
  std::vector<GemmItem> part;
  std::vector<RowItem> fin, init;
  std::vector<ShiftItem> bt;
  std::vector<GemmTileItem> vtile, vfin;   // validator of the finishing requests
  std::vector<GemmDesc> gdescs;
  std::map<int, StepItems> steps;   // by tile size
  std::vector<Req*> fresh, fin_reqs;
  int shifts = 0;
  for (Req* r : s->live) {
    if (r->st == St::Ready) {
      if (r->in.m - r->first == 0) {
        finish(s, r);
        continue;
      }
      r->st = St::Active;
      r->fresh = true;
    }
    if (r->st == St::Active) shifts += r->in.m - r->first;
  }
  // one layout for the whole tick (both layouts give identical results)
  const int pack_req = s->ctx->pack;
  std::vector<ProbDev> descs;
  for (Req* r : s->live) {
    if (r->st != St::Active) continue;
    ProbDev P;
    describe(s, r, P);
    if (r->fresh) {
      fresh.push_back(r);
      r->J = P.T - 1;
      add_norm_items(P, r->slot, part, fin);
      add_init_items(P, r->slot, init);
    }
    const int pack = choose_pack(pack_req, r->b, shifts);
    add_step_items(P, r->slot, r->J, pack, steps[r->b * 8 + pack]);
    if (r->J == 0) {
      add_bt_items(P, r->slot, bt);
      fin_reqs.push_back(r);
      // validation of this solve's vectors (shifts first .. m-1)
      GemmDesc G;
      const int mm = r->in.m - r->first;
      resid_describe(G, r->in.n, r->dH, r->ldH, mm, r->dwr + r->first,
                     r->dvr + (size_t)r->first * r->ldvr, r->ldvr,
                     (double*)(r->ws + prob_workspace_bytes(r->in.n, r->in.m, r->b)));
      gdescs.push_back(G);
      add_gemm_items(r->slot, r->in.n, mm, kGemmColsRes, vtile);
      add_resid_final_items(r->slot, mm, vfin);
    }
    if (r->fresh) descs.push_back(P);
  }
  if (fresh.empty() && steps.empty()) return 0;

  // pack the item tables into the pinned staging buffer of this parity
  size_t need = align_up(part.size() * sizeof(GemmItem)) + align_up(fin.size() * sizeof(RowItem)) +
                align_up(init.size() * sizeof(RowItem)) + align_up(bt.size() * sizeof(ShiftItem)) +
                align_up(vtile.size() * sizeof(GemmTileItem)) +
                align_up(vfin.size() * sizeof(GemmTileItem));
  for (auto& kv : steps)
    need += align_up(kv.second.k1.size() * sizeof(K1Item)) +
            align_up(kv.second.gb.size() * sizeof(GemmItem));
  need = std::max(need, kAlign);
  if (need > s->cap_items[par]) {
    size_t cap = std::max(need, 2 * s->cap_items[par]);
    if (s->h_items[par]) (void)hipHostFree(s->h_items[par]);
    if (s->d_items[par]) (void)hipFree(s->d_items[par]);
    s->h_items[par] = nullptr;
    s->d_items[par] = nullptr;
    s->cap_items[par] = 0;
    hipError_t e = hipHostMalloc((void**)&s->h_items[par], cap, 0);
    if (e == hipSuccess) e = hipMalloc(&s->d_items[par], cap);
    if (e != hipSuccess) {
      last_hip_error() = e;
      rc = HSRQ_ENOMEM;
      return 0;
    }
    s->cap_items[par] = cap;
  }
  size_t pos = 0;
  char* h = s->h_items[par];
  char* d = (char*)s->d_items[par];
  auto put = [&](const void* src, size_t sz) -> const void* {
    if (sz) std::memcpy(h + pos, src, sz);
    const void* dp = d + pos;
    pos += align_up(sz);
    return dp;
  };
  const GemmItem* d_part = (const GemmItem*)put(part.data(), part.size() * sizeof(GemmItem));
  const RowItem* d_fin = (const RowItem*)put(fin.data(), fin.size() * sizeof(RowItem));
  const RowItem* d_init = (const RowItem*)put(init.data(), init.size() * sizeof(RowItem));
  const ShiftItem* d_bt = (const ShiftItem*)put(bt.data(), bt.size() * sizeof(ShiftItem));
  const GemmTileItem* d_vtile =
      (const GemmTileItem*)put(vtile.data(), vtile.size() * sizeof(GemmTileItem));
  const GemmTileItem* d_vfin =
      (const GemmTileItem*)put(vfin.data(), vfin.size() * sizeof(GemmTileItem));
  struct Dev {
    int b, pack;
    const K1Item* k1;
    const GemmItem* gb;
    int nk1, ngb;
  };
  std::vector<Dev> dev;
  for (auto& kv : steps) {
    Dev x;
    x.b = kv.first / 8;
    x.pack = kv.first % 8;
    x.k1 = (const K1Item*)put(kv.second.k1.data(), kv.second.k1.size() * sizeof(K1Item));
    x.gb = (const GemmItem*)put(kv.second.gb.data(), kv.second.gb.size() * sizeof(GemmItem));
    x.nk1 = (int)kv.second.k1.size();
    x.ngb = (int)kv.second.gb.size();
    dev.push_back(x);
  }
  auto fail = [&](hipError_t e) {
    last_hip_error() = e;
    rc = HSRQ_EHIP;
    return 0;
  };
  hipError_t e = hipMemcpyAsync(d, h, pos, hipMemcpyHostToDevice, s->stream);
  if (e != hipSuccess) return fail(e);
  // descriptors of (re)started requests
  for (size_t i = 0; i < descs.size(); ++i) s->h_probs[par][i] = descs[i];
  for (size_t i = 0; i < descs.size(); ++i) {
    e = hipMemcpyAsync(s->d_probs + fresh[i]->slot, s->h_probs[par] + i, sizeof(ProbDev),
                       hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return fail(e);
  }
  for (size_t i = 0; i < gdescs.size(); ++i) {
    s->h_gdesc[par][i] = gdescs[i];
    e = hipMemcpyAsync(s->d_gdesc + fin_reqs[i]->slot, s->h_gdesc[par] + i, sizeof(GemmDesc),
                       hipMemcpyHostToDevice, s->stream);
    if (e != hipSuccess) return fail(e);
  }
  // wait for the copies of H and wr of the fresh requests
  for (Req* r : fresh)
    if (r->in.location == HSRQ_ON_HOST && r->its < 0) {
      e = hipStreamWaitEvent(s->stream, r->ev_copy, 0);
      if (e != hipSuccess) return fail(e);
    }
  const ProbDev* P = s->d_probs;
  if ((e = launch_norms(P, d_part, (int)part.size(), d_fin, (int)fin.size(), s->stream)) != hipSuccess)
    return fail(e);
  if ((e = launch_init(P, d_init, (int)init.size(), s->stream)) != hipSuccess) return fail(e);
  for (const Dev& x : dev)
    if ((e = launch_sweep(x.pack, x.b, P, x.k1, x.nk1, s->stream)) != hipSuccess) return fail(e);
  for (const Dev& x : dev)
    if ((e = launch_update(P, x.gb, x.ngb, s->stream)) != hipSuccess) return fail(e);
  if ((e = launch_backtransform(P, d_bt, (int)bt.size(), s->stream)) != hipSuccess) return fail(e);
  if ((e = launch_resid(s->d_gdesc, d_vtile, (int)vtile.size(), d_vfin, (int)vfin.size(),
                        s->stream)) != hipSuccess)
    return fail(e);
  for (size_t i = 0; i < fin_reqs.size(); ++i) {
    Req* r = fin_reqs[i];
    ProbDev Q;
    describe(s, r, Q);
    e = hipMemcpyAsync(r->st_pinned, Q.status, (r->in.m - r->first) * sizeof(int),
                       hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(r->res_pinned, gdescs[i].res, (r->in.m - r->first) * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream);
    if (e != hipSuccess) return fail(e);
  }
  if ((e = hipEventRecord(s->ev_tick[par], s->stream)) != hipSuccess) return fail(e);
  for (Req* r : s->live)
    if (r->st == St::Active) {
      r->fresh = false;
      r->J -= 1;
      if (r->J < 0) r->st = St::Finishing;
    }
  s->finishing[par] = fin_reqs;
  s->tick_pending[par] = true;
  s->ticks += 1;
  return 1;
}

// Completion of the tick of parity par (lock held while processing).
int retire_tick(hsrq_server* s, int par, std::unique_lock<std::mutex>& lk) {
  if (!s->tick_pending[par]) return HSRQ_OK;
  lk.unlock();
  const hipError_t e = hipEventSynchronize(s->ev_tick[par]);
  lk.lock();
  s->tick_pending[par] = false;
  if (e != hipSuccess) {
    last_hip_error() = e;
    return HSRQ_EHIP;
  }
  for (Req* r : s->finishing[par]) {
    const int rc = complete(s, r);
    if (rc) return rc;
  }
  s->finishing[par].clear();
  s->cv_done.notify_all();
  return HSRQ_OK;
}

void worker_main(hsrq_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  int par = 0;
  for (;;) {
    int rc = admit(s);
    int work = 0;
    if (!rc) work = enqueue_tick(s, par, rc);
    // keep at most two ticks in flight: retire the previous one
    if (!rc) rc = retire_tick(s, par ^ 1, lk);
    if (rc) {
      s->worker_rc = rc;
      for (Req* r : s->live)
        if (r->st != St::Done) {
          r->st = St::Failed;
          r->rc = rc;
        }
      for (Req* r : s->queue) {
        r->st = St::Failed;
        r->rc = rc;
      }
      s->cv_done.notify_all();
      return;
    }
    if (work) {
      par ^= 1;
      continue;
    }
    if (s->tick_pending[par]) {
      rc = retire_tick(s, par, lk);
      continue;
    }
    bool ready = false, copying = false;
    for (Req* r : s->live) {
      ready = ready || r->st == St::Ready;
      copying = copying || r->st == St::Copying;
    }
    if (ready) continue;   // restarted requests: next tick
    if (copying) {         // host-to-device copies in flight: poll them
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(50));
      lk.lock();
      continue;
    }
    if (s->stop) return;
    s->cv.wait(lk);
  }
}

void release(hsrq_server* s, Req* r) {
  if (r->off != SIZE_MAX) s->arena.release(r->off, r->footprint);
  if (r->slot >= 0) s->free_slots.push_back(r->slot);
  s->live.erase(std::remove(s->live.begin(), s->live.end(), r), s->live.end());
  s->by_id.erase(r->id);
  if (r->ev_copy) (void)hipEventDestroy(r->ev_copy);
  if (r->st_pinned) (void)hipHostFree(r->st_pinned);
  if (r->res_pinned) (void)hipHostFree(r->res_pinned);
  delete r;
  s->cv.notify_all();   // memory became free: admission may proceed
}

}  // namespace

extern "C" {

int hsrq_server_create(hsrq_context* ctx, size_t hbm_bytes, int max_active, int tlsz,
                       hsrq_server** out) {
  if (!ctx || !out || hbm_bytes == 0 || max_active <= 0 || tlsz < 0 || tlsz > 256)
    return HSRQ_EINVAL;
  *out = nullptr;
  HIP_TRY(hipSetDevice(ctx->device));
  hsrq_server* s = new (std::nothrow) hsrq_server();
  if (!s) return HSRQ_ENOMEM;
  s->ctx = ctx;
  s->device = ctx->device;
  s->tile = tlsz;
  s->max_active = max_active;
  auto cleanup = [&](int rc) {
    hsrq_server_destroy(s);
    return rc;
  };
  hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&s->d_arena, hbm_bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_probs, (size_t)max_active * sizeof(ProbDev));
  if (e == hipSuccess) e = hipMalloc((void**)&s->d_gdesc, (size_t)max_active * sizeof(GemmDesc));
  for (int p = 0; p < 2 && e == hipSuccess; ++p) {
    e = hipHostMalloc((void**)&s->h_probs[p], (size_t)max_active * sizeof(ProbDev), 0);
    if (e == hipSuccess)
      e = hipHostMalloc((void**)&s->h_gdesc[p], (size_t)max_active * sizeof(GemmDesc), 0);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s->ev_tick[p], hipEventDisableTiming);
  }
  if (e != hipSuccess) {
    last_hip_error() = e;
    return cleanup(e == hipErrorOutOfMemory ? HSRQ_ENOMEM : HSRQ_EHIP);
  }
  s->arena.init((char*)s->d_arena, hbm_bytes);
  for (int i = max_active - 1; i >= 0; --i) s->free_slots.push_back(i);
  try {
    s->worker = std::thread(worker_main, s);
  } catch (...) {
    return cleanup(HSRQ_ENOMEM);
  }
  *out = s;
  return HSRQ_OK;
}

int hsrq_server_submit(hsrq_server* s, const hsrq_request* req, int64_t* id) {
  if (!s || !req || !id) return HSRQ_EINVAL;
  const hsrq_request& q = *req;
  if (q.n <= 0 || q.m < 0 || q.ldH < q.n || q.ldvr < q.n || !q.H ||
      (q.m > 0 && (!q.wr || !q.vr)) || (q.location != HSRQ_ON_HOST && q.location != HSRQ_ON_DEVICE))
    return HSRQ_EINVAL;
  Req* r = new (std::nothrow) Req();
  if (!r) return HSRQ_ENOMEM;
  r->in = q;
  r->b = auto_tile_server(s, q.n);
  r->footprint = footprint(q, r->b);
  if (r->footprint > s->arena.size()) {
    delete r;
    return HSRQ_EBUSY;   // can never fit this server's HBM budget
  }
  r->wr.resize(q.m);
  r->status.assign(q.m, 0);
  r->res.assign(q.m, NAN);
  hipError_t e = hipSetDevice(s->device);
  if (e == hipSuccess && q.m > 0) {
    if (q.location == HSRQ_ON_HOST)
      std::memcpy(r->wr.data(), q.wr, q.m * sizeof(double));
    else
      e = hipMemcpy(r->wr.data(), q.wr, q.m * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (e == hipSuccess) e = hipHostMalloc((void**)&r->st_pinned, std::max(q.m, 1) * sizeof(int), 0);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&r->res_pinned, std::max(q.m, 1) * sizeof(double), 0);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&r->ev_copy, hipEventDisableTiming);
  if (e != hipSuccess) {
    last_hip_error() = e;
    if (r->st_pinned) (void)hipHostFree(r->st_pinned);
    if (r->res_pinned) (void)hipHostFree(r->res_pinned);
    if (r->ev_copy) (void)hipEventDestroy(r->ev_copy);
    delete r;
    return HSRQ_EHIP;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->worker_rc) {
    (void)hipHostFree(r->st_pinned);
    (void)hipHostFree(r->res_pinned);
    (void)hipEventDestroy(r->ev_copy);
    delete r;
    return s->worker_rc;
  }
  r->id = s->next_id++;
  r->t_submit = Clock::now();
  if (!s->any) {
    s->t_first = r->t_submit;
    s->any = true;
  }
  s->submitted += 1;
  s->by_id[r->id] = r;
  s->queue.push_back(r);
  *id = r->id;
  s->cv.notify_all();
  return HSRQ_OK;
}

int hsrq_server_poll(hsrq_server* s, int64_t id) {
  if (!s) return HSRQ_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->by_id.find(id);
  if (it == s->by_id.end()) return HSRQ_ENOTFOUND;
  if (it->second->st == St::Failed) return it->second->rc;
  return it->second->st == St::Done ? 1 : 0;
}

int hsrq_server_wait(hsrq_server* s, int64_t id, double* wr_out, int* status, double* latency_ms) {
  if (!s) return HSRQ_EINVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->by_id.find(id);
  if (it == s->by_id.end()) return HSRQ_ENOTFOUND;
  Req* r = it->second;
  if (r->waited) return HSRQ_EBUSY;   // another thread waits for it
  r->waited = true;
  s->cv_done.wait(lk, [&] { return r->st == St::Done || r->st == St::Failed; });
  int rc = r->st == St::Failed ? r->rc : HSRQ_OK;
  lk.unlock();   // r stays allocated until release(); the worker skips Done requests
  if (!rc) {
    const int n = r->in.n, m = r->in.m;
    if (r->in.location == HSRQ_ON_HOST && m > 0) {
      // results leave HBM only now
      hipError_t e = hipSetDevice(s->device);
      if (e == hipSuccess)
        e = hipMemcpy2DAsync(r->in.vr, (size_t)r->in.ldvr * sizeof(double), r->dvr,
                             (size_t)r->ldvr * sizeof(double), n * sizeof(double), m,
                             hipMemcpyDeviceToHost, s->copy);
      if (e == hipSuccess) e = hipStreamSynchronize(s->copy);
      if (e != hipSuccess) {
        last_hip_error() = e;
        rc = HSRQ_EHIP;
      }
    }
    if (wr_out && m > 0) std::memcpy(wr_out, r->wr.data(), m * sizeof(double));
    if (status && m > 0) std::memcpy(status, r->status.data(), m * sizeof(int));
    if (latency_ms)
      *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
  }
  lk.lock();
  release(s, r);
  return rc;
}

int hsrq_server_get_stats(hsrq_server* s, hsrq_server_stats* out) {
  if (!s || !out) return HSRQ_EINVAL;
  std::lock_guard<std::mutex> lk(s->mu);
  std::memset(out, 0, sizeof(*out));
  out->submitted = s->submitted;
  out->completed = s->completed;
  out->eigvecs = s->eigvecs;
  out->validated = s->validated;
  out->worst_residual = s->worst_residual;
  out->ticks = s->ticks;
  out->queued = (int64_t)s->queue.size();
  int64_t active = 0;
  for (Req* r : s->live) active += (r->st != St::Done) ? 1 : 0;
  out->active = active;
  out->hbm_bytes = s->arena.size();
  out->hbm_in_use = s->arena.used();
  out->hbm_peak = s->arena.peak();
  if (!s->lat_ms.empty()) {
    std::vector<double> v(s->lat_ms);
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) {
      const double x = q * (v.size() - 1);
      const size_t i = (size_t)x;
      const double f = x - i;
      return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
    };
    out->p50_ms = pct(0.50);
    out->p99_ms = pct(0.99);
    out->span_s = std::chrono::duration<double>(s->t_last - s->t_first).count();
  }
  return HSRQ_OK;
}

void hsrq_server_destroy(hsrq_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // admitted requests complete; requests still queued are dropped
    s->stop = true;
    s->cv.notify_all();
    s->cv_done.wait(lk, [&] {
      if (s->worker_rc || !s->worker.joinable()) return true;
      for (Req* r : s->live)
        if (r->st != St::Done && r->st != St::Failed) return false;
      return true;
    });
  }
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (Req* r : std::vector<Req*>(s->live)) release(s, r);
  for (Req* r : s->queue) {
    if (r->ev_copy) (void)hipEventDestroy(r->ev_copy);
    if (r->st_pinned) (void)hipHostFree(r->st_pinned);
    if (r->res_pinned) (void)hipHostFree(r->res_pinned);
    delete r;
  }
  for (int p = 0; p < 2; ++p) {
    if (s->h_items[p]) (void)hipHostFree(s->h_items[p]);
    if (s->d_items[p]) (void)hipFree(s->d_items[p]);
    if (s->h_probs[p]) (void)hipHostFree(s->h_probs[p]);
    if (s->h_gdesc[p]) (void)hipHostFree(s->h_gdesc[p]);
    if (s->ev_tick[p]) (void)hipEventDestroy(s->ev_tick[p]);
  }
  if (s->d_probs) (void)hipFree(s->d_probs);
  if (s->d_gdesc) (void)hipFree(s->d_gdesc);
  if (s->d_arena) (void)hipFree(s->d_arena);
  if (s->copy) (void)hipStreamDestroy(s->copy);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

}  // extern "C"
