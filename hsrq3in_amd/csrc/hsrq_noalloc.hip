// hsrq_noalloc.hip -- hsrq_solve_device_noalloc: the solve of
// hsrq_solve_device (same tile, same kernels, bit-identical results) for
// callers that cannot allocate device memory.  The workspace, the item tables
// and the problem descriptor live in one file-scope device array, `scratch`,
// of HSRQ_NOALLOC_SCRATCH_BYTES per GPU; the host side stages its tables in
// static host buffers.  The call makes no device or host allocation: it
// copies, launches and synchronizes on the context stream only (no lookahead
// stream, no events; the schedule is the serial one, whose results equal the
// overlapped one's bit for bit).
//
// Every call on a GPU shares that GPU's array, which keeps the last call's
// contents (the kernels write everything they read: nothing needs clearing),
// and every call shares the host staging buffers, so calls are serialized by
// one process-wide lock.  Without it, two overlapping calls on one GPU would
// race on the array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/hsrq.h"
#include "hsrq_host.h"

// The allocation-free entry point's device memory: workspace, item tables
// and descriptor.  In host code `scratch` names a host-side stand-in, not a
// device address; only the address from hipGetSymbolAddress (per device,
// noalloc_setup) is passed to kernels and copies.
static __device__ __align__(256) unsigned char scratch[HSRQ_NOALLOC_SCRATCH_BYTES];

namespace hsrq {

namespace {

constexpr size_t kScratchBytes = HSRQ_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxDevices = 64;
constexpr int kMaxShifts = HSRQ_NOALLOC_MAX_SHIFTS;
constexpr int kMaxSteps = 1024;                // tile columns per problem
constexpr size_t kStageBytes = 8u << 20;       // descriptor + item tables

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};                 // device address of `scratch`
hipError_t g_setup[kMaxDevices] = {};          // result of the set-up per device
bool g_done[kMaxDevices] = {};

alignas(256) char g_stage[kStageBytes];
int g_status[kMaxShifts], g_st2[kMaxShifts], g_order[kMaxShifts];
double g_wr[kMaxShifts], g_wr2[kMaxShifts];
size_t g_k1_off[kMaxSteps], g_gb_off[kMaxSteps];
int g_nk1[kMaxSteps], g_ngb[kMaxSteps];

// push_back sinks of the item builders: one counts, one writes into the
// staging buffer (bounds fixed by the count).
template <class T>
struct Count {
  size_t n = 0;
  void push_back(const T&) { ++n; }
};
template <class T>
struct Fill {
  T* p;
  size_t n = 0;
  void push_back(const T& x) { p[n++] = x; }
};

// One pass (the first solve or one restart) of problem q with start vector
// `its` (-1: eps3 * 1, >= 0: dlaein's its-th vector).  Status bits of the
// pass's m shifts go to st; *ws_dev receives the workspace address (idle
// after the pass).
int run_pass(hsrq_context* ctx, char* dptr, const hsrq_problem& q, int b, int pack, int its,
             int* st, char** ws_dev) {
  ProbDev P;
  prob_describe(P, q, b, its, ctx->exact, dptr);   // shape only (pointers unused)
  if (P.T > kMaxSteps) return HSRQ_ESCRATCH;
  // sizes of the tables
  Count<GemmItem> cpart;
  Count<RowItem> cfin, cinit;
  Count<ShiftItem> cbt;
  add_norm_items(P, 0, cpart, cfin);
  add_init_items(P, 0, cinit);
  add_bt_items(P, 0, cbt);
  size_t off = align_up(sizeof(ProbDev));
  const size_t part_off = off;
  off += align_up(cpart.n * sizeof(GemmItem));
  const size_t fin_off = off;
  off += align_up(cfin.n * sizeof(RowItem));
  const size_t init_off = off;
  off += align_up(cinit.n * sizeof(RowItem));
  const size_t bt_off = off;
  off += align_up(cbt.n * sizeof(ShiftItem));
  const int steps = P.T;
  for (int s = 0; s < steps; ++s) {
    Count<K1Item> k1;
    Count<GemmItem> gb;
    add_step_items(P, 0, P.T - 1 - s, pack, k1, gb);
    g_k1_off[s] = off;
    g_nk1[s] = (int)k1.n;
    off += align_up(k1.n * sizeof(K1Item));
    g_gb_off[s] = off;
    g_ngb[s] = (int)gb.n;
    off += align_up(gb.n * sizeof(GemmItem));
  }
  const size_t stage = off, ws_off = align_up(off);
  if (stage > kStageBytes || ws_off + prob_workspace_bytes(q.n, q.m, b) > kScratchBytes)
    return HSRQ_ESCRATCH;
  char* ws = dptr + ws_off;
  prob_describe(P, q, b, its, ctx->exact, ws);
  // fill the staging buffer: descriptor, then the tables
  std::memcpy(g_stage, &P, sizeof(ProbDev));
  Fill<GemmItem> part{(GemmItem*)(g_stage + part_off)};
  Fill<RowItem> fin{(RowItem*)(g_stage + fin_off)};
  Fill<RowItem> init{(RowItem*)(g_stage + init_off)};
  Fill<ShiftItem> bt{(ShiftItem*)(g_stage + bt_off)};
  add_norm_items(P, 0, part, fin);
  add_init_items(P, 0, init);
  add_bt_items(P, 0, bt);
  for (int s = 0; s < steps; ++s) {
    Fill<K1Item> k1{(K1Item*)(g_stage + g_k1_off[s])};
    Fill<GemmItem> gb{(GemmItem*)(g_stage + g_gb_off[s])};
    add_step_items(P, 0, P.T - 1 - s, pack, k1, gb);
  }
  hipStream_t st_ = ctx->stream;
  HIP_TRY(hipMemcpyAsync(dptr, g_stage, stage, hipMemcpyHostToDevice, st_));
  const ProbDev* dP = (const ProbDev*)dptr;
  HIP_TRY(launch_norms(dP, (const GemmItem*)(dptr + part_off), (int)cpart.n,
                       (const RowItem*)(dptr + fin_off), (int)cfin.n, st_));
  HIP_TRY(launch_init(dP, (const RowItem*)(dptr + init_off), (int)cinit.n, st_));
  for (int s = 0; s < steps; ++s) {
    HIP_TRY(launch_sweep(pack, b, dP, (const K1Item*)(dptr + g_k1_off[s]), g_nk1[s], st_));
    if (g_ngb[s] > 0)
      HIP_TRY(launch_update(dP, (const GemmItem*)(dptr + g_gb_off[s]), g_ngb[s], st_, 0));
  }
  HIP_TRY(launch_backtransform(dP, (const ShiftItem*)(dptr + bt_off), (int)cbt.n, st_));
  HIP_TRY(hipMemcpyAsync(st, P.status, q.m * sizeof(int), hipMemcpyDeviceToHost, st_));
  HIP_TRY(hipStreamSynchronize(st_));
  *ws_dev = ws;
  return HSRQ_OK;
}

}  // namespace

// Set-up of one device (called by hsrq_context_create with the device
// current): the device address of its copy of `scratch`.  A failure is kept
// and reported by the entry point; the rest of the library does not need it.
void noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_done[device]) return;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_setup[device] = e;
  g_ptr[device] = e == hipSuccess ? p : nullptr;
  g_done[device] = true;
}

}  // namespace hsrq

using namespace hsrq;

extern "C" int hsrq_solve_device_noalloc(hsrq_context* ctx, int n, const double* dH, int ldH,
                                         int m, double* dwr, double* dvr, int ldvr, int tlsz,
                                         int* status) {
  if (!ctx) return HSRQ_EINVAL;
  hsrq_problem q = {n, dH, ldH, m, dwr, dvr, ldvr};
  if (validate_problem(q)) return HSRQ_EINVAL;
  if (ctx->precision != HSRQ_PREC_FP64) return HSRQ_EUNSUPPORTED;
  if (m > kMaxShifts) return HSRQ_ESCRATCH;
  if (m == 0) return HSRQ_OK;
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return HSRQ_EUNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_done[dev] || !g_ptr[dev]) {
    last_hip_error() = g_done[dev] ? g_setup[dev] : hipErrorNotInitialized;
    return HSRQ_EHIP;
  }
  HIP_TRY(hipSetDevice(dev));
  char* dptr = (char*)g_ptr[dev];
  // the tile and sweep layout of hsrq_solve_device for one problem
  const int b = (tlsz > 0 || ctx->tile > 0) ? choose_tile(ctx, tlsz, n)
                                            : choose_tile(ctx, n < 4096 ? 64 : 128, 0);
  const int pack = choose_pack(ctx->pack, b, m);
  char* ws = nullptr;
  int rc = run_pass(ctx, dptr, q, b, pack, -1, g_status, &ws);
  if (rc) return rc;
  // restarts (hsrq_solve_device's semantics: dlaein vectors its = 0, 1, ...
  // for the unconverged tail, converged shifts first, stable)
  hipStream_t st = ctx->stream;
  int first = 0;
  for (int its = 0; its < n; ++its) {
    const int cnt = m - first;
    int nc = 0;
    for (int r = first; r < m; ++r) nc += (g_status[r] & HSRQ_CONVERGED) ? 1 : 0;
    if (nc == cnt) break;
    int t = 0;
    for (int r = first; r < m; ++r)
      if (g_status[r] & HSRQ_CONVERGED) g_order[t++] = r;
    for (int r = first; r < m; ++r)
      if (!(g_status[r] & HSRQ_CONVERGED)) g_order[t++] = r;
    HIP_TRY(hipMemcpy(g_wr, dwr, m * sizeof(double), hipMemcpyDeviceToHost));
    std::memcpy(g_wr2, g_wr, m * sizeof(double));
    std::memcpy(g_st2, g_status, m * sizeof(int));
    // the columns move through the idle workspace of the last pass (at least
    // 4 n m doubles of that pass; cnt <= its m)
    double* tmp = (double*)ws;
    for (int k = 0; k < cnt; ++k) {
      const int src = g_order[k];
      g_wr2[first + k] = g_wr[src];
      g_st2[first + k] = g_status[src];
      HIP_TRY(hipMemcpyAsync(tmp + (size_t)k * n, dvr + (size_t)src * ldvr, n * sizeof(double),
                             hipMemcpyDeviceToDevice, st));
    }
    HIP_TRY(hipMemcpy2DAsync(dvr + (size_t)first * ldvr, ldvr * sizeof(double), tmp,
                             n * sizeof(double), n * sizeof(double), cnt,
                             hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(dwr, g_wr2, m * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::memcpy(g_status, g_st2, m * sizeof(int));
    first += nc;
    hsrq_problem tail = q;
    tail.m = m - first;
    tail.wr = dwr + first;
    tail.vr = dvr + (size_t)first * ldvr;
    rc = run_pass(ctx, dptr, tail, b, choose_pack(ctx->pack, b, tail.m), its, g_status + first,
                  &ws);
    if (rc) return rc;
  }
  if (status) std::memcpy(status, g_status, m * sizeof(int));
  return HSRQ_OK;
}
