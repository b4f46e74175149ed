// hsrq_post.cpp -- the steps after the solve, on this library's FP64 MFMA
// products (hsrq_gemm.hip):
//
//  * back-transformation: with A0 = Z H Z^T (Z orthogonal, the accumulated
//    reflectors of a Hessenberg reduction) and H x = lambda x, A0 (Z x) =
//    lambda (Z x); for m eigenvectors one product V = Z X (2 n^2 m flops);
//  * residual validator (validation.c:79-171 at scale): ||(H - lambda_k I)
//    x_k||_2 / (||H||_oo ||x_k||_2 2^-52) for every eigenvector of any number
//    of problems, in three launches (row maxima, the Hessenberg product with
//    its per-panel reduction, the per-vector combination).
//
// Descriptors and item tables go through a device buffer kept by the context
// and reused in stream order (a later call's upload is ordered after the
// earlier kernels that read it).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/hsrq.h"
#include "hsrq_gemm.h"
#include "hsrq_host.h"

namespace hsrq {

void release_post(hsrq_context* ctx) {
  if (ctx && ctx->post) {
    (void)hipFree(ctx->post);
    ctx->post = nullptr;
    ctx->post_cap = 0;
  }
}

namespace {

// Device copy of `bytes` of host staging in the context's buffer.
int upload(hsrq_context* ctx, const std::vector<char>& host, char** dev) {
  if (host.size() > ctx->post_cap) {
    release_post(ctx);
    const size_t cap = std::max(host.size(), (size_t)64 << 10);
    HIP_TRY(hipMalloc(&ctx->post, cap));
    ctx->post_cap = cap;
  }
  HIP_TRY(hipMemcpyAsync(ctx->post, host.data(), host.size(), hipMemcpyHostToDevice,
                         ctx->stream));
  *dev = static_cast<char*>(ctx->post);
  return HSRQ_OK;
}

template <class T>
size_t put(std::vector<char>& h, const std::vector<T>& v) {
  const size_t off = h.size();
  h.resize(off + align_up(std::max<size_t>(v.size() * sizeof(T), 1)));
  if (!v.empty()) std::memcpy(h.data() + off, v.data(), v.size() * sizeof(T));
  return off;
}

}  // namespace

int residuals_run(hsrq_context* ctx, int count, const hsrq_problem* probs, double* res) {
  // per problem its workspace; the residuals of all problems contiguous at
  // the end (one copy to the host)
  std::vector<size_t> off(count + 1, 0);
  size_t total_m = 0;
  for (int i = 0; i < count; ++i) {
    off[i + 1] = off[i] + align_up(resid_workspace_doubles(probs[i].n, probs[i].m) * 8) / 8;
    total_m += probs[i].m;
  }
  double* ws = nullptr;
  HIP_TRY(hipMalloc(&ws, std::max<size_t>(off[count] + total_m, 1) * sizeof(double)));
  double* const res_all = ws + off[count];
  std::vector<GemmDesc> descs(count);
  std::vector<GemmTileItem> tiles, fin;
  size_t rpos = 0;
  for (int i = 0; i < count; ++i) {
    const hsrq_problem& q = probs[i];
    resid_describe(descs[i], q.n, q.H, q.ldH, q.m, q.wr, q.vr, q.ldvr, ws + off[i]);
    descs[i].res = res_all + rpos;
    rpos += q.m;
    if (q.m == 0) continue;
    add_gemm_items(i, q.n, q.m, kGemmColsRes, tiles);
    add_resid_final_items(i, q.m, fin);
  }
  std::vector<char> h;
  const size_t od = put(h, descs), o2 = put(h, tiles), o3 = put(h, fin);
  char* d = nullptr;
  int rc = upload(ctx, h, &d);
  hipError_t e = hipSuccess;
  if (!rc)
    e = launch_resid((const GemmDesc*)(d + od), (const GemmTileItem*)(d + o2), (int)tiles.size(),
                     (const GemmTileItem*)(d + o3), (int)fin.size(), ctx->stream);
  if (!rc && e == hipSuccess && total_m > 0)
    e = hipMemcpyAsync(res, res_all, total_m * sizeof(double), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (!rc && e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(ws);
  if (rc) return rc;
  if (e != hipSuccess) {
    last_hip_error() = e;
    return HSRQ_EHIP;
  }
  return HSRQ_OK;
}

}  // namespace hsrq

using namespace hsrq;

extern "C" int hsrq_backtransform_device(hsrq_context* ctx, int n, int m, const double* dZ,
                                         int ldZ, const double* dX, int ldX, double* dV,
                                         int ldV) {
  if (!ctx || n <= 0 || m < 0 || !dZ || ldZ < n || ldX < n || ldV < n) return HSRQ_EINVAL;
  if (m == 0) return HSRQ_OK;
  if (!dX || !dV) return HSRQ_EINVAL;
  // V is written while Z and X are read: no overlap
  const char* v0 = reinterpret_cast<const char*>(dV);
  const char* v1 = reinterpret_cast<const char*>(dV + (size_t)ldV * (m - 1) + n);
  auto overlaps = [&](const double* p, size_t count) {
    const char* a = reinterpret_cast<const char*>(p);
    const char* b = reinterpret_cast<const char*>(p + count);
    return a < v1 && v0 < b;
  };
  if (overlaps(dZ, (size_t)ldZ * (n - 1) + n) || overlaps(dX, (size_t)ldX * (m - 1) + n))
    return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  GemmDesc D{};
  D.A = dZ;
  D.ldA = ldZ;
  D.B = dX;
  D.ldB = ldX;
  D.C = dV;
  D.ldC = ldV;
  D.n = n;
  D.m = m;
  std::vector<GemmDesc> descs(1, D);
  std::vector<GemmTileItem> tiles;
  add_gemm_items(0, n, m, kGemmColsBT, tiles);
  std::vector<char> h;
  const size_t od = put(h, descs), ot = put(h, tiles);
  char* d = nullptr;
  const int rc = upload(ctx, h, &d);
  if (rc) return rc;
  HIP_TRY(launch_gemm((const GemmDesc*)(d + od), (const GemmTileItem*)(d + ot), (int)tiles.size(),
                      ctx->stream));
  return HSRQ_OK;
}

extern "C" int hsrq_residuals_device(hsrq_context* ctx, int n, const double* dH, int ldH, int m,
                                     const double* dwr, const double* dvr, int ldvr,
                                     double* res) {
  if (!ctx || n <= 0 || m < 0 || ldH < n || ldvr < n || !dH || (m > 0 && (!dwr || !dvr || !res)))
    return HSRQ_EINVAL;
  if (m == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  hsrq_problem q = {n, dH, ldH, m, const_cast<double*>(dwr), const_cast<double*>(dvr), ldvr};
  return residuals_run(ctx, 1, &q, res);
}

extern "C" int hsrq_residuals_batch_device(hsrq_context* ctx, int count,
                                           const hsrq_problem* probs, double* res) {
  if (!ctx || count < 0 || (count > 0 && (!probs || !res))) return HSRQ_EINVAL;
  for (int i = 0; i < count; ++i)
    if (validate_problem(probs[i])) return HSRQ_EINVAL;
  if (count == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  return residuals_run(ctx, count, probs, res);
}
