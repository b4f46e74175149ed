// hsrq_blas.cpp -- the step after the solve: back-transformation of the
// eigenvectors of H to those of the matrix it was reduced from.
//
// If A = Z H Z^T (Z orthogonal: the accumulated Householder reflectors of a
// Hessenberg reduction) and H x = lambda x, then A (Z x) = lambda (Z x).  For
// m eigenvectors this is one plain FP64 GEMM, V = Z X (2 n^2 m flops), run by
// rocBLAS on the context stream; the handle is made on first use and lives
// with the context.
#include <rocblas/rocblas.h>

#include "hsrq_host.h"

namespace hsrq {

void release_blas(hsrq_context* ctx) {
  if (ctx && ctx->blas) {
    (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->blas));
    ctx->blas = nullptr;
  }
}

namespace {

int blas_rc(rocblas_status s) {
  switch (s) {
    case rocblas_status_success: return HSRQ_OK;
    case rocblas_status_memory_error: return HSRQ_ENOMEM;
    case rocblas_status_invalid_handle:
    case rocblas_status_invalid_pointer:
    case rocblas_status_invalid_size:
    case rocblas_status_invalid_value: return HSRQ_EINVAL;
    default: return HSRQ_EHIP;
  }
}

int blas_handle(hsrq_context* ctx, rocblas_handle* h) {
  if (!ctx->blas) {
    rocblas_handle nh = nullptr;
    const int rc = blas_rc(rocblas_create_handle(&nh));
    if (rc) return rc;
    ctx->blas = nh;
  }
  *h = static_cast<rocblas_handle>(ctx->blas);
  return blas_rc(rocblas_set_stream(*h, ctx->stream));
}

}  // namespace
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
  rocblas_handle h = nullptr;
  int rc = blas_handle(ctx, &h);
  if (rc) return rc;
  const double one = 1.0, zero = 0.0;
  rc = blas_rc(rocblas_set_pointer_mode(h, rocblas_pointer_mode_host));
  if (rc) return rc;
  return blas_rc(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n, m, n, &one,
                               dZ, ldZ, dX, ldX, &zero, dV, ldV));
}
