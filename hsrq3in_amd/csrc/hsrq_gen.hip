// hsrq_gen.hip -- test-problem generation on the device, the step before the
// solve: Gaussian upper Hessenberg matrices from a counter-based generator,
// Hessenberg matrices with a prescribed real spectrum (H = U B U^-1), and the
// exact power-of-two grading H <- D H D^-1 with a range check.
//
// Gaussian: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11) -- a 128-bit counter and a 64-bit
// key, ten rounds of two 32x32->64 multiplies, Weyl-sequence key increments
// -- and the Box-Muller transform.  Entry (i, j) of the n x n matrix has the
// linear index e = j n + i (independent of ldH); the pair p = e / 2 is the
// counter (p, 0), the seed the key, and e's parity picks the cosine (even) or
// sine (odd) variate of the pair.  Each entry is a function of (seed, n, i, j)
// alone: the same matrix for any launch shape.
//
// Constructed spectrum: B = diag(lam) + subdiag(nu), U = diag(d) +
// superdiag(u), H = U B U^-1 from H U = U B, column by column:
//   H(:, j) = ((U B)(:, j) - u_{j-1} H(:, j-1)) / d_j,
// which is a first-order recurrence along each ROW: row i needs only its own
// previous entry.  One thread per row, all threads of a wavefront stepping j
// together, so each step stores 64 consecutive doubles of column j.  The
// operations and their order are those of the column build in
// hsrq3in_amd/workloads.py (constructed_spectrum), without contraction, so
// the two agree bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <climits>
#include <vector>

#include "hsrq_host.h"

namespace hsrq {
namespace {

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

__device__ __forceinline__ void philox4x32_10(uint32_t (&x)[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += kPhiloxW0;
      k1 += kPhiloxW1;
    }
    const uint64_t p0 = (uint64_t)kPhiloxM0 * x[0];
    const uint64_t p1 = (uint64_t)kPhiloxM1 * x[2];
    const uint32_t y0 = (uint32_t)(p1 >> 32) ^ x[1] ^ k0;
    const uint32_t y1 = (uint32_t)p1;
    const uint32_t y2 = (uint32_t)(p0 >> 32) ^ x[3] ^ k1;
    const uint32_t y3 = (uint32_t)p0;
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
    x[3] = y3;
  }
}

// The top 52 bits u of hi:lo mapped to (u + 1/2) 2^-52, in the open interval
// (0, 1): the largest, 1 - 2^-53, is exact (with 53 bits it would round to 1).
__device__ __forceinline__ double unit_open(uint32_t hi, uint32_t lo) {
  const uint64_t u = ((uint64_t)hi << 20) | (lo >> 12);
  return ((double)u + 0.5) * 0x1p-52;
}

__global__ __launch_bounds__(256) void k_gaussian_hessenberg(int n, uint64_t seed, double* H,
                                                             int ldH) {
  const int j = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double val = 0.0;
  if (i <= j + 1) {
    const uint64_t e = (uint64_t)j * (uint64_t)n + (uint64_t)i;
    const uint64_t p = e >> 1;
    uint32_t x[4] = {(uint32_t)p, (uint32_t)(p >> 32), 0u, 0u};
    philox4x32_10(x, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double u1 = unit_open(x[0], x[1]);
    const double u2 = unit_open(x[2], x[3]);
    const double rad = sqrt(-2.0 * log(u1));
    const double ang = 6.283185307179586 * u2;  // 2 pi, rounded
    val = (e & 1) ? rad * sin(ang) : rad * cos(ang);
  }
  H[i + (int64_t)j * ldH] = val;
}

// prm: lam[n], nu[n-1], d[n], u[n-1] (nu, u padded by one unused entry).
__global__ __launch_bounds__(64) void k_constructed(int n, const double* __restrict__ prm,
                                                    double* H, int ldH) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= n) return;
  const double* lam = prm;
  const double* nu = prm + n;
  const double* d = prm + 2 * n;
  const double* u = prm + 3 * n;
  double prev = 0.0;  // H(i, j-1)
  for (int j = 0; j < n; ++j) {
    double c;
    if (i > j + 1) {
      c = 0.0;
    } else if (i == j + 1) {
      c = (d[j + 1] * nu[j]) / d[j];
    } else {
      double b;
      if (i == j) {
        b = d[j] * lam[j] + (j < n - 1 ? u[j] * nu[j] : 0.0);
      } else if (i == j - 1) {
        b = u[j - 1] * lam[j];
      } else {
        b = 0.0;
      }
      if (j >= 1) {
        const double t = u[j - 1] * prev;
        b = b - t;
      }
      c = b / d[j];
    }
    H[i + (int64_t)j * ldH] = c;
    prev = c;
  }
}

// Range check of the grading: over the nonzero entries of the Hessenberg part,
// the largest and smallest binary exponent E (|h| = f 2^E, f in [1/2, 1)) after
// scaling by 2^(k_i - k_j), and whether any scaled entry rounds (it does iff
// scaling it back does not restore it: a result below the normal range that
// drops bits).  rng[0] = max, rng[1] = min, rng[2] = inexact flag (start
// INT_MIN, INT_MAX, 0).
__global__ __launch_bounds__(256) void k_grade_range(int n, const double* __restrict__ H, int ldH,
                                                     const int* __restrict__ k, int* rng) {
  const int j = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  int emax = INT_MIN, emin = INT_MAX, inexact = 0;
  if (i < n && i <= j + 1) {
    const double h = H[i + (int64_t)j * ldH];
    if (h != 0.0) {
      const int dk = k[i] - k[j];
      int e;
      (void)frexp(h, &e);
      e += dk;
      emax = e;
      emin = e;
      // |dk| <= 2^21: the scaled value saturates to 0 or Inf well inside int range
      if (e <= 1024) inexact = ldexp(ldexp(h, dk), -dk) != h;
    }
  }
  __shared__ int smax[4], smin[4], sinx[4];
  for (int off = 32; off > 0; off >>= 1) {
    emax = max(emax, __shfl_xor(emax, off, 64));
    emin = min(emin, __shfl_xor(emin, off, 64));
    inexact |= __shfl_xor(inexact, off, 64);
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smax[w] = emax;
    smin[w] = emin;
    sinx[w] = inexact;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    emax = max(max(smax[0], smax[1]), max(smax[2], smax[3]));
    emin = min(min(smin[0], smin[1]), min(smin[2], smin[3]));
    inexact = sinx[0] | sinx[1] | sinx[2] | sinx[3];
    if (emax != INT_MIN) atomicMax(&rng[0], emax);
    if (emin != INT_MAX) atomicMin(&rng[1], emin);
    if (inexact) atomicOr(&rng[2], 1);
  }
}

// H(i, j) <- 2^(k_i - k_j) H(i, j) for i <= j + 1 (entries below stay as they are).
__global__ __launch_bounds__(256) void k_grade(int n, double* H, int ldH,
                                               const int* __restrict__ k) {
  const int j = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n || i > j + 1) return;
  double* h = H + i + (int64_t)j * ldH;
  *h = ldexp(*h, k[i] - k[j]);
}

// Largest order: the generators launch one grid row per column.
constexpr int kGenMaxN = 65535;

}  // namespace
}  // namespace hsrq

using namespace hsrq;

extern "C" {

int hsrq_gaussian_hessenberg_device(hsrq_context* ctx, int n, uint64_t seed, double* dH,
                                    int ldH) {
  if (!ctx || n <= 0 || !dH || ldH < n) return HSRQ_EINVAL;
  if (n > kGenMaxN) return HSRQ_EUNSUPPORTED;
  HIP_TRY(hipSetDevice(ctx->device));
  const dim3 grid((n + 255) / 256, n);
  hipLaunchKernelGGL(k_gaussian_hessenberg, grid, dim3(256), 0, ctx->stream, n, seed, dH, ldH);
  HIP_TRY(hipGetLastError());
  return HSRQ_OK;
}

int hsrq_constructed_hessenberg_device(hsrq_context* ctx, int n, const double* lam,
                                       const double* nu, const double* d, const double* u,
                                       double* dH, int ldH) {
  if (!ctx || n <= 0 || !lam || !d || !dH || ldH < n) return HSRQ_EINVAL;
  if (n > 1 && (!nu || !u)) return HSRQ_EINVAL;
  if (n > kGenMaxN) return HSRQ_EUNSUPPORTED;
  for (int i = 0; i < n; ++i)
    if (!(d[i] != 0.0)) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<double> prm(4 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) {
    prm[i] = lam[i];
    prm[2 * (size_t)n + i] = d[i];
  }
  for (int i = 0; i + 1 < n; ++i) {
    prm[(size_t)n + i] = nu[i];
    prm[3 * (size_t)n + i] = u[i];
  }
  double* dp = nullptr;
  HIP_TRY(hipMalloc(&dp, prm.size() * sizeof(double)));
  hipError_t e = hipMemcpyAsync(dp, prm.data(), prm.size() * sizeof(double),
                                hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_constructed, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, n, dp, dH,
                       ldH);
    e = hipGetLastError();
  }
  // the parameters must outlive the kernel: wait before freeing them
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = es;
  (void)hipFree(dp);
  HIP_TRY(e);
  return HSRQ_OK;
}

int hsrq_grade_device(hsrq_context* ctx, int n, double* dH, int ldH, const int* k, int exact,
                      int* emax_out, int* emin_out) {
  if (!ctx || n <= 0 || !dH || ldH < n || !k) return HSRQ_EINVAL;
  if (n > kGenMaxN) return HSRQ_EUNSUPPORTED;
  for (int i = 0; i < n; ++i)
    if (k[i] < -(1 << 20) || k[i] > (1 << 20)) return HSRQ_EINVAL;
  HIP_TRY(hipSetDevice(ctx->device));
  std::vector<int> hk((size_t)n + 3);
  for (int i = 0; i < n; ++i) hk[i] = k[i];
  hk[n] = INT_MIN;
  hk[n + 1] = INT_MAX;
  hk[n + 2] = 0;
  int* dk = nullptr;
  HIP_TRY(hipMalloc(&dk, hk.size() * sizeof(int)));
  const dim3 grid((n + 255) / 256, n);
  int rng[3] = {INT_MIN, INT_MAX, 0};
  hipError_t e = hipMemcpyAsync(dk, hk.data(), hk.size() * sizeof(int), hipMemcpyHostToDevice,
                                ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_grade_range, grid, dim3(256), 0, ctx->stream, n, dH, ldH, dk, dk + n);
    e = hipGetLastError();
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(rng, dk + n, 3 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  // overflow: f 2^E > DBL_MAX for E > 1024
  const bool over = rng[0] != INT_MIN && rng[0] > 1024;
  const bool under = rng[2] != 0;
  int rc = HSRQ_OK;
  if (e == hipSuccess) {
    if (emax_out) *emax_out = rng[0];
    if (emin_out) *emin_out = rng[1];
    if (over || (exact && under)) {
      rc = HSRQ_ERANGE;
    } else {
      hipLaunchKernelGGL(k_grade, grid, dim3(256), 0, ctx->stream, n, dH, ldH, dk);
      e = hipGetLastError();
      const hipError_t es = hipStreamSynchronize(ctx->stream);
      if (e == hipSuccess) e = es;
    }
  }
  (void)hipFree(dk);
  HIP_TRY(e);
  return rc;
}

}  // extern "C"
