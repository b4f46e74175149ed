// hsrq_cplx.hip -- inverse iteration with complex shifts sigma = a + i b on a
// real upper Hessenberg H: the eigenvectors of complex-conjugate eigenpairs,
// which the real-shift path (and the reference's driver, main.c:95-97) does
// not produce.  The reference ships the robust complex primitives but no
// caller (robust.c:639-671 robust_cmplx_div, :677-767 solve_1x1_cmplx_system).
//
// One workgroup per shift, the whole vector on chip: for n <= 512, 512
// threads with one row each in registers; for 512 < n <= 8192, 1024 threads
// (16 wavefronts: more latency hiding for the loads of H) where thread t
// holds the running column of rows t + 1024 q, q < RPT <= 8, in registers
// and the right-hand side y in LDS -- or, up to n = 2048 with more shifts
// than CUs, 512 threads with RPT <= 4 rows each in registers (two
// workgroups per CU).  Above n = 8192 k_csolve_big keeps the vector in
// global memory.  The sweep runs right to left over the columns of H - sigma I:
//   * the running column v (column st after the rotations so far) and the
//     next column h = H(:, st-1) (- sigma on the diagonal) are combined by the
//     unitary M = [[c, -conj(s)], [s, c]] that zeroes row st of the new column
//     st-1:  c = |v_st| / d, s = -(hsub / d) conj(ph), ph = v_st / |v_st|,
//     d = sqrt(|v_st|^2 + hsub^2), so R(st,st) = ph d and
//       r_i = c v_i - conj(s) h_i  (column st of R),
//       v_i <- c h_i + s v_i       (the next running column);
//   * the back substitution with R is fused into the sweep, column-oriented:
//     y_st <- y_st / R(st,st) (pivot raised to smin if |R| < smin, as
//     robust_cmplx_div; the division protected by a power of two), then
//     y_i <- y_i - y_st r_i, with running bounds deciding whether a
//     power-of-two protection of the update can be needed (then the exact
//     bound is reduced over the workgroup);
//   * x = U_{n-1} ... U_1 z from the stored (c, s): a sequential complex
//     recurrence over LDS-staged chunks.
// Every step is one barrier: after its row updates, the wavefront owning row
// st-1 computes the next step's scalar chain (rotation, quotient, bounds)
// from that row and publishes it in LDS; every thread applies it next step.
// Output: column 2k of vr the real part, 2k+1 the imaginary part of x_k,
// ||x_k||_2 = 1 with its first entry of largest modulus real and positive;
// status bit 1 converged (dlaein's growth test, as the real path), 2 a pivot
// was raised to smin.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <algorithm>
#include <climits>
#include <vector>

#include "hsrq_device.h"
#include "hsrq_host.h"

namespace hsrq {
namespace {

constexpr int kCThreads = 512;      // n <= 512 and the global-memory kernel
constexpr int kCWaves = kCThreads / 64;
constexpr int kCBigThreads = 1024;  // 512 < n <= 8192, y in LDS
constexpr int kCMaxRpt = 8;
constexpr int kCMaxN = kCBigThreads * kCMaxRpt;

struct CArgs {
  int n;
  const double* H;
  int64_t ldH;
  const double* wr;
  const double* wi;
  double* vr;
  int64_t ldvr;
  int* status;
  double* work;        // 3 n per shift: c, Re s, Im s of rotation st at index st
  const double* rs;    // row sums ||H(i,:)||_1
  int its;             // start vector: -1 eps3 (1, ..., 1); >= 0 the its-th restart vector
  const int* idx;      // shift of each workgroup (nullptr: workgroup k solves shift k)
};

// Entry `row` of the start vector (inverse_iteration.c:19-57, dlaein's choice):
// eps3 everywhere, or for restart its: eps3 / (sqrt(n) + 1) with eps3 in row 0
// and eps3 sqrt(n) subtracted in row n-1-its.
__device__ __forceinline__ double start_y(int row, int n, int its, double eps3) {
  if (its < 0) return eps3;
  const double sq = sqrt((double)n);
  double y = (row == 0) ? eps3 : eps3 / (sq + 1.0);
  if (row == n - 1 - its) y -= eps3 * sq;
  return y;
}
// A bound of |y_i| for that start vector.
__device__ __forceinline__ double start_bound(int n, int its, double eps3) {
  return its < 0 ? eps3 : eps3 * (1.0 + sqrt((double)n));
}

// sqrt(a^2 + b^2 + c^2) with power-of-two scaling.
__device__ __forceinline__ double norm3(double a, double b, double c) {
  const double m = fmax(fabs(a), fmax(fabs(b), fabs(c)));
  if (m == 0.0 || !isfinite(m)) return m;
  const int e = ilogb(m);
  const double as = ldexp(a, -e), bs = ldexp(b, -e), cs = ldexp(c, -e);
  return ldexp(sqrt((as * as + bs * bs) + cs * cs), e);
}

// Workgroup reductions over W wavefronts (every thread of the workgroup
// calls them).
template <int W = kCWaves>
__device__ __forceinline__ double block_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < W; ++w) r = fmax(r, red[w]);
  return r;
}

template <int W = kCWaves>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int w = 1; w < W; ++w) r = r + red[w];
  return r;
}

template <int W = kCWaves>
__device__ __forceinline__ int block_min_int(int v, int* redi) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) redi[threadIdx.x >> 6] = v;
  __syncthreads();
  int r = redi[0];
#pragma unroll
  for (int w = 1; w < W; ++w) r = min(r, redi[w]);
  return r;
}

// NT threads per workgroup, RPT rows per thread; YL: the right-hand side y
// lives in LDS instead of registers (NT = 1024: the 128 registers a thread
// has at 4 wavefronts per SIMD hold the running column and the ring of H
// columns only).
constexpr int kCChunk = 512;   // rows per back-transform chunk
template <int RPT, bool YL, int NT>
__global__ __launch_bounds__(NT) void k_csolve(CArgs A) {
  static_assert(YL || NT == kCChunk, "register-resident y: 512 threads");
  constexpr int NW = NT / 64;
  __shared__ double ylr[YL ? RPT * NT : 1], yli[YL ? RPT * NT : 1];
  struct Rec {   // one pivot step: rotation, quotient, bounds, protections
    double c, sr, si, qr, qi, by;
    int e1, slow, info;
  };
  __shared__ Rec rec[2];                 // by step parity
  __shared__ double red[NW];
  __shared__ int redi[NW];
  __shared__ double stg[5][kCChunk];   // back-transform chunk: Re z, Im z, c, Re s, Im s
  __shared__ double xo[2][YL ? 1 : kCChunk + 1];

  const int k = A.idx ? A.idx[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x;
  const int n = A.n;
  const double* __restrict__ H = A.H;
  const int64_t ldH = A.ldH;
  const double sa = A.wr[k], sb = A.wi[k];
  double* __restrict__ C = A.work + (int64_t)k * 3 * n;
  double* __restrict__ SR = C + n;
  double* __restrict__ SI = C + 2 * n;

  double hinf = 0.0;
  for (int i = tid; i < n; i += NT) hinf = fmax(hinf, A.rs[i]);
  hinf = block_max<NW>(hinf, red);

  const double eps = 0x1p-53;
  const double smlnum = fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps));
  const double asig = norm3(sa, sb, 0.0);
  const double smin = fmax(eps * asig, smlnum);
  const double eps3 = hinf * 0x1p-53;
  const double slack = 1.0 + 0x1p-40;
  const double hm = hinf + asig;
  // |h_i| <= hm, and a rotation keeps |v_i|^2 <= |h_i|^2 + |v_i|^2, so after j
  // steps |v_i| <= hm sqrt(j + 1) <= hm sqrt(n); |r_i| <= sqrt(|v_i|^2 + |h_i|^2),
  // and N(z) = |Re z| + |Im z| <= sqrt(2) |z|: br bounds N(r_i) in every step
  const double br = M_SQRT2 * hm * sqrt((double)n + 1.0) * slack;

  // H columns ride a ring of D register slots: the column of step st is
  // loaded D steps before it is used (the step loop is unrolled by D so that
  // every slot index is a compile-time constant).
  // (depths measured on MI355X; at 8 rows per thread D = 1 keeps the
  // spills of the 128-register budget of 1024 threads to a few)
  constexpr int D = NT > kCChunk ? (RPT == 1 ? 4 : (RPT <= 4 ? 2 : 1)) : (RPT <= 2 ? 8 : 2);
  double vr[RPT], vi[RPT], yrr[YL ? 1 : RPT], yir[YL ? 1 : RPT];
#define YR(q) (*(YL ? &ylr[tid + NT * (q)] : &yrr[YL ? 0 : (q)]))
#define YI(q) (*(YL ? &yli[tid + NT * (q)] : &yir[YL ? 0 : (q)]))
  double ring[D][RPT], rsub[D];
  auto load_col = [&](int st, double (&h)[RPT], double& hs) {
    // column st-1 (rows 0 .. st-1 are used) and its subdiagonal entry
    // H(st, st-1).  The loads are unconditional, from clamped in-bounds
    // addresses, and their values are not touched until the step that uses
    // them (which reads only rows <= st-1): no branch or select makes the
    // compiler wait for them early.
    const double* hcol = H + (int64_t)max(st - 1, 0) * ldH;
#pragma unroll
    for (int q = 0; q < RPT; ++q) h[q] = hcol[min(tid + NT * q, n - 1)];
    const int sc = min(max(st, 1), n - 1);
    hs = H[n >= 2 ? sc + (int64_t)(sc - 1) * ldH : 0];
  };
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int row = tid + NT * q;
    const bool in = row < n;
    vr[q] = in ? H[row + (int64_t)(n - 1) * ldH] : 0.0;
    vi[q] = 0.0;
    YR(q) = in ? start_y(row, n, A.its, eps3) : 0.0;
    YI(q) = 0.0;
    if (row == n - 1) {
      vr[q] -= sa;
      vi[q] = -sb;
    }
  }
#pragma unroll
  for (int u = 0; u < D; ++u) load_col(n - 1 - u, ring[u], rsub[u]);
  __syncthreads();

  // The scalar chain of pivot step st -- rotation from (v_st, H(st, st-1)),
  // division of y_st by R(st,st), bounds -- is computed once, by the
  // wavefront that owns row st (right after it updated that row in step
  // st+1), and published with the barrier; every thread then applies it.
  const int wave = tid >> 6, lane = tid & 63;
  auto chain = [&](int st, double psr, double psi, double ysr, double ysi, double hsub,
                   double by) {
    double c = 1.0, sr = 0.0, si = 0.0, pr = 1.0, pi = 0.0, av, d;
    {   // |v| and d = sqrt(|v|^2 + hsub^2) under one power-of-two scaling
      const double h = st >= 1 ? hsub : 0.0;
      const double mm = fmax(fabs(psr), fmax(fabs(psi), fabs(h)));
      if (mm == 0.0 || !isfinite(mm)) {
        av = fmax(fabs(psr), fabs(psi));
        d = mm;
      } else {
        const int e = ilogb(mm);
        const double a = ldexp(psr, -e), b = ldexp(psi, -e), g = ldexp(h, -e);
        const double s2v = a * a + b * b;
        av = ldexp(sqrt(s2v), e);
        d = ldexp(sqrt(s2v + g * g), e);
      }
    }
    const bool tiny = d < smin;          // pivot raised to smin (robust_cmplx_div)
    const double deff = tiny ? smin : d;
    const double id = 1.0 / deff;
    if (av > 0.0) {
      const double ia = 1.0 / av;
      pr = psr * ia;
      pi = psi * ia;
    }
    if (st >= 1 && d > 0.0) {
      const double idr = tiny ? 1.0 / d : id;
      c = av * idr;
      const double t = hsub * idr;
      sr = -(t * pr);
      si = t * pi;
    }
    // y_st / (ph d), the division protected by 2^e1
    const double qpr = tiny ? 1.0 : pr, qpi = tiny ? 0.0 : pi;
    double wr = ysr * qpr + ysi * qpi;
    double wi = ysi * qpr - ysr * qpi;
    const int e1 = pow2_exp(protect_real_division(fabs(wr) + fabs(wi), deff));
    if (e1 != 0) {
      wr = ldexp(wr, e1);
      wi = ldexp(wi, e1);
      by = ldexp(by, e1);
    }
    const double qr = wr * id, qi = wi * id;
    const double byn = (by + (fabs(qr) + fabs(qi)) * br) * slack;
    if (lane == 0) {
      Rec& o = rec[st & 1];
      o.c = c;
      o.sr = sr;
      o.si = si;
      o.qr = qr;
      o.qi = qi;
      o.by = byn;
      o.e1 = e1;
      o.slow = !(byn <= kOmega);
      o.info = tiny ? 1 : 0;
      if (st >= 1) {
        C[st] = c;
        SR[st] = sr;
        SI[st] = si;
      }
    }
  };
  // value of register slot q_own of the thread holding row `row`, in every lane
  auto owner_value = [&](const double (&a)[RPT], int row) -> double {
    const int qo = row / NT;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < RPT; ++q) v = (q == qo) ? a[q] : v;
    return readlane(v, row & 63);
  };

  // y of row `row` in every lane: from LDS, or from the owner's register
  auto owner_y = [&](const double (&a)[YL ? 1 : RPT], const double* l, int row) -> double {
    if constexpr (YL) {
      return l[row];
    } else {
      return owner_value(a, row);
    }
  };

  int E = 0, info = 0;
  double by = start_bound(n, A.its, eps3);
  if (wave == ((n - 1) % NT) >> 6)
    chain(n - 1, owner_value(vr, n - 1), owner_value(vi, n - 1), owner_y(yrr, ylr, n - 1),
          owner_y(yir, yli, n - 1), rsub[0], by);
  __syncthreads();

  auto step = [&](int st, const double (&hc)[RPT]) {
    const Rec R = rec[st & 1];
    const double c = R.c, sr = R.sr, si = R.si;
    double qr = R.qr, qi = R.qi;
    E += R.e1;
    info |= R.info;
    if (R.e1 != 0) {
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        YR(q) = ldexp(YR(q), R.e1);
        YI(q) = ldexp(YI(q), R.e1);
      }
    }
    by = R.by;
    if (R.slow) {   // uniform: exact bound, scaled by 2^-kk
      const double nq = fabs(qr) + fabs(qi);
      const int kk = nq > 0.0 ? ilogb(nq) : 0;
      const double nqs = ldexp(nq, -kk);
      double a = 0.0;
#pragma unroll
      for (int q = 0; q < RPT; ++q) {
        const int row = tid + NT * q;
        if (row <= st - 1) {
          const double hr = (row == st - 1) ? hc[q] - sa : hc[q];
          const double hi = (row == st - 1) ? -sb : 0.0;
          const double rr = c * vr[q] - (sr * hr + si * hi);
          const double ri = c * vi[q] - (sr * hi - si * hr);
          a = fmax(a, nqs * (fabs(rr) + fabs(ri)) + ldexp(fabs(YR(q)) + fabs(YI(q)), -kk));
        }
      }
      a = block_max<NW>(a, red);
      int e2 = 0;
      if (a > ldexp(kOmega, -kk)) e2 = -(ilogb(a) + kk) - 1;
      if (e2 != 0) {
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
          YR(q) = ldexp(YR(q), e2);
          YI(q) = ldexp(YI(q), e2);
        }
        qr = ldexp(qr, e2);
        qi = ldexp(qi, e2);
        E += e2;
      }
      by = ldexp(a, kk + e2) * slack;
    }
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int row = tid + NT * q;
      if (row <= st - 1) {
        const double hr = (row == st - 1) ? hc[q] - sa : hc[q];
        const double hi = (row == st - 1) ? -sb : 0.0;
        const double rr = fma(c, vr[q], -fma(sr, hr, si * hi));
        const double ri = fma(c, vi[q], fma(si, hr, -(sr * hi)));
        const double nvr = fma(c, hr, fma(sr, vr[q], -(si * vi[q])));
        const double nvi = fma(c, hi, fma(sr, vi[q], si * vr[q]));
        YR(q) = fma(-qr, rr, fma(qi, ri, YR(q)));
        YI(q) = fma(-qr, ri, fma(-qi, rr, YI(q)));
        vr[q] = nvr;
        vi[q] = nvi;
      } else if (row == st) {
        YR(q) = qr;
        YI(q) = qi;
      }
    }
  };
  for (int st0 = n - 1; st0 >= 1; st0 -= D) {
#pragma unroll
    for (int u = 0; u < D; ++u) {
      const int st = st0 - u;
      if (st >= 1) {
        step(st, ring[u]);
        load_col(st - D, ring[u], rsub[u]);   // the column of step st - D
        if (wave == ((st - 1) % NT) >> 6)
          chain(st - 1, owner_value(vr, st - 1), owner_value(vi, st - 1),
                owner_y(yrr, ylr, st - 1), owner_y(yir, yli, st - 1), rsub[(u + 1) % D], by);
        __syncthreads();
      }
    }
  }
  {   // apply step 0: y_0 = y_0 / R(0,0)
    const Rec R = rec[0];
    E += R.e1;
    info |= R.info;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int row = tid + NT * q;
      if (row == 0) {
        YR(q) = R.qr;
        YI(q) = R.qi;
      } else if (R.e1 != 0) {
        YR(q) = ldexp(YR(q), R.e1);
        YI(q) = ldexp(YI(q), R.e1);
      }
    }
  }
  // the rotations written by thread 0 are read by all threads below
  __threadfence_block();
  __syncthreads();

  // x = U_{n-1} ... U_1 z by one thread, a sequential complex recurrence over
  // chunks of 512 rows whose rotations are staged in LDS; x of row j - 1 is
  // produced at element j.
  if constexpr (YL) {
    // z in LDS: x overwrites it in place (element j reads z_j, writes
    // x_{j-1}, whose z was read at element j - 1)
    double tr = 0.0, ti = 0.0;
    for (int g0 = 0; g0 < n; g0 += kCChunk) {
      const int cnt = min(kCChunk, n - g0);
      if (tid < cnt) {
        const int row = g0 + tid;
        stg[2][tid] = row >= 1 ? C[row] : 1.0;
        stg[3][tid] = row >= 1 ? SR[row] : 0.0;
        stg[4][tid] = row >= 1 ? SI[row] : 0.0;
      }
      __syncthreads();
      if (tid == 0) {
        for (int j = 0; j < cnt; ++j) {
          const int row = g0 + j;
          const double zr = ylr[row], zi = yli[row];
          if (row == 0) {
            tr = zr;
            ti = zi;
          } else {
            const double cc = stg[2][j], s_r = stg[3][j], s_i = stg[4][j];
            ylr[row - 1] = cc * tr - (s_r * zr + s_i * zi);
            yli[row - 1] = cc * ti - (s_r * zi - s_i * zr);
            const double ntr = (s_r * tr - s_i * ti) + cc * zr;
            const double nti = (s_r * ti + s_i * tr) + cc * zi;
            tr = ntr;
            ti = nti;
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      ylr[n - 1] = tr;
      yli[n - 1] = ti;
    }
    __syncthreads();
  } else {
    // z in registers: chunk q is slot q of every thread (NT = 512 rows),
    // staged with its rotations; xo[j] receives x of row 512 q + j - 1
    double tr = 0.0, ti = 0.0;   // the entering value (thread 0)
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int g0 = NT * q;
      if (g0 < n) {
        const int row = g0 + tid;
        if (row < n) {
          stg[0][tid] = YR(q);
          stg[1][tid] = YI(q);
          stg[2][tid] = row >= 1 ? C[row] : 1.0;
          stg[3][tid] = row >= 1 ? SR[row] : 0.0;
          stg[4][tid] = row >= 1 ? SI[row] : 0.0;
        }
        __syncthreads();
        if (tid == 0) {
          const int cnt = min(NT, n - g0);
          for (int j = 0; j < cnt; ++j) {
            const double zr = stg[0][j], zi = stg[1][j];
            if (g0 + j == 0) {
              tr = zr;
              ti = zi;
            } else {
              const double cc = stg[2][j], s_r = stg[3][j], s_i = stg[4][j];
              xo[0][j] = cc * tr - (s_r * zr + s_i * zi);
              xo[1][j] = cc * ti - (s_r * zi - s_i * zr);
              const double ntr = (s_r * tr - s_i * ti) + cc * zr;
              const double nti = (s_r * ti + s_i * tr) + cc * zi;
              tr = ntr;
              ti = nti;
            }
          }
        }
        __syncthreads();
        if (tid + 1 < NT && row + 1 < n) {   // produced at element tid + 1
          YR(q) = xo[0][tid + 1];
          YI(q) = xo[1][tid + 1];
        }
        if (q > 0 && tid == NT - 1) {        // last row of the previous chunk
          YR(q > 0 ? q - 1 : 0) = xo[0][0];
          YI(q > 0 ? q - 1 : 0) = xo[1][0];
        }
        __syncthreads();
      }
    }
    if (tid == 0) {
      xo[0][0] = tr;
      xo[1][0] = ti;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RPT; ++q)
      if (tid + NT * q == n - 1) {
        YR(q) = xo[0][0];
        YI(q) = xo[1][0];
      }
  }

  // ||x||_2, the first entry of largest modulus, normalisation
  double mx = 0.0;
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (tid + NT * q < n) mx = fmax(mx, fmax(fabs(YR(q)), fabs(YI(q))));
  mx = block_max<NW>(mx, red);
  double nrm = mx;
  int p = 0;
  if (mx > 0.0 && isfinite(mx)) {
    const int e = ilogb(mx);
    double sum = 0.0, best = -1.0;
    int bi = INT_MAX;
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int row = tid + NT * q;
      if (row < n) {
        const double a = ldexp(YR(q), -e), b = ldexp(YI(q), -e);
        const double t = a * a + b * b;
        sum += t;
        if (t > best) {
          best = t;
          bi = row;
        }
      }
    }
    sum = block_sum<NW>(sum, red);
    const double bmax = block_max<NW>(best, red);
    p = block_min_int<NW>(best == bmax ? bi : INT_MAX, redi);
    nrm = ldexp(sqrt(sum), e);
  }
  const bool conv = isfinite(nrm) && nrm > ldexp(0.1 / sqrt((double)n), E);
  __syncthreads();
#pragma unroll
  for (int q = 0; q < RPT; ++q)
    if (tid + NT * q == p) {
      xo[0][0] = YR(q);
      xo[1][0] = YI(q);
    }
  __syncthreads();
  double pr = 1.0, pi = 0.0;
  const bool scale = nrm > 0.0 && isfinite(nrm);
  if (scale) {
    const double ap = norm3(xo[0][0], xo[1][0], 0.0);
    pr = xo[0][0] / ap;
    pi = xo[1][0] / ap;
  }
  double* outr = A.vr + (int64_t)(2 * k) * A.ldvr;
  double* outi = A.vr + (int64_t)(2 * k + 1) * A.ldvr;
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    const int row = tid + NT * q;
    if (row < n) {
      const double a = YR(q), b = YI(q);
      outr[row] = scale ? (a * pr + b * pi) / nrm : a;
      outi[row] = scale ? (b * pr - a * pi) / nrm : b;
    }
  }
  if (tid == 0) A.status[k] = (conv ? HSRQ_CONVERGED : 0) | (info ? HSRQ_PERTURBED : 0);
#undef YR
#undef YI
}

// Orders above 8192: the same sweep with the vector in global memory (per
// shift 4 n doubles: Re v, Im v, Re y, Im y), rows strided over the threads
// in runtime loops.  Every step reads and writes the active rows once (L2 /
// HBM traffic instead of registers), so it is slower per row; results follow
// the same equations and operation order as k_csolve.
__global__ __launch_bounds__(kCThreads) void k_csolve_big(CArgs A, double* __restrict__ state) {
  struct Rec {
    double c, sr, si, qr, qi, by;
    int e1, slow, info;
  };
  __shared__ Rec rec[2];
  __shared__ double red[kCWaves];
  __shared__ int redi[kCWaves];
  __shared__ double stg[5][kCThreads];

  const int k = A.idx ? A.idx[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int n = A.n;
  const double* __restrict__ H = A.H;
  const int64_t ldH = A.ldH;
  const double sa = A.wr[k], sb = A.wi[k];
  double* __restrict__ C = A.work + (int64_t)k * 3 * n;
  double* __restrict__ SR = C + n;
  double* __restrict__ SI = C + 2 * n;
  double* __restrict__ VR = state + (int64_t)k * 4 * n;
  double* __restrict__ VI = VR + n;
  double* __restrict__ YR = VI + n;
  double* __restrict__ YI = YR + n;

  double hinf = 0.0;
  for (int i = tid; i < n; i += kCThreads) hinf = fmax(hinf, A.rs[i]);
  hinf = block_max(hinf, red);
  const double eps = 0x1p-53;
  const double smlnum = fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps));
  const double asig = norm3(sa, sb, 0.0);
  const double smin = fmax(eps * asig, smlnum);
  const double eps3 = hinf * 0x1p-53;
  const double slack = 1.0 + 0x1p-40;
  const double hm = hinf + asig;
  const double br = M_SQRT2 * hm * sqrt((double)n + 1.0) * slack;

  for (int row = tid; row < n; row += kCThreads) {
    double v = H[row + (int64_t)(n - 1) * ldH];
    if (row == n - 1) v -= sa;
    VR[row] = v;
    VI[row] = (row == n - 1) ? -sb : 0.0;
    YR[row] = start_y(row, n, A.its, eps3);
    YI[row] = 0.0;
  }

  auto chain = [&](int st, double psr, double psi, double ysr, double ysi, double hsub,
                   double by) {
    double c = 1.0, sr = 0.0, si = 0.0, pr = 1.0, pi = 0.0, av, d;
    {
      const double h = st >= 1 ? hsub : 0.0;
      const double mm = fmax(fabs(psr), fmax(fabs(psi), fabs(h)));
      if (mm == 0.0 || !isfinite(mm)) {
        av = fmax(fabs(psr), fabs(psi));
        d = mm;
      } else {
        const int e = ilogb(mm);
        const double a = ldexp(psr, -e), b = ldexp(psi, -e), g = ldexp(h, -e);
        const double s2v = a * a + b * b;
        av = ldexp(sqrt(s2v), e);
        d = ldexp(sqrt(s2v + g * g), e);
      }
    }
    const bool tiny = d < smin;
    const double deff = tiny ? smin : d;
    const double id = 1.0 / deff;
    if (av > 0.0) {
      const double ia = 1.0 / av;
      pr = psr * ia;
      pi = psi * ia;
    }
    if (st >= 1 && d > 0.0) {
      const double idr = tiny ? 1.0 / d : id;
      c = av * idr;
      const double t = hsub * idr;
      sr = -(t * pr);
      si = t * pi;
    }
    const double qpr = tiny ? 1.0 : pr, qpi = tiny ? 0.0 : pi;
    double wr = ysr * qpr + ysi * qpi;
    double wi = ysi * qpr - ysr * qpi;
    const int e1 = pow2_exp(protect_real_division(fabs(wr) + fabs(wi), deff));
    if (e1 != 0) {
      wr = ldexp(wr, e1);
      wi = ldexp(wi, e1);
      by = ldexp(by, e1);
    }
    const double qr = wr * id, qi = wi * id;
    const double byn = (by + (fabs(qr) + fabs(qi)) * br) * slack;
    if (lane == 0) {
      Rec& o = rec[st & 1];
      o.c = c;
      o.sr = sr;
      o.si = si;
      o.qr = qr;
      o.qi = qi;
      o.by = byn;
      o.e1 = e1;
      o.slow = !(byn <= kOmega);
      o.info = tiny ? 1 : 0;
      if (st >= 1) {
        C[st] = c;
        SR[st] = sr;
        SI[st] = si;
      }
    }
  };
  auto scale_y = [&](int e) {
    for (int row = tid; row < n; row += kCThreads) {
      YR[row] = ldexp(YR[row], e);
      YI[row] = ldexp(YI[row], e);
    }
  };

  int E = 0, info = 0;
  double by = start_bound(n, A.its, eps3);
  if (wave == 0) {
    // row n-1 holds its initial values, known to every thread
    const double hsub = n >= 2 ? H[(n - 1) + (int64_t)(n - 2) * ldH] : 0.0;
    chain(n - 1, H[(n - 1) + (int64_t)(n - 1) * ldH] - sa, -sb, start_y(n - 1, n, A.its, eps3),
          0.0, hsub, by);
  }
  __syncthreads();

  for (int st = n - 1; st >= 1; --st) {
    const Rec R = rec[st & 1];
    const double c = R.c, sr = R.sr, si = R.si;
    double qr = R.qr, qi = R.qi;
    E += R.e1;
    info |= R.info;
    if (R.e1 != 0) scale_y(R.e1);   // each thread its own rows: no barrier needed
    by = R.by;
    const double* __restrict__ hc = H + (int64_t)(st - 1) * ldH;
    if (R.slow) {
      const double nq = fabs(qr) + fabs(qi);
      const int kk = nq > 0.0 ? ilogb(nq) : 0;
      const double nqs = ldexp(nq, -kk);
      double a = 0.0;
      for (int row = tid; row <= st - 1; row += kCThreads) {
        const double hr = (row == st - 1) ? hc[row] - sa : hc[row];
        const double hi = (row == st - 1) ? -sb : 0.0;
        const double rr = c * VR[row] - (sr * hr + si * hi);
        const double ri = c * VI[row] - (sr * hi - si * hr);
        a = fmax(a, nqs * (fabs(rr) + fabs(ri)) + ldexp(fabs(YR[row]) + fabs(YI[row]), -kk));
      }
      a = block_max(a, red);
      int e2 = 0;
      if (a > ldexp(kOmega, -kk)) e2 = -(ilogb(a) + kk) - 1;
      if (e2 != 0) {
        scale_y(e2);
        qr = ldexp(qr, e2);
        qi = ldexp(qi, e2);
        E += e2;
      }
      by = ldexp(a, kk + e2) * slack;
    }
    double o0 = 0.0, o1 = 0.0, o2 = 0.0, o3 = 0.0;   // row st-1 after the update
    for (int row = tid; row <= st - 1; row += kCThreads) {
      const double hr = (row == st - 1) ? hc[row] - sa : hc[row];
      const double hi = (row == st - 1) ? -sb : 0.0;
      const double vr = VR[row], vi = VI[row];
      const double rr = fma(c, vr, -fma(sr, hr, si * hi));
      const double ri = fma(c, vi, fma(si, hr, -(sr * hi)));
      const double nvr = fma(c, hr, fma(sr, vr, -(si * vi)));
      const double nvi = fma(c, hi, fma(sr, vi, si * vr));
      const double nyr = fma(-qr, rr, fma(qi, ri, YR[row]));
      const double nyi = fma(-qr, ri, fma(-qi, rr, YI[row]));
      VR[row] = nvr;
      VI[row] = nvi;
      YR[row] = nyr;
      YI[row] = nyi;
      if (row == st - 1) {
        o0 = nvr;
        o1 = nvi;
        o2 = nyr;
        o3 = nyi;
      }
    }
    if (tid == st % kCThreads) {
      YR[st] = qr;
      YI[st] = qi;
    }
    if (wave == ((st - 1) % kCThreads) >> 6) {
      const int ol = (st - 1) & 63;
      const double hsub = st >= 2 ? H[(st - 1) + (int64_t)(st - 2) * ldH] : 0.0;
      chain(st - 1, readlane(o0, ol), readlane(o1, ol), readlane(o2, ol), readlane(o3, ol), hsub,
            by);
    }
    __syncthreads();
  }
  {
    const Rec R = rec[0];
    E += R.e1;
    info |= R.info;
    for (int row = tid; row < n; row += kCThreads) {
      if (row == 0) {
        YR[0] = R.qr;
        YI[0] = R.qi;
      } else if (R.e1 != 0) {
        YR[row] = ldexp(YR[row], R.e1);
        YI[row] = ldexp(YI[row], R.e1);
      }
    }
  }
  // every thread reads rows and rotations written by others below
  __threadfence_block();
  __syncthreads();

  // x = U_{n-1} ... U_1 z in place of z: chunks staged in LDS, the recurrence
  // by thread 0 (x_{i-1} overwrites z_{i-1}, already staged)
  double tr = 0.0, ti = 0.0;
  for (int g0 = 0; g0 < n; g0 += kCThreads) {
    const int row = g0 + tid;
    if (row < n) {
      stg[0][tid] = YR[row];
      stg[1][tid] = YI[row];
      stg[2][tid] = row >= 1 ? C[row] : 1.0;
      stg[3][tid] = row >= 1 ? SR[row] : 0.0;
      stg[4][tid] = row >= 1 ? SI[row] : 0.0;
    }
    __syncthreads();
    if (tid == 0) {
      const int cnt = min(kCThreads, n - g0);
      for (int j = 0; j < cnt; ++j) {
        const double zr = stg[0][j], zi = stg[1][j];
        if (g0 + j == 0) {
          tr = zr;
          ti = zi;
        } else {
          const double cc = stg[2][j], s_r = stg[3][j], s_i = stg[4][j];
          YR[g0 + j - 1] = cc * tr - (s_r * zr + s_i * zi);
          YI[g0 + j - 1] = cc * ti - (s_r * zi - s_i * zr);
          const double ntr = (s_r * tr - s_i * ti) + cc * zr;
          const double nti = (s_r * ti + s_i * tr) + cc * zi;
          tr = ntr;
          ti = nti;
        }
      }
      if (g0 + cnt == n) {
        YR[n - 1] = tr;
        YI[n - 1] = ti;
      }
    }
    __syncthreads();
  }
  __threadfence_block();
  __syncthreads();

  double mx = 0.0;
  for (int row = tid; row < n; row += kCThreads) mx = fmax(mx, fmax(fabs(YR[row]), fabs(YI[row])));
  mx = block_max(mx, red);
  double nrm = mx;
  int p = 0;
  if (mx > 0.0 && isfinite(mx)) {
    const int e = ilogb(mx);
    double sum = 0.0, best = -1.0;
    int bi = INT_MAX;
    for (int row = tid; row < n; row += kCThreads) {
      const double a = ldexp(YR[row], -e), b = ldexp(YI[row], -e);
      const double t = a * a + b * b;
      sum += t;
      if (t > best) {
        best = t;
        bi = row;
      }
    }
    sum = block_sum(sum, red);
    const double bmax = block_max(best, red);
    p = block_min_int(best == bmax ? bi : INT_MAX, redi);
    nrm = ldexp(sqrt(sum), e);
  }
  const bool conv = isfinite(nrm) && nrm > ldexp(0.1 / sqrt((double)n), E);
  double pr = 1.0, pi = 0.0;
  const bool scale = nrm > 0.0 && isfinite(nrm);
  if (scale) {
    const double xr = YR[p], xi = YI[p];
    const double ap = norm3(xr, xi, 0.0);
    pr = xr / ap;
    pi = xi / ap;
  }
  double* outr = A.vr + (int64_t)(2 * k) * A.ldvr;
  double* outi = A.vr + (int64_t)(2 * k + 1) * A.ldvr;
  for (int row = tid; row < n; row += kCThreads) {
    const double a = YR[row], b = YI[row];
    outr[row] = scale ? (a * pr + b * pi) / nrm : a;
    outi[row] = scale ? (b * pr - a * pi) / nrm : b;
  }
  if (tid == 0) A.status[k] = (conv ? HSRQ_CONVERGED : 0) | (info ? HSRQ_PERTURBED : 0);
}

template <int RPT, bool YL = false, int NT = kCThreads>
hipError_t launch_csolve(const CArgs& a, int m, hipStream_t st) {
  hipLaunchKernelGGL((k_csolve<RPT, YL, NT>), dim3(m), dim3(NT), 0, st, a);
  return hipGetLastError();
}

}  // namespace
}  // namespace hsrq

using namespace hsrq;

extern "C" int hsrq_solve_complex_device(hsrq_context* ctx, int n, const double* dH, int ldH,
                                         int m, const double* dwr, const double* dwi,
                                         double* dvr, int ldvr, int max_restarts, int* status) {
  if (!ctx || n <= 0 || m < 0 || !dH || ldH < n || ldvr < n || max_restarts < 0 ||
      (m > 0 && (!dwr || !dwi || !dvr)))
    return HSRQ_EINVAL;
  if (m == 0) return HSRQ_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  const bool big = n > kCMaxN;
  // rotations (3 n per shift), row sums (n), their partial sums (64 n), the
  // vectors of the large-order kernel (4 n per shift); status and the shift
  // list of a restart pass (m each)
  const size_t nd = 3 * (size_t)n * m + 65 * (size_t)n + (big ? 4 * (size_t)n * m : 0);
  char* buf = nullptr;
  HIP_TRY(hipMalloc(&buf, nd * sizeof(double) + 2 * (size_t)m * sizeof(int)));
  double* work = reinterpret_cast<double*>(buf);
  double* rs = work + 3 * (size_t)n * m;
  double* part = rs + n;
  double* state = part + 64 * (size_t)n;
  int* dst = reinterpret_cast<int*>(buf + nd * sizeof(double));
  int* didx = dst + m;
  CArgs a{n, dH, ldH, dwr, dwi, dvr, ldvr, dst, work, rs, -1, nullptr};
  const int rpt = (n + kCBigThreads - 1) / kCBigThreads;   // rows per thread above 512
  auto pass = [&](int count) -> hipError_t {
    if (big) {
      hipLaunchKernelGGL(k_csolve_big, dim3(count), dim3(kCThreads), 0, ctx->stream, a, state);
      return hipGetLastError();
    }
    if (n <= kCThreads) return launch_csolve<1>(a, count, ctx->stream);
    // Up to 2048 rows the 512-thread register layout fits two workgroups per
    // CU (the 1024-thread one, one): with more shifts than CUs it finishes
    // first, with at most one shift per CU the 1024-thread one does
    // (measured on MI355X: n = 2000, 512 shifts 3.6 vs 5.3 ms; 256 shifts
    // 2.67 vs 2.60 ms; n = 1000, 256 shifts 1.40 vs 1.17 ms).
    if (rpt <= 2 && count > std::max(ctx->cus, 1)) {
      if (n <= 2 * kCThreads) return launch_csolve<2>(a, count, ctx->stream);
      return launch_csolve<4>(a, count, ctx->stream);
    }
    if (rpt <= 1) return launch_csolve<1, true, kCBigThreads>(a, count, ctx->stream);
    if (rpt <= 2) return launch_csolve<2, true, kCBigThreads>(a, count, ctx->stream);
    if (rpt <= 4) return launch_csolve<4, true, kCBigThreads>(a, count, ctx->stream);
    return launch_csolve<8, true, kCBigThreads>(a, count, ctx->stream);
  };
  std::vector<int> st(m, 0), todo;
  hipError_t e = launch_rowsums(n, dH, ldH, part, rs, ctx->stream);
  if (e == hipSuccess) e = pass(m);
  if (e == hipSuccess)
    e = hipMemcpyAsync(st.data(), dst, (size_t)m * sizeof(int), hipMemcpyDeviceToHost,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  // restarts (dlaein): the unconverged shifts again, with the its-th start
  // vector, until they converge or max_restarts passes (at most n) are done
  for (int its = 0; e == hipSuccess && its < std::min(max_restarts, n); ++its) {
    todo.clear();
    for (int k = 0; k < m; ++k)
      if (!(st[k] & HSRQ_CONVERGED)) todo.push_back(k);
    if (todo.empty()) break;
    e = hipMemcpyAsync(didx, todo.data(), todo.size() * sizeof(int), hipMemcpyHostToDevice,
                       ctx->stream);
    a.its = its;
    a.idx = didx;
    if (e == hipSuccess) e = pass((int)todo.size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(st.data(), dst, (size_t)m * sizeof(int), hipMemcpyDeviceToHost,
                         ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  const hipError_t es = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = es;
  (void)hipFree(buf);
  HIP_TRY(e);
  if (status) std::copy(st.begin(), st.end(), status);
  return HSRQ_OK;
}
