// hsrq_device.h -- device-side scalar core of the robust shifted Hessenberg
// solve for gfx950: Givens generation, overflow protection with power-of-two
// scaling factors, and wavefront (64-lane) helpers.
//
// Scaling convention.  Every protected quantity carries an integer exponent
// E <= 0 (occasionally +1, see solve_1x1): stored = true * 2^E.  Factors are
// the reference's protect_* factors (robust.c:144-285) rounded DOWN to a
// power of two, so every rescaling is exact and a protection is triggered in
// exactly the situations the reference's bounds trigger it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hsrq {

// Global-memory view of a generic pointer.  Loads and stores through it are
// global_* instructions, which count in vmcnt only; flat ones count in
// lgkmcnt as well, so every LDS wait behind them would also wait for the
// memory access (no overlap of a chunk's loads with the LDS work before it).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gm(T* p) {
  return (__attribute__((address_space(1))) T*)p;
}
// gm(p) or p itself (flat), chosen per kernel by measurement: the one-shift
// sweep with global accesses (n = 8000, m = 800: 6.30 -> 6.19 ms per solve),
// the packed sweep and the panel passes with flat ones (the 256-problem
// batch: 21.2 ms, against 21.8 ms with global accesses in either).
template <bool G, typename T>
__device__ __forceinline__ auto gsel(T* p) {
  if constexpr (G)
    return gm(p);
  else
    return p;
}

constexpr double kOmega = 1.0e+300;      // robust.c:14
constexpr double kOmegaInv = 1.0e-300;   // robust.c:15

// ---------------------------------------------------------------------------
// Power-of-two helpers
// ---------------------------------------------------------------------------

// Exponent e such that 2^e <= alpha, for alpha in (0, 1]; 0 for alpha >= 1.
__device__ __forceinline__ int pow2_exp(double alpha) {
  if (alpha >= 1.0) return 0;
  return ilogb(alpha);
}

// ---------------------------------------------------------------------------
// Robust scalar primitives (robust.c), returning the reference's factor.
// ---------------------------------------------------------------------------

// robust.c:144-166: alpha such that (alpha*b)/t cannot overflow.
__device__ __forceinline__ double protect_real_division(double b, double t) {
  double scale = 1.0;
  const double ab = fabs(b), at = fabs(t);
  if (at < kOmegaInv) {
    if (ab > at * kOmega) scale = (at * kOmega) / ab;
  } else {
    if (1.0 > at && ab > at * kOmega) scale = 1.0 / ab;
  }
  return scale;
}

// robust.c:175-185
__device__ __forceinline__ double protect_sum(double x, double y) {
  double scale = 1.0;
  if ((x > 0 && y > 0) || (x < 0 && y < 0))
    if (fabs(x) > kOmega - fabs(y)) scale = 0.5;
  return scale;
}

// robust.c:244-262: alpha such that (alpha*y) - t*(alpha*x) cannot overflow.
// When t and x are both below 2^500 their product is exact enough to decide
// without the reference's division (the two tests differ only within a
// rounding of the threshold Omega - y, where either answer is safe).
__device__ __forceinline__ double protect_update(double tnorm, double xnorm, double ynorm) {
  double scale = 1.0;
  if (xnorm <= 1.0) {
    if (tnorm * xnorm > kOmega - ynorm) scale = 0.5;
  } else if (tnorm < 0x1p500 && xnorm < 0x1p500) {
    if (tnorm * xnorm > kOmega - ynorm) scale = 0.5 / xnorm;
  } else {
    if (tnorm > (kOmega - ynorm) / xnorm) scale = 0.5 / xnorm;
  }
  return scale;
}

// robust.c:270-285
__device__ __forceinline__ double protect_double_update(double t1, double x1, double t2,
                                                        double x2, double y) {
  return fmin(protect_update(t1, x1, y), protect_update(t2, x2, y));
}

// robust.c:453-477 with power-of-two factors: x := x / (t - lambda).
// Returns the exponent of the factor applied to x (phi = alpha / s with
// s in {1, 1/2}); sets info if |t - lambda| was raised to smin.
__device__ __forceinline__ int solve_1x1(double smin, double t, double lambda, double& x,
                                         int& info) {
  const double s = protect_sum(t, -lambda);
  double csr = (s * t) - (s * lambda);
  if (fabs(csr) < smin) {
    csr = smin;
    info = 1;
  }
  const int ea = pow2_exp(protect_real_division(x, csr));
  x = ldexp(x, ea) / csr;
  return ea + (s == 1.0 ? 0 : 1);
}

// LAPACK 3.10 dlartg (Anderson 2017): c >= 0, r has the sign of f.  The
// reference transposes the rotation by negating s (reduce.c:14-24); so does
// this function: c*f - s*g = r, s*f + c*g = 0.
__device__ __forceinline__ void givens(double f, double g, double& c, double& s) {
  const double safmin = 0x1p-1022;
  const double safmax = 0x1p+1022;
  const double rtmin = 0x1p-511;                   // sqrt(safmin)
  const double rtmax = 0x1.6a09e667f3bcdp+510;     // sqrt(safmax / 2), rounded
  const double f1 = fabs(f), g1 = fabs(g);
  if (g == 0.0) {
    c = 1.0;
    s = -0.0;
  } else if (f == 0.0) {
    c = 0.0;
    s = -copysign(1.0, g);
  } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
    const double d = sqrt(f * f + g * g);
    c = f1 / d;
    const double r = copysign(d, f);
    s = -(g / r);
  } else {
    const double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
    const double fs = f / u, gs = g / u;
    const double d = sqrt(fs * fs + gs * gs);
    c = fabs(fs) / d;
    const double r = copysign(d, f);
    s = -(gs / r);
  }
}

// 1/sqrt(x) for x in a safe normal range: hardware estimate y (relative
// error below 2^-22) refined by one third-order step of the series
// (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + ..., e = 1 - x y^2:
// y <- y + y e (1/2 + 3e/8), error ~(5/16) e^3 < 2^-64 plus rounding
// (~1 ulp); four dependent operations instead of two Newton steps' six.
__device__ __forceinline__ double rsqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-x * y, y, 1.0);
  return fma(y * e, fma(0.375, e, 0.5), y);
}

// ---------------------------------------------------------------------------
// Wavefront helpers (wave64)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double readlane(double x, int lane) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, lane);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), lane);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_move(double old, double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint64_t o = __builtin_bit_cast(uint64_t, old);
  const uint32_t lo = __builtin_amdgcn_update_dpp((uint32_t)o, (uint32_t)u, CTRL, ROW_MASK, 0xf, false);
  const uint32_t hi = __builtin_amdgcn_update_dpp((uint32_t)(o >> 32), (uint32_t)(u >> 32), CTRL,
                                                  ROW_MASK, 0xf, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Maximum over the 64 lanes of a wavefront (all lanes active), returned
// uniformly.  DPP quad permutations, row rotations and row broadcasts.
__device__ __forceinline__ double wave_max(double x) {
  x = fmax(x, dpp_move<0xB1>(x, x));         // quad_perm [1,0,3,2]
  x = fmax(x, dpp_move<0x4E>(x, x));         // quad_perm [2,3,0,1]
  x = fmax(x, dpp_move<0x124>(x, x));        // row_ror:4
  x = fmax(x, dpp_move<0x128>(x, x));        // row_ror:8
  x = fmax(x, dpp_move<0x142, 0xa>(x, x));   // row_bcast:15
  x = fmax(x, dpp_move<0x143, 0xc>(x, x));   // row_bcast:31
  return readlane(x, 63);
}

// Maximum over the 16 lanes of each DPP row, in every lane of the row.
__device__ __forceinline__ double row16_max(double x) {
  x = fmax(x, dpp_move<0xB1>(x, x));    // quad_perm [1,0,3,2]
  x = fmax(x, dpp_move<0x4E>(x, x));    // quad_perm [2,3,0,1]
  x = fmax(x, dpp_move<0x124>(x, x));   // row_ror:4
  x = fmax(x, dpp_move<0x128>(x, x));   // row_ror:8
  return x;
}

// Lane L of every DPP row (16 lanes) broadcast to the whole row.
template <int L>
__device__ __forceinline__ double bcast16(double x) {
  return dpp_move<0x150 + L>(x, x);   // row_newbcast:L
}

__device__ __forceinline__ double bcast16_rt(double x, int l) {
  switch (l & 15) {
    case 0: return bcast16<0>(x);
    case 1: return bcast16<1>(x);
    case 2: return bcast16<2>(x);
    case 3: return bcast16<3>(x);
    case 4: return bcast16<4>(x);
    case 5: return bcast16<5>(x);
    case 6: return bcast16<6>(x);
    case 7: return bcast16<7>(x);
    case 8: return bcast16<8>(x);
    case 9: return bcast16<9>(x);
    case 10: return bcast16<10>(x);
    case 11: return bcast16<11>(x);
    case 12: return bcast16<12>(x);
    case 13: return bcast16<13>(x);
    case 14: return bcast16<14>(x);
    default: return bcast16<15>(x);
  }
}

__device__ __forceinline__ double wave_sum(double x) {
  x = x + dpp_move<0xB1>(0.0, x);
  x = x + dpp_move<0x4E>(0.0, x);
  x = x + dpp_move<0x124>(0.0, x);
  x = x + dpp_move<0x128>(0.0, x);
  x = x + dpp_move<0x142, 0xa>(0.0, x);
  x = x + dpp_move<0x143, 0xc>(0.0, x);
  return readlane(x, 63);
}

// Inclusive scan over lane index of the affine maps t -> a*t + b
// (composition: later lane applied after earlier lanes).  Returns (A, B) of
// the composed map for lanes 0..lane.
__device__ __forceinline__ void wave_affine_scan(double& a, double& b) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double pa = __shfl_up(a, off, 64);
    const double pb = __shfl_up(b, off, 64);
    if (lane >= off) {
      b = a * pb + b;
      a = a * pa;
    }
  }
}

// Inclusive prefix product over lane index.
__device__ __forceinline__ double wave_prefix_prod(double a) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const double p = __shfl_up(a, off, 64);
    if (lane >= off) a = a * p;
  }
  return a;
}

// LAPACK 3.10 dnrm2 thresholds (Blue's algorithm), exact powers of two
// (shifted_Hessenberg_solve.c:15-18).
constexpr double kTsml = 0x1p-511;
constexpr double kTbig = 0x1p+486;
constexpr double kSsml = 0x1p+537;
constexpr double kSbig = 0x1p-538;

struct Acc3 {
  double asml, amed, abig;
};

__device__ __forceinline__ void acc3_add(Acc3& a, double v) {
  const double x = fabs(v);
  if (x > kTbig) {
    const double t = x * kSbig;
    a.abig += t * t;
  } else if (x < kTsml) {
    const double t = x * kSsml;
    a.asml += t * t;
  } else {
    a.amed += x * x;
  }
}

// Combination of the three accumulators (shifted_Hessenberg_solve.c:198-229).
__device__ __forceinline__ double acc3_norm(Acc3 a) {
  double ynrm, scale = 1.0;
  if (a.abig > 0.0) {
    ynrm = a.abig + (a.amed * kSbig) * kSbig;
    scale = 1.0 / kSbig;
  } else if (a.asml > 0.0) {
    if (a.amed > 0.0) {
      const double amed = sqrt(a.amed);
      const double asml = sqrt(a.asml) / kSsml;
      const double ymin = asml > amed ? amed : asml;
      const double ymax = asml > amed ? asml : amed;
      ynrm = ymax * ymax * (1.0 + (ymin / ymax) * (ymin / ymax));
    } else {
      scale = 1.0 / kSsml;
      ynrm = a.asml;
    }
  } else {
    ynrm = a.amed;
  }
  return scale * sqrt(ynrm);
}

}  // namespace hsrq
