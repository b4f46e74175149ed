// hsrq_kernels.hip -- gfx950 kernels of the robust tiled inverse iteration.
//
// One problem is one fused right-to-left sweep over tile columns J:
//   k_diag(J)    one wavefront per shift: Givens generation on the diagonal
//                tile fused with the robust back substitution (reference
//                reduce.c:27-68 + solve-task.c:19-225), then the coefficient
//                vectors W_J (cross-over reduction) and Z_J (update) and the
//                per-(tile row, shift) scaling exponents of the update;
//   k_update(J)  FP64 MFMA: [R~_out | dY](0:l) = H(0:l, l-1 : l+b-2) [W | Z]
//                with a scaled epilogue that adds the cross-over terms and
//                updates Y (reference reduce.c:73-143 + update-task.c:16-262
//                in one pass over the panel, all shifts at once);
//   k_backtransform  x = Q^T y, 2-norm, convergence, normalisation
//                (shifted_Hessenberg_solve.c:104-245) as an affine scan.
// Work is described by item tables, so one launch serves any number of
// problems (batching and continuous batching use the same kernels).
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include <type_traits>

#include "hsrq_device.h"
#include "hsrq_internal.h"

namespace hsrq {

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// K0: norms of H.  rowinf[I] = ||H(I, :)||_oo (the reference's
// dlange('I', n, n, H), inverse_iteration.c:154, taken per tile row) and
// diagn[I] = ||H(I, I)||_oo of the diagonal tile (solve-task.c:30-34), over
// the Hessenberg part (j >= i - 1; entries below are zero by contract).
// Pass 1: absolute row sums of 64 x 64 blocks (tile-aligned like the row
// blocks) into part[cb * n + i]; pass 2: fixed-order sums and maxima per
// tile (deterministic, no atomics).
// ---------------------------------------------------------------------------
__host__ __device__ inline bool norm_block_needed(const ProbDev& P, int rb, int cb) {
  int t1, r0, nr, t2, c0, nc;
  rowblock_range(P, rb, t1, r0, nr);
  rowblock_range(P, cb, t2, c0, nc);
  return c0 + nc - 1 >= r0 - 1;
}

__global__ __launch_bounds__(256) void k_norm_part(const ProbDev* __restrict__ probs,
                                                   const GemmItem* __restrict__ items) {
  __shared__ double red[4][64];
  const GemmItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  int t1, r0, nr, t2, c0, nc;
  rowblock_range(P, it.rb, t1, r0, nr);
  rowblock_range(P, it.cb, t2, c0, nc);
  const int row = threadIdx.x & 63, g = threadIdx.x >> 6;
  double a = 0.0;
  if (row < nr) {
    const int i = r0 + row;
    const auto* h = gm(P.H + i);
    double v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int cc = g + 4 * q;
      const int j = c0 + cc;
      v[q] = (cc < nc && j >= i - 1) ? fabs(h[(int64_t)j * P.ldH]) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) a += v[q];
  }
  red[g][row] = a;
  __syncthreads();
  if (threadIdx.x < nr)
    gm(P.part)[(int64_t)it.cb * P.n + r0 + threadIdx.x] =
        (red[0][threadIdx.x] + red[1][threadIdx.x]) + (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_norm_final(const ProbDev* __restrict__ probs,
                                                    const RowItem* __restrict__ items) {
  __shared__ double red[2][4];
  const RowItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int I = it.rb;
  const int l = tile_first_row(P, I), nb = tile_rows(P, I);
  const int r = threadIdx.x;
  double rs = 0.0, ds = 0.0;
  if (r < nb) {
    const int rb = tile_first_rowblock(P, I) + r / kRowBlock;
    const int cbd0 = tile_first_rowblock(P, I), cbd1 = cbd0 + tile_num_rowblocks(P, I);
    // the column blocks a row needs are cb >= cb_lo (the Hessenberg part):
    // loaded 8 at a time (independent loads in flight), added in cb order
    int cb_lo = 0;
    while (cb_lo < P.nrb && !norm_block_needed(P, rb, cb_lo)) ++cb_lo;
    const auto* __restrict__ pc = gm(P.part + l + r);
    for (int cb0 = cb_lo; cb0 < P.nrb; cb0 += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        v[q] = cb0 + q < P.nrb ? pc[(int64_t)(cb0 + q) * P.n] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int cb = cb0 + q;
        if (cb < P.nrb) {
          rs += v[q];
          if (cb >= cbd0 && cb < cbd1) ds += v[q];
        }
      }
    }
  }
  const int wave = threadIdx.x >> 6;
  const double mr = wave_max(rs), md = wave_max(ds);
  if ((threadIdx.x & 63) == 0) {
    red[0][wave] = mr;
    red[1][wave] = md;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    gm(P.rowinf)[I] = fmax(fmax(red[0][0], red[0][1]), fmax(red[0][2], red[0][3]));
    gm(P.diagn)[I] = fmax(fmax(red[1][0], red[1][1]), fmax(red[1][2], red[1][3]));
  }
}

// ---------------------------------------------------------------------------
// Init: starting vectors (reference inverse_iteration.c:19-57), initial
// cross-over columns H(:, n-1) - mu e_{n-1} (reduce.c:239-246), exponents 0,
// per-row-block norms.  Workgroup = (problem, row block, 4 shifts).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_init(const ProbDev* __restrict__ probs,
                                              const RowItem* __restrict__ items) {
  const RowItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int k = it.k0 + wave;
  if (k >= P.m) return;  // whole wavefront exits; no barriers below
  double hinf = 0.0;
  for (int t = lane; t < P.T; t += 64) hinf = fmax(hinf, gm(P.rowinf)[t]);
  hinf = wave_max(hinf);
  const double eps3 = hinf * 0x1p-53;   // inverse_iteration.c:157-158
  if (it.rb == 0 && k == 0 && lane == 0) {
    gm(P.scal)[0] = hinf;
    gm(P.scal)[1] = eps3;
  }
  int tile, r0, nr;
  rowblock_range(P, it.rb, tile, r0, nr);
  const int n = P.n;
  double yv = 0.0;
  const int its = P.its_k ? P.its_k[k] : P.its;
  if (lane < nr) {
    const int row = r0 + lane;
    if (its == -2) {
      yv = gm(P.Y)[row + (int64_t)k * P.ldY];   // given right-hand side
    } else if (its < 0) {
      yv = eps3;
    } else {
      const double sq = sqrt((double)n);
      yv = (row == 0) ? eps3 : eps3 / (sq + 1.0);
      if (row == n - 1 - its) yv -= eps3 * sq;
    }
    if (its != -2) gm(P.Y)[row + (int64_t)k * P.ldY] = yv;
    double h = P.H[row + (int64_t)(n - 1) * P.ldH];
    if (row == n - 1) h -= gm(P.wr)[k];
    gm(P.Rt[(P.T - 1) & 1])[row + (int64_t)k * n] = h;
  }
  const double mx = wave_max(fabs(yv));
  if (lane == 0) gm(P.Ynb[P.T & 1])[(int64_t)it.rb * P.m + k] = mx;
  if (it.rb == 0) {
    for (int t = lane; t < P.T; t += 64) gm(P.Ep[P.T & 1])[(int64_t)t * P.m + k] = 0;
    if (lane == 0) P.status[k] = 0;
  }
}

// ---------------------------------------------------------------------------
// Power-of-two equilibration of H for the FP32 panel pass (once per solve):
// ccol[j] = -ilogb(max_i |H(i, j)|) over the Hessenberg part of column j, then
// rrow[i] = ilogb(max_j |H(i, j)| 2^ccol[j]).  For H = D H0 D^-1 with D a
// diagonal of powers of two the scaled entries H(i, j) 2^(ccol[j] - rrow[i])
// are of the size of H0's however wide D's range, so the conversion to float
// keeps them.  One wavefront per column; rows in blocks of 64 (threads read a
// column's consecutive rows together).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_eq_col(const ProbDev* __restrict__ probs) {
  const ProbDev& P = probs[0];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= P.n) return;   // whole wavefront
  const int rows = min(j + 2, P.n);
  const auto* __restrict__ col = gm(P.H + (int64_t)j * P.ldH);
  double a = 0.0;
  for (int i = lane; i < rows; i += 64) a = fmax(a, fabs(col[i]));
  a = wave_max(a);
  if (lane == 0) gm(P.ccol)[j] = (a > 0.0 && isfinite(a)) ? -ilogb(a) : 0;
}

// Row maxima in two passes: block (64 rows, column chunk blockIdx.y of
// gridDim.y) into part[chunk * n + row], then the maximum over the chunks.
__global__ __launch_bounds__(256) void k_eq_row(const ProbDev* __restrict__ probs) {
  const ProbDev& P = probs[0];
  __shared__ double red[4][64];
  const int r = threadIdx.x & 63, ph = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + r;
  const int64_t ldH = P.ldH;
  const int j0 = max((int)blockIdx.x * 64 - 1, 0);   // signed: block 0 starts at column 0
  const int len = (P.n - j0 + gridDim.y - 1) / gridDim.y;
  const int ja = j0 + blockIdx.y * len, jb = min(ja + len, P.n);
  double a = 0.0;
  for (int j = ja + ph; j < jb; j += 4)
    if (i < P.n && j >= i - 1) a = fmax(a, ldexp(fabs(P.H[i + (int64_t)j * ldH]), gm(P.ccol)[j]));
  red[ph][r] = a;
  __syncthreads();
  if (ph == 0 && i < P.n)
    gm(P.part)[(int64_t)blockIdx.y * P.n + i] =
        fmax(fmax(red[0][r], red[1][r]), fmax(red[2][r], red[3][r]));
}

__global__ __launch_bounds__(256) void k_eq_rowfin(const ProbDev* __restrict__ probs, int chunks) {
  const ProbDev& P = probs[0];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.n) return;
  double a = 0.0;
  for (int c = 0; c < chunks; ++c) a = fmax(a, gm(P.part)[(int64_t)c * P.n + i]);
  gm(P.rrow)[i] = (a > 0.0 && isfinite(a)) ? ilogb(a) : 0;
}

// H32(i, j) = H(i, j) 2^(ccol[j] - rrow[i]) in float over the Hessenberg part
// (rows i <= j + 1; the panel passes read no other), one wavefront per column.
__global__ __launch_bounds__(256) void k_eq_conv(const ProbDev* __restrict__ probs) {
  const ProbDev& P = probs[0];
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= P.n) return;
  const int rows = min(j + 2, P.n);
  const auto* __restrict__ col = gm(P.H + (int64_t)j * P.ldH);
  auto* __restrict__ out = gm(P.H32 + (int64_t)j * P.n);
  const int cj = gm(P.ccol)[j];
  for (int i = lane; i < rows; i += 64) out[i] = (float)ldexp(col[i], cj - gm(P.rrow)[i]);
}

// Exponent of the largest |x_r| 2^-ccol[l - 1 + r] over a shift's coefficient
// column x (rows r < nb of the tile starting at row l), 0 if x is 0: the scale
// of the [W | Z] column in the FP32 panel pass.  Per lane: its rows' best.
__device__ __forceinline__ double coef_exp_part(const ProbDev& P, int l, double x, int r, int nb,
                                                double cur) {
  if (r < nb && x != 0.0) cur = fmax(cur, (double)(ilogb(x) - gm(P.ccol)[l - 1 + r]));
  return cur;
}

// ---------------------------------------------------------------------------
// Scale exponents of the panel pass of tile column J for tile row I < J and
// shift k: consolidate the scales of y_I and Z_J (update-task.c:68-88), then
// one protection for the whole fused update (update-task.c:103-248):
// |update(i)| <= (||H(I,:)||_oo + |mu|) ||Z|| + |R~(i)| |Zlast| and
// |R~(i)| <= ||(H - mu I)(i,:)||_2, so t_I = 2 (||H(I,:)||_oo + |mu|).
// Reads the state the panel pass of tile column J+1 left (Ep, Ynb of parity
// J+1) and tile J's final exponent; returns tile I's new exponent (for
// parity J) and the epilogue exponents sy (of y) and sz (of the update).
// Every block that updates rows of tile I derives the same values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int update_exponents(const ProbDev& P, int I, int J, int k, double mu,
                                                int& sy, int& sz) {
  const int m = P.m;
  const auto* __restrict__ Yb = gm(P.Ynb[(J + 1) & 1]);
  const int rb0 = tile_first_rowblock(P, I), nrbI = tile_num_rowblocks(P, I);
  double yn = 0.0;
  for (int rb = rb0; rb < rb0 + nrbI; ++rb) yn = fmax(yn, Yb[(int64_t)rb * m + k]);
  const int EI = gm(P.Ep[(J + 1) & 1])[(int64_t)I * m + k];
  const int EJ = gm(P.Ef)[(int64_t)J * m + k];
  const double zn = gm(P.aux[J & 1])[4 * k + 2];
  const int p = (int)gm(P.aux[J & 1])[4 * k + 3];
  const int c0 = min(EI, EJ);
  const double yI = ldexp(yn, c0 - EI);
  const double zJ = ldexp(zn, c0 - EJ);
  const double tI = 2.0 * (gm(P.rowinf)[I] + fabs(mu));
  const int ec = c0 + pow2_exp(protect_update(tI, zJ, yI));
  sy = ec - EI;
  sz = ec - EJ + p;
  return ec;
}

// ---------------------------------------------------------------------------
// The panel pass of tile column J+1 over the rows of tile J, for one shift k:
// the rows the sweep of tile J starts from.  It runs at the start of that
// sweep (same launch, no separate critical-path GEMM launch):
//   acc = H(tile J rows, l'-1 : l'+b-2) [W_k | Z_k],  l' = first row of J+1,
//   R~ = acc_W + P_k R~in,  update = acc_Z + Zlast_k R~in  (minus mu B(0,.)
//   on row l'-1), y = 2^sy y - 2^sz update      (update-task.c:103-248),
// in FP64 FMAs in column order, the panel staged CW columns at a time through
// the sweep's LDS chunk buffer (rows of the tile, column-major, ld BMAX).
// Lanes hold rows r0 + STRIDE q.  Every sweep layout computes each row the
// same way, so the layouts stay bit-identical.
// ---------------------------------------------------------------------------
template <int RPL, int STRIDE, int CW, int BMAX, int NL, int SPW, bool GLOBAL>
__device__ __forceinline__ void own_rows_update(const ProbDev& P, int J, int k0, int k, int sidx,
                                                int l, int nb, int r0, double mu, double* Hs,
                                                double* Bsh, double (&pf)[NL], double (&v)[RPL],
                                                double (&y)[RPL], int& E0) {
  const int b = P.b, m = P.m, n = P.n;
  // epilogue operands first: their loads overlap the panel product
  const auto* aux = gsel<GLOBAL>(P.aux[(J + 1) & 1]);
  const auto* __restrict__ Rin = gsel<GLOBAL>(P.Rt[(J + 1) & 1] + (int64_t)k * n + l);
  const auto* __restrict__ yg = gsel<GLOBAL>(P.Y + (int64_t)k * P.ldY + l);
  double rin[RPL], yo[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = r0 + STRIDE * q;
    rin[q] = r < nb ? Rin[r] : 0.0;
    yo[q] = r < nb ? yg[r] : 0.0;
  }
  const auto* __restrict__ Bg = gsel<GLOBAL>(P.Bm[(J + 1) & 1]);
  const double Pk = aux[4 * k + 0], Zl = aux[4 * k + 1];
  const double b0w = Bg[(int64_t)(2 * k) * b], b0z = Bg[(int64_t)(2 * k + 1) * b];
  const int64_t ldH = P.ldH;
  const auto* __restrict__ Hq = gsel<GLOBAL>(P.H + (int64_t)(l + nb - 1) * ldH + l);
  // [W | Z] of the workgroup's SPW shifts, CW columns per chunk, in LDS as
  // Bsh[c * 2 SPW + 2 s + {0: W, 1: Z}]
  constexpr int NBT = CW * 2 * SPW, NB = (NBT + 255) / 256;
  double pb[NB];
  constexpr int CSTEP = 256 / BMAX;   // thread t: row t % BMAX, columns t / BMAX + CSTEP i
  const int hr = threadIdx.x % BMAX, hcb = threadIdx.x / BMAX;
  const int64_t hstep = (int64_t)CSTEP * ldH;
  auto stage = [&](int chunk) __attribute__((always_inline)) {
    const auto* __restrict__ src = Hq + hr + (int64_t)(chunk * CW + hcb) * ldH;
    if (nb == BMAX) {   // full tile
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = src[i * hstep];
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = hr < nb ? src[i * hstep] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int idx = threadIdx.x + 256 * i;
      const int j = idx / CW, c = idx - j * CW;
      const int ks = min(k0 + (j >> 1), m - 1);
      pb[i] = idx < NBT ? Bg[(int64_t)(2 * ks + (j & 1)) * b + chunk * CW + c] : 0.0;
    }
  };
  double aw[RPL], az[RPL];
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    aw[q] = 0.0;
    az[q] = 0.0;
  }
  const int nch = b / CW;
  stage(0);
  int sy, sz;   // after the first chunk's loads are issued: the latencies overlap
  E0 = update_exponents(P, J, J + 1, k, mu, sy, sz);
  for (int ch = 0; ch < nch; ++ch) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NL; ++i) Hs[threadIdx.x + 256 * i] = pf[i];
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int idx = threadIdx.x + 256 * i;
      const int j = idx / CW, c = idx - j * CW;
      if (idx < NBT) Bsh[c * 2 * SPW + j] = pb[i];
    }
    __syncthreads();
    if (ch + 1 < nch) stage(ch + 1);
    // G columns per group, the next group's LDS reads issued before this
    // group's FMAs (software pipelined; the column order of the sums is kept).
    // G = 1 where registers are short (the packed layouts).
    constexpr int G = RPL == 2 ? 4 : (RPL == 1 ? 2 : 1);
    double hb[2][G][RPL], wb[2][G], zb[2][G];
    auto ld = [&](int buf, int c0) __attribute__((always_inline)) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        wb[buf][g] = Bsh[(c0 + g) * 2 * SPW + 2 * sidx];
        zb[buf][g] = Bsh[(c0 + g) * 2 * SPW + 2 * sidx + 1];
#pragma unroll
        for (int q = 0; q < RPL; ++q) hb[buf][g][q] = Hs[(c0 + g) * BMAX + r0 + STRIDE * q];
      }
    };
    ld(0, 0);
#pragma unroll
    for (int c0 = 0; c0 < CW; c0 += G) {
      const int cur = (c0 / G) & 1;
      if (c0 + G < CW) ld(cur ^ 1, c0 + G);
#pragma unroll
      for (int g = 0; g < G; ++g)
#pragma unroll
        for (int q = 0; q < RPL; ++q) {
          aw[q] = fma(hb[cur][g][q], wb[cur][g], aw[q]);
          az[q] = fma(hb[cur][g][q], zb[cur][g], az[q]);
        }
    }
  }
  __syncthreads();   // the chunk buffer is the sweep's again
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = r0 + STRIDE * q;
    if (r < nb) {
      double vw = fma(Pk, rin[q], aw[q]);
      double vz = fma(Zl, rin[q], az[q]);
      if (r == nb - 1) {
        vw -= mu * b0w;
        vz -= mu * b0z;
      }
      v[q] = vw;
      y[q] = ldexp(yo[q], sy) - ldexp(vz, sz);
    } else {
      v[q] = 0.0;
      y[q] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// K1: diagonal tile J, one wavefront per shift, RPL rows per lane.
// ---------------------------------------------------------------------------
// a[q] for a uniform slot q, as a binary tree of selects on the bits of q
// (keeps the register array out of memory).
template <int N>
__device__ __forceinline__ double pick_tree(const double* a, int q, int bit) {
  if constexpr (N == 1) {
    return a[0];
  } else {
    const double lo = pick_tree<N / 2>(a, q, bit >> 1);
    const double hi = pick_tree<N / 2>(a + N / 2, q, bit >> 1);
    return (q & bit) ? hi : lo;
  }
}
// compile-time int argument of the generic lambdas below
template <int V>
using IC = std::integral_constant<int, V>;

// a[q] of the first NS slots (NS = RPL by default)
template <int RPL, int NS = RPL>
__device__ __forceinline__ double pick(const double (&a)[RPL], int q) {
  static_assert((NS & (NS - 1)) == 0 && NS <= RPL, "power of two");
  return pick_tree<NS>(a, q, NS / 2);
}

// 128-row tiles: at most 128 VGPRs and 16-column H chunks (33 KiB of LDS),
// so that three panel-pass workgroups of at most 128 VGPRs and 40 KiB run
// beside a sweep workgroup on every CU (n = 8000, m = 800: 6.56 -> 6.31 ms
// against 64-column chunks, up to 193 VGPRs and two panel-pass workgroups).
template <int RPL>
constexpr int diag_chunk_cols() { return RPL >= 2 ? 16 : 64; }   // tile columns per LDS chunk
// LDS of the one-shift sweep, in doubles: the H chunk (column-major, ld
// BMAX) and the [W | Z] chunk of the workgroup's 4 shifts
template <int RPL>
constexpr int diag_smem_doubles() { return diag_chunk_cols<RPL>() * (64 * RPL + 2 * 4); }

template <int RPL>
__device__ __forceinline__ void diag_body(const ProbDev* __restrict__ probs,
                                          const K1Item* __restrict__ items, int bid, double* smem) {
  __builtin_amdgcn_s_setprio(2);   // latency-bound: ahead of bulk panel passes
  constexpr int BMAX = 64 * RPL;
  constexpr int CW = diag_chunk_cols<RPL>();
  double* Hs = smem;               // chunk, column-major, ld BMAX
  double* Bsh = smem + CW * BMAX;  // [W | Z] chunk of the 4 shifts

  const K1Item it = items[bid];
  const ProbDev& P = probs[it.prob];
  const int J = it.J;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int m = P.m, n = P.n;
  const int kraw = it.k0 + wave;
  const bool valid = kraw < m;
  const int k = valid ? kraw : m - 1;
  const int l = tile_first_row(P, J), nb = tile_rows(P, J);
  const double mu = gm(P.wr)[k];
  const auto* __restrict__ H = gm(P.H);
  const int64_t ldH = P.ldH;
  const auto* Rin = gm(P.Rt[J & 1] + (int64_t)k * n + l);
  auto* yg = gm(P.Y + (int64_t)k * P.ldY + l);

  // H tile chunks: CW columns each, staged through registers one chunk ahead
  // so the global loads overlap the sweep of the previous chunk.
  constexpr int NL = CW * BMAX / 256;
  double pf[NL];

  // thread t loads row t % BMAX of columns t / BMAX + CSTEP i of the chunk
  constexpr int CSTEP = 256 / BMAX;
  static_assert(256 % BMAX == 0 && CW % CSTEP == 0, "chunk tiling");
  auto issue = [&](int chunk) {
    const int c0 = chunk * CW;
    const int nc = min(CW, nb - c0);
    const int r = threadIdx.x % BMAX, cb = threadIdx.x / BMAX;
    const auto* __restrict__ src = H + (l + r) + (int64_t)(l + c0 + cb) * ldH;
    const int64_t step = (int64_t)CSTEP * ldH;
    if (nc == CW && nb == BMAX) {   // full chunk of a full tile
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = src[i * step];
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = (cb + CSTEP * i < nc && r < nb) ? src[i * step] : 0.0;
    }
  };

  double v[RPL], y[RPL], cr[RPL], sr[RPL];
  int E;   // scale exponent of the tile's y
  if (J + 1 < P.T) {   // rows of this tile after the panel pass of tile column J+1
    own_rows_update<RPL, 64, CW, BMAX, NL, 4, true>(P, J, it.k0, k, wave, l, nb, lane, mu, Hs, Bsh, pf,
                                             v, y, E);
  } else {       // the last tile column starts from the initial vectors
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = lane + 64 * q;
      v[q] = r < nb ? Rin[r] : 0.0;
      y[q] = r < nb ? yg[r] : 0.0;
    }
    E = gm(P.Ep[(J + 1) & 1])[(int64_t)J * m + k];
  }
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    cr[q] = 1.0;
    sr[q] = 0.0;
  }
  const double Hnorm = gm(P.diagn)[J];
  const double eps = 0x1p-53;
  const double smlnum = fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps));   // s_H_s.c:70-71
  const double smin = fmax(eps * fabs(mu), smlnum);                        // solve-task.c:47
  int info = 0;
  const double slack = 1.0 + 0x1p-40;
  const bool exact = P.force_exact != 0;
  int cur = -1;
  auto column = [&](int col) -> const double* {
    const int ch = col / CW;
    if (ch != cur) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NL; ++i) Hs[threadIdx.x + 256 * i] = pf[i];
      __syncthreads();
      if (ch > 0) issue(ch - 1);
      cur = ch;
    }
    return &Hs[(col - ch * CW) * BMAX];
  };

  // Upper bounds of |y| and |v| over the rows still active.
  double By, Bv;
  {
    double ly = 0.0, lv = 0.0;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      ly = fmax(ly, fabs(y[q]));
      lv = fmax(lv, fabs(v[q]));
    }
    By = wave_max(ly);
    Bv = wave_max(lv);
  }

  // ---- v-chain step st: rotation G_st from the running column
  // (reduce.c:45-51 / solve-task.c:171), column st of R for rows < st
  // (r_i = c v_i - s h_i, h_{st-1} = H(st-1,st-1) - mu) and the next running
  // column (v_i = c h_i + s v_i).  The pivot R(st,st) = c v_st - s h_sub is
  // the rotation's r = sign(v_st) ||(v_st, h_sub)|| itself, so it is taken
  // from the rotation with its reciprocal.  Fast form: reciprocal square root
  // (in LAPACK 3.10 dlartg's unscaled range); `rare` flags every input
  // outside it, which the full form (dlartg's branches) handles.
  struct VOut {
    double c, s, r, rinv, br, bv;
  };
  // NSc: the row slots that can hold active rows (the first NS; see the
  // step loop); nr and nv are written for those only
  auto vstep = [&](auto NSc, int st, const double* hc, bool full, VOut& o, double (&nr)[RPL],
                   double (&nv)[RPL]) -> bool {
    constexpr int NS = decltype(NSc)::value % 100;
    constexpr bool A0 = decltype(NSc)::value >= 100;   // slot 0: all rows active, none st-1
    // (with A0, row st lies in the last of the NS slots)
    const double vs = readlane(A0 ? v[NS - 1] : pick<RPL, NS>(v, st >> 6), st & 63);
    const double hsub = hc[st];
    const double hdd = hc[st - 1] - mu;
    double c, s, r, rinv;
    bool rare = false;
    if (full) {
      givens(vs, hsub, c, s);
      r = c * vs - s * hsub;
      rinv = 1.0 / r;
    } else {
      const double f1 = fabs(vs), g1 = fabs(hsub);
      rare = !((f1 > 0x1p-511) & (f1 < 0x1.6a09e667f3bcdp+510) & (g1 > 0x1p-511) &
               (g1 < 0x1.6a09e667f3bcdp+510));
      const double d2 = fma(vs, vs, hsub * hsub);
      const double id = rsqrt_nr(d2);
      rinv = copysign(id, vs);
      r = d2 * rinv;
      c = vs * rinv;
      s = -(hsub * rinv);
    }
    const double Hm = fmax(Hnorm, fabs(hdd));
    o.c = c;
    o.s = s;
    o.r = r;
    o.rinv = rinv;
    o.br = fma(fabs(c), Bv, fabs(s) * Hm) * slack;
    o.bv = fma(fabs(c), Hm, fabs(s) * Bv) * slack;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int row = lane + 64 * q;
      if (A0 && q == 0) {
        const double h = hc[row];
        nr[q] = fma(c, v[q], -(s * h));
        nv[q] = fma(c, h, s * v[q]);
      } else {
        const bool act = row <= st - 1;
        const double h = (row == st - 1) ? hdd : hc[row];
        nr[q] = act ? fma(c, v[q], -(s * h)) : 0.0;
        nv[q] = act ? fma(c, h, s * v[q]) : v[q];
      }
    }
    return rare;
  };

  // ---- y-chain step st, fast form: y(st) /= R(st,st) and
  // y(0:st-1) -= y(st) R(0:st-1, st) when no protection can trigger
  // (|R(st,st)| >= smin, the quotient and the update provably below Omega
  // from the running bounds).
  auto ystep_fast = [&](auto NSc, int st, const VOut& o, const double (&r)[RPL],
                        double (&ny)[RPL], double& byn) -> bool {
    constexpr int NS = decltype(NSc)::value % 100;
    constexpr bool A0 = decltype(NSc)::value >= 100;
    const double yj = readlane(A0 ? y[NS - 1] : pick<RPL, NS>(y, st >> 6), st & 63);
    const double acsr = fabs(o.r);
    bool rare = !(acsr >= smin) | ((acsr < 1.0) & (fabs(yj) > acsr * kOmega));
    const double yq = yj * o.rinv;
    byn = fma(fabs(yq), o.br, By) * slack;
    rare = rare | !(byn <= kOmega);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      // r is 0 on the rows outside 0 .. st-1 (vstep), where the update
      // leaves y as it is (yq is finite when this result is used)
      const int row = lane + 64 * q;
      const double u = fma(-yq, r[q], y[q]);
      ny[q] = (A0 && q == 0) ? u : ((row == st) ? yq : u);
    }
    return rare;
  };

  // ---- y-chain step st, protected form.  solve_1x1_real_system
  // (robust.c:453-477) on the pivot R(st,st) with power-of-two factors:
  // |R| < smin -> smin (info), division protected against overflow; then ONE
  // protection of the whole update: with a = max_i (|y_i| + |y(st)|
  // |R(i,st)|) (exact, one wavefront reduction, in a scaled form) the tile is
  // scaled by a power of two below 1/a whenever a > Omega.
  auto ystep_slow = [&](int st, const VOut& o, const double (&r)[RPL]) {
    double yj = readlane(pick<RPL>(y, st >> 6), st & 63);
    double csr = o.r, cinv = o.rinv;
    if (fabs(csr) < smin) {
      csr = smin;
      cinv = 1.0 / smin;
      info = 1;
    }
    const int e1 = pow2_exp(protect_real_division(yj, csr));
    yj = ldexp(yj, e1) * cinv;
    if (e1 != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q)
        if (lane + 64 * q != st) y[q] = ldexp(y[q], e1);
      E += e1;
    }
    const double ay = fabs(yj);
    const int kk = (ay > 0.0) ? ilogb(ay) : 0;
    const double ays = ldexp(ay, -kk);
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
      if (lane + 64 * q <= st - 1) a = fmax(a, fma(ays, fabs(r[q]), ldexp(fabs(y[q]), -kk)));
    a = wave_max(a);
    int e2 = 0;
    if (a > ldexp(kOmega, -kk)) e2 = -(ilogb(a) + kk) - 1;
    if (e2 != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q) y[q] = ldexp(y[q], e2);
      yj = ldexp(yj, e2);
      E += e2;
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int row = lane + 64 * q;
      if (row <= st - 1) y[q] = fma(-yj, r[q], y[q]);
      if (row == st) y[q] = yj;
    }
    By = ldexp(a, kk + e2) * slack;
  };

  auto keep_rotation = [&](auto NSc, int st, const VOut& o) {
    constexpr int NS = decltype(NSc)::value % 100;
    constexpr int Q0 = decltype(NSc)::value >= 100 ? 1 : 0;   // slot 0 owns no row st
    const int own = (lane == (st & 63)) ? (st >> 6) : -1;
#pragma unroll
    for (int q = Q0; q < NS; ++q) {
      cr[q] = (own == q) ? o.c : cr[q];
      sr[q] = (own == q) ? o.s : sr[q];
    }
  };

  if (nb >= 2) {
    issue((nb - 2) / CW);
    VOut cv;
    double rr[RPL], nr[RPL], nv[RPL], ny[RPL];
    {
      const double* hc = column(nb - 2);
      if (vstep(IC<RPL>{}, nb - 1, hc, false, cv, nr, nv))
        vstep(IC<RPL>{}, nb - 1, hc, true, cv, nr, nv);
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        v[q] = nv[q];
        rr[q] = nr[q];
      }
      Bv = cv.bv;
      keep_rotation(IC<RPL>{}, nb - 1, cv);
    }
    // Steady state: v-step s-1 and y-step s are independent and interleave.
    // Steps s hi .. lo over the first NS row slots: the slots from NS on hold
    // finished rows only (v and y kept, R column rr already 0 there).
    auto steps = [&](auto NSc, int hi, int lo) {
      constexpr int NS = decltype(NSc)::value % 100;
      for (int s = hi; s >= lo; --s) {
        const double* hc = column(s - 2);
        VOut nvo;
        double byn;
        const bool rv = vstep(NSc, s - 1, hc, false, nvo, nr, nv);
        const bool ry = ystep_fast(NSc, s, cv, rr, ny, byn);
        if (rv) vstep(NSc, s - 1, hc, true, nvo, nr, nv);
        if (ry || exact) {
          ystep_slow(s, cv, rr);
        } else {
#pragma unroll
          for (int q = 0; q < NS; ++q) y[q] = ny[q];
          By = byn;
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          v[q] = nv[q];
          rr[q] = nr[q];
        }
        Bv = nvo.bv;
        cv = nvo;
        keep_rotation(NSc, s - 1, cv);
      }
    };
    if constexpr (RPL >= 4) {
      // below s = 64 NS (y-step s, v-step s - 1) every active row and row s
      // lie in the first NS slots: all, half, then a quarter of them
      steps(IC<RPL>{}, nb - 1, 32 * RPL);
      steps(IC<RPL / 2>{}, min(nb - 1, 32 * RPL - 1), 16 * RPL);
      steps(IC<RPL / 4>{}, min(nb - 1, 16 * RPL - 1), 2);
    } else if constexpr (RPL == 2) {
      // from s = 66 on, slot 0 (rows 0 .. 63) holds active rows only, none
      // of them row s or s - 2 (IC<100 + NS>: that slot without predicates);
      // below s = 64 every active row and row s lie in slot 0
      steps(IC<100 + RPL>{}, nb - 1, 66);
      steps(IC<RPL>{}, min(nb - 1, 65), 64);
      steps(IC<RPL / 2>{}, min(nb - 1, 63), 2);
    } else if constexpr (RPL >= 2) {
      // below s = 32 RPL (y-step s, v-step s - 1) every active row and row s
      // lie in the lower half of the slots
      constexpr int MID = 32 * RPL;
      steps(IC<RPL>{}, nb - 1, MID);
      steps(IC<RPL / 2>{}, min(nb - 1, MID - 1), 2);
    } else {
      steps(IC<RPL>{}, nb - 1, 2);
    }
    {
      double byn;
      if (ystep_fast(IC<RPL>{}, 1, cv, rr, ny, byn) || exact) {
        ystep_slow(1, cv, rr);
      } else {
#pragma unroll
        for (int q = 0; q < RPL; ++q) y[q] = ny[q];
        By = byn;
      }
    }
  }

  // Final rotation and y(0) (solve-task.c:104-118, :212-218).
  {
    double v0 = readlane(v[0], 0);
    if (J > 0) {
      const double hleft = H[l + (int64_t)(l - 1) * ldH];
      double c0 = 1.0, s0 = 0.0;
      if (hleft != 0.0) givens(v0, hleft, c0, s0);
      v0 = v0 * c0 - hleft * s0;
      if (lane == 0) {
        cr[0] = c0;
        sr[0] = s0;
      }
    }
    double y0 = readlane(y[0], 0);
    double csr = v0;
    if (fabs(csr) < smin) {
      csr = smin;
      info = 1;
    }
    const int e = pow2_exp(protect_real_division(y0, csr));
    y0 = ldexp(y0, e) / csr;
    if (e != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q) y[q] = ldexp(y[q], e);
      E += e;
    }
    if (lane == 0) y[0] = y0;
  }

  if (valid) {
    auto* Cg = gm(P.C + (int64_t)k * n + l);
    auto* Sg = gm(P.S + (int64_t)k * n + l);
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = lane + 64 * q;
      if (r < nb) {
        yg[r] = y[q];
        Cg[r] = cr[q];
        Sg[r] = sr[q];
      }
    }
    if (lane == 0) {
      gm(P.Ef)[(int64_t)J * m + k] = E;
      if (info) atomicOr(&P.status[k], 2);
    }
  }
  if (J == 0) return;

  // ---- Coefficients of the panel pass (J >= 1, nb = b, full tile) ----
  // W: w_r = c_r prod_{i<r} s_i,  P = prod_{i<b} s_i.
  // Z: t_r = s_r t_{r-1} + c_r y_r (t_{-1} = 0);  Z_0 = -s_0 y_0,
  //    Z_r = c_r t_{r-1} - s_r y_r (1 <= r < b),  Zlast = t_{b-1}
  //    (update-task.c:109-117, :187-198).
  double w[RPL], z[RPL];
  double carry_p = 1.0, carry_t = 0.0;
  double zmax = 0.0;
  // Scans run over segments of 16 rows with carries taken segment by segment
  // in row order -- the association of the packed kernel, so that both
  // layouts produce bit-identical coefficients.
  const int sg16 = lane >> 4, t16 = lane & 15;
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = lane + 64 * q;
    const bool in = r < nb;
    const double sq = in ? sr[q] : 1.0;
    const double cq = in ? cr[q] : 0.0;
    double inc = sq;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const double pv = __shfl_up(inc, off, 16);
      if (t16 >= off) inc = inc * pv;
    }
    double ex = __shfl_up(inc, 1, 16);
    if (t16 == 0) ex = 1.0;
    double a = in ? sq : 1.0, bb = in ? cq * y[q] : 0.0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const double pa = __shfl_up(a, off, 16);
      const double pb = __shfl_up(bb, off, 16);
      if (t16 >= off) {
        bb = a * pb + bb;
        a = a * pa;
      }
    }
    double my_cp = carry_p, my_ct = carry_t;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      if (sg16 == g) {
        my_cp = carry_p;
        my_ct = carry_t;
      }
      const double segp = readlane(inc, 16 * g + 15);
      const double sega = readlane(a, 16 * g + 15);
      const double segb = readlane(bb, 16 * g + 15);
      carry_p *= segp;
      carry_t = sega * carry_t + segb;
    }
    w[q] = cq * (ex * my_cp);
    const double tt = a * my_ct + bb;
    double tprev = __shfl_up(tt, 1, 16);
    if (t16 == 0) tprev = my_ct;
    if (r == 0)
      z[q] = -sq * y[q];
    else
      z[q] = in ? cq * tprev - sq * y[q] : 0.0;
    zmax = fmax(zmax, fabs(z[q]));
  }
  const double Pk = carry_p;
  const double zlast = carry_t;   // rows >= nb are identity maps
  const double zn = fmax(wave_max(zmax), fabs(zlast));

  // Prescale so that the panel product cannot overflow:
  // |acc| <= 2 (||H||_oo + |mu|) ||Z||_oo must stay below 2^1020.
  const double hinf = gm(P.scal)[0];
  const double tmax = 2.0 * (hinf + fabs(mu));
  int p = 0;
  if (zn > 0.0 && tmax > 0.0) p = max(0, ilogb(tmax) + ilogb(zn) + 2 - 1020);

  if (P.ccol) {   // FP32 panel pass: exponents of the scaled [W | Z] columns
    double ew = -1e300, ez = -1e300;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      ew = coef_exp_part(P, l, w[q], lane + 64 * q, nb, ew);
      ez = coef_exp_part(P, l, ldexp(z[q], -p), lane + 64 * q, nb, ez);
    }
    ew = wave_max(ew);
    ez = wave_max(ez);
    const int iw = ew > -1e300 ? (int)ew : 0, iz = ez > -1e300 ? (int)ez : 0;
    if (valid) {
      if (lane == 0) {
        gm(P.Be[J & 1])[2 * k] = iw;
        gm(P.Be[J & 1])[2 * k + 1] = iz;
      }
      auto* Fw = gm(P.B32[J & 1] + (int64_t)(2 * k) * P.b);
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int r = lane + 64 * q;
        if (r < nb) {
          const int cr = gm(P.ccol)[l - 1 + r];
          Fw[r] = (float)ldexp(w[q], -cr - iw);
          Fw[P.b + r] = (float)ldexp(ldexp(z[q], -p), -cr - iz);
        }
      }
    }
  }
  if (valid) {
    auto* Bw = gm(P.Bm[J & 1] + (int64_t)(2 * k) * P.b);
    auto* Bz = Bw + P.b;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = lane + 64 * q;
      if (r < nb) {
        Bw[r] = w[q];
        Bz[r] = ldexp(z[q], -p);
      }
    }
    if (lane == 0) {
      gm(P.aux[J & 1])[4 * k + 0] = Pk;
      gm(P.aux[J & 1])[4 * k + 1] = ldexp(zlast, -p);
      gm(P.aux[J & 1])[4 * k + 2] = zn;
      gm(P.aux[J & 1])[4 * k + 3] = (double)p;
    }
  }
}

template <int RPL>
__global__ __launch_bounds__(256, RPL == 2 ? 4 : 1) void k_diag(const ProbDev* __restrict__ probs,
                                                              const K1Item* __restrict__ items) {
  __shared__ double smem[diag_smem_doubles<RPL>()];
  diag_body<RPL>(probs, items, blockIdx.x, smem);
}

// ---------------------------------------------------------------------------
// K1, packed form: 16 lanes per shift, 4 shifts per wavefront, 16 per
// workgroup.  Same algorithm and arithmetic as k_diag; the per-shift scalar
// chain (rotation, pivot, protections) is executed once for four shifts,
// broadcasts inside a shift's 16-lane DPP row use row_newbcast, so the hot
// loop is unrolled by 16 to give every broadcast a compile-time source lane.
// ---------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ double bc16(double x, int lrt) {
  if constexpr (L < 0)
    return bcast16_rt(x, lrt);
  else
    return bcast16<L>(x);
}

// Three wavefronts per SIMD (at most 168 VGPRs) where the rows per lane allow
// it: the packed sweeps run where many shifts are in flight, and their rate
// is set by how many wavefronts hide each other's latency.
template <int RPL>
__global__ __launch_bounds__(256, RPL <= 8 ? 3 : 1) void k_diag16(const ProbDev* __restrict__ probs,
                                                const K1Item* __restrict__ items) {
  __builtin_amdgcn_s_setprio(2);   // latency-bound: ahead of bulk panel passes
  constexpr int BMAX = 16 * RPL;
  constexpr int CW = (BMAX <= 64) ? 64 : 16;   // LDS chunk columns (registers of the staging)
  __shared__ double Hs[CW * BMAX];
  __shared__ double Bsh[CW * 2 * 16];   // [W | Z] chunk of the 16 shifts

  const K1Item it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int J = it.J;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int grp = lane >> 4, t = lane & 15;
  const int m = P.m, n = P.n;
  const int kraw = it.k0 + wave * 4 + grp;
  const bool valid = kraw < m;
  const int k = valid ? kraw : m - 1;
  const int l = tile_first_row(P, J), nb = tile_rows(P, J);
  const double mu = gsel<false>(P.wr)[k];
  const auto* __restrict__ H = gsel<false>(P.H);
  const int64_t ldH = P.ldH;
  const auto* Rin = gsel<false>(P.Rt[J & 1] + (int64_t)k * n + l);
  auto* yg = gsel<false>(P.Y + (int64_t)k * P.ldY + l);

  constexpr int NL = CW * BMAX / 256;
  double pf[NL];

  // thread t loads row t % BMAX of columns t / BMAX + CSTEP i of the chunk
  constexpr int CSTEP = 256 / BMAX;
  static_assert(256 % BMAX == 0 && CW % CSTEP == 0, "chunk tiling");
  auto issue = [&](int chunk) __attribute__((always_inline)) {
    const int c0 = chunk * CW;
    const int nc = min(CW, nb - c0);
    const int r = threadIdx.x % BMAX, cb = threadIdx.x / BMAX;
    const auto* __restrict__ src = H + (l + r) + (int64_t)(l + c0 + cb) * ldH;
    const int64_t step = (int64_t)CSTEP * ldH;
    if (nc == CW && nb == BMAX) {   // full chunk of a full tile
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = src[i * step];
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) pf[i] = (cb + CSTEP * i < nc && r < nb) ? src[i * step] : 0.0;
    }
  };

  double v[RPL], y[RPL];
  // rotations go to C, S as they are made (the lane owning the row stores
  // them), not kept in registers: 2 RPL registers and the selects of a
  // register-array update per step less
  auto* __restrict__ Cg = gsel<false>(P.C + (int64_t)k * n + l);
  auto* __restrict__ Sg = gsel<false>(P.S + (int64_t)k * n + l);
  int E;   // scale exponent of the tile's y
  if (J + 1 < P.T) {   // rows of this tile after the panel pass of tile column J+1
    own_rows_update<RPL, 16, CW, BMAX, NL, 16, false>(P, J, it.k0, k, wave * 4 + grp, l, nb, t, mu, Hs,
                                               Bsh, pf, v, y, E);
  } else {       // the last tile column starts from the initial vectors
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = t + 16 * q;
      v[q] = r < nb ? Rin[r] : 0.0;
      y[q] = r < nb ? yg[r] : 0.0;
    }
    E = gsel<false>(P.Ep[(J + 1) & 1])[(int64_t)J * m + k];
  }
  const double Hnorm = gsel<false>(P.diagn)[J];
  const double eps = 0x1p-53;
  const double smlnum = fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps));   // s_H_s.c:70-71
  const double smin = fmax(eps * fabs(mu), smlnum);                        // solve-task.c:47
  int info = 0;
  const double slack = 1.0 + 0x1p-40;
  const bool exact = P.force_exact != 0;
  int cur = -1;
  auto column = [&](int col) __attribute__((always_inline)) -> const double* {
    const int ch = col / CW;
    if (ch != cur) {
      __syncthreads();
#pragma unroll
      for (int i = 0; i < NL; ++i) Hs[threadIdx.x + 256 * i] = pf[i];
      __syncthreads();
      if (ch > 0) issue(ch - 1);
      cur = ch;
    }
    return &Hs[(col - ch * CW) * BMAX];
  };

  double By, Bv;
  {
    double ly = 0.0, lv = 0.0;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      ly = fmax(ly, fabs(y[q]));
      lv = fmax(lv, fabs(v[q]));
    }
    By = row16_max(ly);
    Bv = row16_max(lv);
  }

  struct VOut {
    double c, s, r, rinv, br, bv, hdd;
  };
  // v-chain step st (see k_diag), in two parts: the scalar rotation (vscal,
  // returns the rare flag; nothing changed yet) and the vector work (vvec:
  // column st of R into nr, the running column updated in place).  Lc:
  // compile-time lane of row st, or -1.  The y chain likewise (yscal, yvec).
  // Every element gets the operations of k_diag in the same order; with the
  // rare cases decided before any vector work, no step needs copies of v, y
  // or R to fall back on (registers for 3 wavefronts per SIMD).
  // NSc: the row slots that can hold active rows (the first NS; NS = RPL / 2
  // once the sweep is past the middle of the tile, s < 8 RPL).
  auto vscal = [&](auto Lc, auto NSc, int st, const double* hc, bool full, VOut& o)
      __attribute__((always_inline)) -> bool {
    constexpr int L = decltype(Lc)::value;
    constexpr int NS = decltype(NSc)::value;
    const double vs = bc16<L>(pick<RPL, NS>(v, st >> 4), st & 15);
    const double hsub = hc[st];
    const double hdd = hc[st - 1] - mu;
    double c, s, r, rinv;
    bool rare = false;
    if (full) {
      givens(vs, hsub, c, s);
      r = c * vs - s * hsub;
      rinv = 1.0 / r;
    } else {
      const double f1 = fabs(vs), g1 = fabs(hsub);
      rare = !((f1 > 0x1p-511) & (f1 < 0x1.6a09e667f3bcdp+510) & (g1 > 0x1p-511) &
               (g1 < 0x1.6a09e667f3bcdp+510));
      const double d2 = fma(vs, vs, hsub * hsub);
      const double id = rsqrt_nr(d2);
      rinv = copysign(id, vs);
      r = d2 * rinv;
      c = vs * rinv;
      s = -(hsub * rinv);
    }
    const double Hm = fmax(Hnorm, fabs(hdd));
    o.c = c;
    o.s = s;
    o.r = r;
    o.rinv = rinv;
    o.br = fma(fabs(c), Bv, fabs(s) * Hm) * slack;
    o.bv = fma(fabs(c), Hm, fabs(s) * Bv) * slack;
    o.hdd = hdd;
    return rare;
  };
  // slots from NS on hold finished rows: v kept, nr already 0 (the step
  // before set it so: its active rows end below 16 NS)
  auto vvec = [&](auto NSc, int st, const double* hc, const VOut& o, double (&nr)[RPL])
      __attribute__((always_inline)) {
    constexpr int NS = decltype(NSc)::value;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int row = t + 16 * q;
      const bool act = row <= st - 1;
      const double h = (row == st - 1) ? o.hdd : hc[row];
      const double vq = v[q];
      nr[q] = act ? fma(o.c, vq, -(o.s * h)) : 0.0;
      v[q] = act ? fma(o.c, h, o.s * vq) : vq;
    }
  };
  auto yscal = [&](auto Lc, auto NSc, int st, const VOut& o, double& yq, double& byn)
      __attribute__((always_inline)) -> bool {
    constexpr int L = decltype(Lc)::value;
    constexpr int NS = decltype(NSc)::value;
    const double yj = bc16<L>(pick<RPL, NS>(y, st >> 4), st & 15);
    const double acsr = fabs(o.r);
    bool rare = !(acsr >= smin) | ((acsr < 1.0) & (fabs(yj) > acsr * kOmega));
    yq = yj * o.rinv;
    byn = fma(fabs(yq), o.br, By) * slack;
    rare = rare | !(byn <= kOmega);
    return rare;
  };
  // r is 0 on the rows outside 0 .. st-1 (vvec), where the update leaves y
  // as it is (yq is finite on this path): one select per slot, for row st
  auto yvec = [&](auto NSc, int st, double yq, const double (&r)[RPL])
      __attribute__((always_inline)) {
    constexpr int NS = decltype(NSc)::value;
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const int row = t + 16 * q;
      const double u = fma(-yq, r[q], y[q]);
      y[q] = (row == st) ? yq : u;
    }
  };
  auto ystep_slow = [&](int st, const VOut& o, const double (&r)[RPL]) __attribute__((always_inline)) {
    double yj = bcast16_rt(pick<RPL>(y, st >> 4), st & 15);
    double csr = o.r, cinv = o.rinv;
    if (fabs(csr) < smin) {
      csr = smin;
      cinv = 1.0 / smin;
      info = 1;
    }
    const int e1 = pow2_exp(protect_real_division(yj, csr));
    yj = ldexp(yj, e1) * cinv;
    if (e1 != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q)
        if (t + 16 * q != st) y[q] = ldexp(y[q], e1);
      E += e1;
    }
    const double ay = fabs(yj);
    const int kk = (ay > 0.0) ? ilogb(ay) : 0;
    const double ays = ldexp(ay, -kk);
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < RPL; ++q)
      if (t + 16 * q <= st - 1) a = fmax(a, fma(ays, fabs(r[q]), ldexp(fabs(y[q]), -kk)));
    a = row16_max(a);
    int e2 = 0;
    if (a > ldexp(kOmega, -kk)) e2 = -(ilogb(a) + kk) - 1;
    if (e2 != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q) y[q] = ldexp(y[q], e2);
      yj = ldexp(yj, e2);
      E += e2;
    }
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int row = t + 16 * q;
      if (row <= st - 1) y[q] = fma(-yj, r[q], y[q]);
      if (row == st) y[q] = yj;
    }
    By = ldexp(a, kk + e2) * slack;
  };
  auto keep_rotation = [&](int st, const VOut& o) __attribute__((always_inline)) {
    if (valid && t == (st & 15)) {
      Cg[st] = o.c;
      Sg[st] = o.s;
    }
  };

  if (nb >= 2) {
    issue((nb - 2) / CW);
    VOut cv;
    double rr[RPL];   // column of R of the pending y step
    {
      const double* hc = column(nb - 2);
      if (vscal(IC<-1>{}, IC<RPL>{}, nb - 1, hc, false, cv))
        vscal(IC<-1>{}, IC<RPL>{}, nb - 1, hc, true, cv);
      vvec(IC<RPL>{}, nb - 1, hc, cv, rr);
      Bv = cv.bv;
      keep_rotation(nb - 1, cv);
    }
    // Fast step s (y-step s with v-step s-1 interleaved), compile-time lane
    // LY of row s: commits and returns true unless a shift of this wave needs
    // a rare case, in which case nothing changed and the general step runs.
    auto fast_step = [&](auto LYc, auto NSc, int s) __attribute__((always_inline)) -> bool {
      constexpr int LY = decltype(LYc)::value;
      constexpr int LV = (LY + 15) & 15;
      const double* hc = column(s - 2);
      VOut nvo;
      double byn, yq;
      const bool rv = vscal(IC<LV>{}, NSc, s - 1, hc, false, nvo);
      const bool ry = yscal(IC<LY>{}, NSc, s, cv, yq, byn);
      if (__any(rv | ry | exact)) return false;   // nothing changed: the general step runs
      yvec(NSc, s, yq, rr);
      vvec(NSc, s - 1, hc, nvo, rr);   // R column of step s-1 replaces that of step s
      By = byn;
      Bv = nvo.bv;
      cv = nvo;
      keep_rotation(s - 1, cv);
      return true;
    };
    // General step s with runtime lanes: fast or protected per shift.
    auto full_step = [&](int s) __attribute__((always_inline)) {
      const double* hc = column(s - 2);
      VOut nvo;
      double byn, yq;
      if (vscal(IC<-1>{}, IC<RPL>{}, s - 1, hc, false, nvo))
        vscal(IC<-1>{}, IC<RPL>{}, s - 1, hc, true, nvo);
      if (yscal(IC<-1>{}, IC<RPL>{}, s, cv, yq, byn) || exact) {
        ystep_slow(s, cv, rr);
      } else {
        yvec(IC<RPL>{}, s, yq, rr);
        By = byn;
      }
      vvec(IC<RPL>{}, s - 1, hc, nvo, rr);
      Bv = nvo.bv;
      cv = nvo;
      keep_rotation(s - 1, cv);
    };
    int s = nb - 1;
    // 16 fast steps s = 16 B + 15 .. 16 B over the first NS row slots; false
    // at the first that fails (s is then that step)
    auto block = [&](auto NSc) __attribute__((always_inline)) -> bool {
      if (!fast_step(IC<15>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<14>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<13>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<12>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<11>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<10>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<9>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<8>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<7>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<6>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<5>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<4>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<3>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<2>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<1>{}, NSc, s)) return false;
      --s;
      if (!fast_step(IC<0>{}, NSc, s)) return false;
      --s;
      return true;
    };
    while (s >= 2) {
      if ((s & 15) == 15 && s >= 17) {
        // past the middle of the tile the upper half of the slots is done:
        // a second copy of the block updates the lower half only
        const bool ok = s >= 8 * RPL ? block(IC<RPL>{}) : block(IC<RPL / 2>{});
        if (!ok) goto general;
        continue;
      }
    general:
      full_step(s);
      --s;
    }
    {
      double byn, yq;
      if (yscal(IC<1>{}, IC<RPL>{}, 1, cv, yq, byn) || exact) {
        ystep_slow(1, cv, rr);
      } else {
        yvec(IC<RPL>{}, 1, yq, rr);
        By = byn;
      }
    }
  }

  // Final rotation and y(0) (solve-task.c:104-118, :212-218).
  {
    double v0 = bcast16<0>(v[0]);
    if (J > 0) {
      const double hleft = H[l + (int64_t)(l - 1) * ldH];
      double c0 = 1.0, s0 = 0.0;
      if (hleft != 0.0) givens(v0, hleft, c0, s0);
      v0 = v0 * c0 - hleft * s0;
      if (valid && t == 0) {
        Cg[0] = c0;
        Sg[0] = s0;
      }
    } else if (valid && t == 0) {   // tile 0: row 0 is not rotated
      Cg[0] = 1.0;
      Sg[0] = 0.0;
    }
    double y0 = bcast16<0>(y[0]);
    double csr = v0;
    if (fabs(csr) < smin) {
      csr = smin;
      info = 1;
    }
    const int e = pow2_exp(protect_real_division(y0, csr));
    y0 = ldexp(y0, e) / csr;
    if (e != 0) {
#pragma unroll
      for (int q = 0; q < RPL; ++q) y[q] = ldexp(y[q], e);
      E += e;
    }
    if (t == 0) y[0] = y0;
  }

  if (valid) {
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = t + 16 * q;
      if (r < nb) yg[r] = y[q];
    }
    if (t == 0) {
      gsel<false>(P.Ef)[(int64_t)J * m + k] = E;
      if (info) atomicOr(&P.status[k], 2);
    }
  }
  if (J == 0) return;

  // Coefficients of the panel pass (see k_diag), scans over the 16 lanes of
  // each shift's DPP row, carried across the RPL row slots.
  double w[RPL], z[RPL];
  double carry_p = 1.0, carry_t = 0.0;
  double zmax = 0.0;
  // the rotations this wave stored (other lanes of the same shift): visible
  // to its loads once the stores completed (one CU, workgroup scope)
  __threadfence_block();
#pragma unroll
  for (int q = 0; q < RPL; ++q) {
    const int r = t + 16 * q;
    const bool in = r < nb;
    const bool ld = valid && in;   // a shift past m stored nothing: its results are dropped
    const double sq = in ? (ld ? Sg[r] : 1.0) : 1.0;
    const double cq = in ? (ld ? Cg[r] : 0.0) : 0.0;
    double inc = sq;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const double pv = __shfl_up(inc, off, 16);
      if (t >= off) inc = inc * pv;
    }
    double ex = __shfl_up(inc, 1, 16);
    if (t == 0) ex = 1.0;
    w[q] = cq * (ex * carry_p);
    carry_p *= bcast16<15>(inc);
    double a = in ? sq : 1.0, bb = in ? cq * y[q] : 0.0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const double pa = __shfl_up(a, off, 16);
      const double pb = __shfl_up(bb, off, 16);
      if (t >= off) {
        bb = a * pb + bb;
        a = a * pa;
      }
    }
    const double tt = a * carry_t + bb;
    double tprev = __shfl_up(tt, 1, 16);
    if (t == 0) tprev = carry_t;
    if (r == 0)
      z[q] = -sq * y[q];
    else
      z[q] = in ? cq * tprev - sq * y[q] : 0.0;
    carry_t = bcast16<15>(tt);
    zmax = fmax(zmax, fabs(z[q]));
  }
  const double Pk = carry_p;
  const double zlast = carry_t;
  const double zn = fmax(row16_max(zmax), fabs(zlast));
  const double hinf = gsel<false>(P.scal)[0];
  const double tmax = 2.0 * (hinf + fabs(mu));
  int p = 0;
  if (zn > 0.0 && tmax > 0.0) p = max(0, ilogb(tmax) + ilogb(zn) + 2 - 1020);

  if (P.ccol) {   // FP32 panel pass: exponents of the scaled [W | Z] columns
    double ew = -1e300, ez = -1e300;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      ew = coef_exp_part(P, l, w[q], t + 16 * q, nb, ew);
      ez = coef_exp_part(P, l, ldexp(z[q], -p), t + 16 * q, nb, ez);
    }
    ew = row16_max(ew);
    ez = row16_max(ez);
    const int iw = ew > -1e300 ? (int)ew : 0, iz = ez > -1e300 ? (int)ez : 0;
    if (valid) {
      if (t == 0) {
        gsel<false>(P.Be[J & 1])[2 * k] = iw;
        gsel<false>(P.Be[J & 1])[2 * k + 1] = iz;
      }
      auto* Fw = gsel<false>(P.B32[J & 1] + (int64_t)(2 * k) * P.b);
#pragma unroll
      for (int q = 0; q < RPL; ++q) {
        const int r = t + 16 * q;
        if (r < nb) {
          const int cr = gsel<false>(P.ccol)[l - 1 + r];
          Fw[r] = (float)ldexp(w[q], -cr - iw);
          Fw[P.b + r] = (float)ldexp(ldexp(z[q], -p), -cr - iz);
        }
      }
    }
  }
  if (valid) {
    auto* Bw = gsel<false>(P.Bm[J & 1] + (int64_t)(2 * k) * P.b);
    auto* Bz = Bw + P.b;
#pragma unroll
    for (int q = 0; q < RPL; ++q) {
      const int r = t + 16 * q;
      if (r < nb) {
        Bw[r] = w[q];
        Bz[r] = ldexp(z[q], -p);
      }
    }
    if (t == 0) {
      gsel<false>(P.aux[J & 1])[4 * k + 0] = Pk;
      gsel<false>(P.aux[J & 1])[4 * k + 1] = ldexp(zlast, -p);
      gsel<false>(P.aux[J & 1])[4 * k + 2] = zn;
      gsel<false>(P.aux[J & 1])[4 * k + 3] = (double)p;
    }
  }
}

// ---------------------------------------------------------------------------
// K2/K3: panel pass for tile column J on FP64 MFMA (v_mfma_f64_16x16x4_f64).
// Block = 64 rows (one row block, inside one tile row I) x 64 columns of
// [W | Z] interleaved (32 shifts).  Four waves, 32 x 32 each.
// ---------------------------------------------------------------------------
constexpr int KC = 16;          // K chunk staged in LDS
constexpr int GM = kRowBlock;   // 64 rows per block (one row block)
constexpr int GN = kGemmCols;   // 64 columns per block (32 shifts)
constexpr int LDA = GM + 4;
constexpr int LDB = GN + 4;
// FP64 tiles, laid out for 64 LDS banks of 4 bytes (ds_read_b64 served per
// 32 lanes): H tile k-major with ld 80 doubles, so the four k rows read by
// one MFMA step fall into alternating bank halves; [W | Z] tile column-major
// with ld 34 doubles, so a column's 32 k values are written contiguously
// and the 16 columns x 2 k read by 32 lanes hit 64 distinct banks.
constexpr int LDA64 = GM + 16;
constexpr int LDB64 = KC + 2;
constexpr int CT = GN / 32;     // 16-column MFMA tiles per wave

// Block = one 64-row block x 64 columns of [W | Z] (32 shifts); four waves,
// 32 rows x 32 columns each (2 x 2 MFMA tiles).  At most 128 VGPRs: four
// blocks per CU by registers and LDS (the 256-problem batch: 22.0 -> 21.4
// ms), three beside a sweep workgroup.
// LDS of the panel pass, in doubles: the H tiles As[2][KC * LDA64] (As[buf][k
// * LDA64 + row]), the [W | Z] tiles Bs[2][GN * LDB64] (Bs[buf][col * LDB64 +
// k]), the epilogue parameters of the block's shifts (5 x GN / 2 doubles,
// then 2 x GN / 2 ints)
constexpr int kUpdSmemDoubles = 2 * KC * LDA64 + 2 * GN * LDB64 + 5 * (GN / 2) + GN / 2;

// Blocks bid0, bid0 + nblk, ... of the item list (LOOP: grid-stride, so that
// a launch beside a diagonal sweep can be capped to leave room for it on
// every CU; otherwise one block per item).
template <bool LOOP>
__device__ __forceinline__ void update_body(const ProbDev* __restrict__ probs,
                                            const GemmItem* __restrict__ items, int count,
                                            int bid0, int nblk, double* smem) {
  auto As = reinterpret_cast<double(*)[KC * LDA64]>(smem);
  auto Bs = reinterpret_cast<double(*)[GN * LDB64]>(smem + 2 * KC * LDA64);
  // epilogue parameters of the block's shifts (loaded during the K loop)
  double* pmu = smem + 2 * KC * LDA64 + 2 * GN * LDB64;
  double* pP = pmu + GN / 2;
  double* pZl = pP + GN / 2;
  double* pb0w = pZl + GN / 2;
  double* pb0z = pb0w + GN / 2;
  int* psy = reinterpret_cast<int*>(pb0z + GN / 2);
  int* psz = psy + GN / 2;
  // per-column maxima of the epilogue, in the H tile's LDS once the K loop
  // is over (4 blocks of 256 threads fit a CU's LDS)
  double(*red)[GN] = reinterpret_cast<double(*)[GN]>(smem);

#pragma nounroll
  for (int item = bid0; item < count; item += nblk) {
    const GemmItem it = items[item];
    const ProbDev& P = probs[it.prob];
    const int J = it.J;
    const int m = P.m, n = P.n, b = P.b;
    const int l = tile_first_row(P, J);
    int tile, r0, nr;
    rowblock_range(P, it.rb, tile, r0, nr);
    const int col0 = it.cb * GN;
    const int ncols = min(GN, 2 * m - col0);

    // the thread index is re-read per item (an empty asm makes it opaque), so
    // that the index arithmetic derived from it is not held in registers
    // across items
    int tid = threadIdx.x;
    if (LOOP) asm volatile("" : "+v"(tid));
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t ldH = P.ldH;

    const auto* __restrict__ Hp = gsel<false>(P.H + (int64_t)(l - 1) * ldH);   // panel column 0 = l-1
    const auto* __restrict__ Bg = gsel<false>(P.Bm[J & 1]);

    // global -> register staging of one K chunk
    constexpr int NPA = KC * GM / 256, NPB = KC * GN / 256;
    double pa[NPA], pb[NPB];
    auto fetch = [&](int kc) {
#pragma unroll
      for (int i = 0; i < NPA; ++i) {
        const int idx = tid + 256 * i;
        const int kk = idx / GM, rr = idx - kk * GM;
        pa[i] = (rr < nr) ? Hp[(r0 + rr) + (int64_t)(kc + kk) * ldH] : 0.0;
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int idx = tid + 256 * i;
        const int cc = idx / KC, k2 = idx - cc * KC;
        pb[i] = (cc < ncols) ? Bg[(kc + k2) + (int64_t)(col0 + cc) * b] : 0.0;
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int i = 0; i < NPA; ++i) {
        const int idx = tid + 256 * i;
        const int kk = idx / GM, rr = idx - kk * GM;
        As[buf][kk * LDA64 + rr] = pa[i];
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int idx = tid + 256 * i;
        const int cc = idx / KC, k2 = idx - cc * KC;
        Bs[buf][cc * LDB64 + k2] = pb[i];
      }
    };

    // D^T = [W|Z]^T H_panel^T: MFMA operand A = B-matrix values, operand B = H
    // values; the result D^T[col][row] sits in register col/4 of lane
    // row + 16 (col % 4): one lane = one matrix row, four columns per tile.
    double4_t acc[2][CT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = (double4_t){0.0, 0.0, 0.0, 0.0};

    const int nchunks = b / KC;
    fetch(0);
    if (tid < GN / 2) {
      const int k = col0 / 2 + tid;
      if (k < m) {
        const auto* aux = gsel<false>(P.aux[J & 1]);
        const double mu = gsel<false>(P.wr)[k];
        pmu[tid] = mu;
        pP[tid] = aux[4 * k + 0];
        pZl[tid] = aux[4 * k + 1];
        pb0w[tid] = Bg[(int64_t)(2 * k) * b];
        pb0z[tid] = Bg[(int64_t)(2 * k + 1) * b];
        int sy, sz;
        const int ec = update_exponents(P, tile, J, k, mu, sy, sz);
        psy[tid] = sy;
        psz[tid] = sz;
        if (it.rb == tile_first_rowblock(P, tile)) gsel<false>(P.Ep[J & 1])[(int64_t)tile * m + k] = ec;
      }
    }
    stash(0);
    __syncthreads();

    // Epilogue operands (R~ in, y), one 16-column group at a time:
    // acc[a][c][r] covers row wm*32 + a*16 + lane%16 and column
    // wn*GN/2 + c*16 + 4*r + lane/16; even lane groups hold W columns (R~
    // out), odd ones Z columns (y).
    const auto* __restrict__ Rin = gsel<false>(P.Rt[J & 1]);
    auto* __restrict__ Rout = gsel<false>(P.Rt[(J + 1) & 1]);
    const int g = lane >> 4;
    const bool isz = (g & 1) != 0;
    const int rl = lane & 15;
    double rin[4][2], yv[4][2];   // one 16-column group at a time (registers)
    auto load_epilogue = [&](int c) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int cl = wn * (GN / 2) + c * 16 + 4 * r + g;
          const int k = (col0 + cl) >> 1;
          const bool okc = cl < ncols;
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int row = wm * 32 + a * 16 + rl;
            const int gi = r0 + row;
            const bool ok = okc && row < nr;
            rin[r][a] = ok ? Rin[gi + (int64_t)k * n] : 0.0;
            yv[r][a] = (ok && isz) ? gsel<false>(P.Y)[gi + (int64_t)k * P.ldY] : 0.0;
          }
        }
    };
    auto mfma_chunk = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
      for (int k4 = 0; k4 < KC; k4 += 4) {
        const int kr = k4 + (lane >> 4);
        double hv[2], zv[CT];
#pragma unroll
        for (int a = 0; a < 2; ++a) hv[a] = As[buf][kr * LDA64 + wm * 32 + a * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < CT; ++c)
          zv[c] = Bs[buf][(wn * (GN / 2) + c * 16 + (lane & 15)) * LDB64 + kr];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv[c], hv[a], acc[a][c], 0, 0, 0);
      }
    };
    for (int ch = 0; ch + 1 < nchunks; ++ch) {
      const int buf = ch & 1;
      fetch((ch + 1) * KC);
      mfma_chunk(buf);
      stash(buf ^ 1);
      __syncthreads();
    }
    mfma_chunk((nchunks - 1) & 1);
    __syncthreads();   // As is reused for the column maxima below

    // Epilogue, one 16-column group at a time.
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      load_epilogue(c);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int cl = wn * (GN / 2) + c * 16 + 4 * r + g;
        const int ks = cl >> 1;
        const int k = (col0 + cl) >> 1;
        const bool okc = cl < ncols;
        double cmax = 0.0;
        if (okc) {
          const double mu = pmu[ks];
          const double coef = isz ? pZl[ks] : pP[ks];
          const double b0 = isz ? pb0z[ks] : pb0w[ks];
          const int sy = psy[ks], sz = psz[ks];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int row = wm * 32 + a * 16 + rl;
            const int gi = r0 + row;
            if (row < nr) {
              double val = fma(coef, rin[r][a], acc[a][c][r]);
              if (gi == l - 1) val -= mu * b0;
              if (!isz) {
                Rout[gi + (int64_t)k * n] = val;
              } else {
                const double ynew = ldexp(yv[r][a], sy) - ldexp(val, sz);
                gsel<false>(P.Y)[gi + (int64_t)k * P.ldY] = ynew;
                cmax = fmax(cmax, fabs(ynew));
              }
            }
          }
        }
        const double x = row16_max(cmax);
        if (rl == 0) red[wm][cl] = x;
      }
    }
    __syncthreads();
    if (tid < GN) {
      const int cl = tid;
      if (cl < ncols && (cl & 1)) {
        const int k = (col0 + cl) >> 1;
        gsel<false>(P.Ynb[J & 1])[(int64_t)it.rb * m + k] = fmax(red[0][cl], red[1][cl]);
      }
    }
    if (!LOOP) break;
    __syncthreads();   // LDS is reused by the block's next item
  }
}

template <bool LOOP>
__global__ __launch_bounds__(256, 4) void k_update(const ProbDev* __restrict__ probs,
                                                const GemmItem* __restrict__ items, int count) {
  __shared__ double smem[kUpdSmemDoubles];
  update_body<LOOP>(probs, items, count, blockIdx.x, gridDim.x, smem);
}

// One step of a plan whose panel passes overlap the sweeps: blocks [0, nk1)
// run the one-shift sweeps of the step, blocks from nk1pad on (nk1 rounded
// up to the 8 XCDs, so that the pass keeps its XCD-aware block order) the
// panel pass of the step before, capped and looping over its items.  The
// sweep of step s needs the sweep and the panel pass of step s - 1's
// launch, the pass of step s - 1 the same: one stream, no cross-stream
// waits (a dependent launch on one stream starts about 3.5 us after the
// one before, one that also waits for another stream about 8.5 us).
template <int RPL>
__global__ __launch_bounds__(256, 4) void k_step(const ProbDev* __restrict__ probs,
                                              const K1Item* __restrict__ k1, int nk1, int nk1pad,
                                              const GemmItem* __restrict__ gb, int ngb) {
  constexpr int SD = diag_smem_doubles<RPL>();
  __shared__ double smem[SD > kUpdSmemDoubles ? SD : kUpdSmemDoubles];
  if ((int)blockIdx.x < nk1pad) {
    if ((int)blockIdx.x < nk1) diag_body<RPL>(probs, k1, blockIdx.x, smem);
  } else {
    update_body<true>(probs, gb, ngb, blockIdx.x - nk1pad, gridDim.x - nk1pad, smem);
  }
}

// ---------------------------------------------------------------------------
// Panel pass with FP32 MFMA (v_mfma_f32_16x16x4_f32) and FP32 accumulation:
// the optional reduced-precision update path.  Same blocking and epilogue as
// k_update.  The panel and [W | Z] columns are equilibrated by powers of two
// before the conversion to float: H(i, j) 2^(c_j - r_i) (column exponents of
// H, then row exponents, once per solve: k_eq_col, k_eq_row) and B(j, s)
// 2^(-c_j - e_s) (e_s from the sweep that emits the column), so graded
// matrices (entries 2^+-800 and more) keep their small terms; the FP32 result
// is scaled back by 2^(r_i + e_s) and finished in FP64 (cross-over terms,
// shift correction, scaled update of y).  FP32 MFMA
// layout: D[i][j] in register i%4 of lane j + 16 (i/4); here D = D^T[col][row],
// so a lane holds one row and four adjacent columns, W (even) and Z (odd).
// ---------------------------------------------------------------------------
template <bool LOOP>
__global__ __launch_bounds__(256, 4) void k_update32(const ProbDev* __restrict__ probs,
                                                  const GemmItem* __restrict__ items, int count) {
  __shared__ float As[2][KC * LDA];
  __shared__ float Bs[2][KC * LDB];
  __shared__ double pmu[GN / 2], pP[GN / 2], pZl[GN / 2], pb0w[GN / 2], pb0z[GN / 2];
  __shared__ int psy[GN / 2], psz[GN / 2];
  __shared__ double red[2][GN];
  // power-of-two equilibration of the block: column exponents of the panel
  // (b <= 256 columns), row exponents of the column-scaled panel, exponents
  // of the [W | Z] columns scaled by the inverse column exponents
  __shared__ int rrow[GM], ecol[GN];

  // LOOP: grid-stride over the items, so that a launch beside a diagonal
  // sweep can be capped to leave room for it on every CU; otherwise one
  // block per item.
#pragma nounroll
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    const GemmItem it = items[item];
    const ProbDev& P = probs[it.prob];
    const int J = it.J;
    const int m = P.m, n = P.n, b = P.b;
    const int l = tile_first_row(P, J);
    int tile, r0, nr;
    rowblock_range(P, it.rb, tile, r0, nr);
    const int col0 = it.cb * GN;
    const int ncols = min(GN, 2 * m - col0);

    // the thread index is re-read per item (an empty asm makes it opaque), so
    // that the index arithmetic derived from it is not held in registers
    // across items
    int tid = threadIdx.x;
    if (LOOP) asm volatile("" : "+v"(tid));
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1;

    const auto* __restrict__ Bg = gsel<false>(P.Bm[J & 1]);
    // equilibrated float operands (k_eq_conv, the sweep): panel columns
    // l-1 .. l+b-2, the block's rows; [W | Z] columns col0 ..
    const auto* __restrict__ Hf = gsel<false>(P.H32 + (int64_t)(l - 1) * n + r0);
    const auto* __restrict__ Bf = gsel<false>(P.B32[J & 1] + (int64_t)col0 * b);

    constexpr int NPA = KC * GM / 256, NPB = KC * GN / 256;
    float pa[NPA], pb[NPB];
    auto fetch = [&](int kc) {
#pragma unroll
      for (int i = 0; i < NPA; ++i) {
        const int idx = tid + 256 * i;
        const int kk = idx / GM, rr = idx - kk * GM;
        pa[i] = (rr < nr) ? Hf[rr + (int64_t)(kc + kk) * n] : 0.0f;
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int idx = tid + 256 * i;
        const int cc = idx / KC, k2 = idx - cc * KC;
        pb[i] = (cc < ncols) ? Bf[(kc + k2) + (int64_t)cc * b] : 0.0f;
      }
    };
    auto stash = [&](int buf) {
#pragma unroll
      for (int i = 0; i < NPA; ++i) {
        const int idx = tid + 256 * i;
        const int kk = idx / GM, rr = idx - kk * GM;
        As[buf][kk * LDA + rr] = pa[i];
      }
#pragma unroll
      for (int i = 0; i < NPB; ++i) {
        const int idx = tid + 256 * i;
        const int cc = idx / KC, k2 = idx - cc * KC;
        Bs[buf][k2 * LDB + cc] = pb[i];
      }
    };

    float4_t acc[2][CT];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[a][c] = (float4_t){0.f, 0.f, 0.f, 0.f};

    const int nchunks = b / KC;
    if (tid < GN / 2) {
      const int k = col0 / 2 + tid;
      if (k < m) {
        const auto* aux = gsel<false>(P.aux[J & 1]);
        const double mu = gsel<false>(P.wr)[k];
        pmu[tid] = mu;
        pP[tid] = aux[4 * k + 0];
        pZl[tid] = aux[4 * k + 1];
        pb0w[tid] = Bg[(int64_t)(2 * k) * b];
        pb0z[tid] = Bg[(int64_t)(2 * k + 1) * b];
        int sy, sz;
        const int ec = update_exponents(P, tile, J, k, mu, sy, sz);
        psy[tid] = sy;
        psz[tid] = sz;
        if (it.rb == tile_first_rowblock(P, tile)) gsel<false>(P.Ep[J & 1])[(int64_t)tile * m + k] = ec;
      }
    }
    // Scale-back exponents of this block's rows and [W | Z] columns.
    if (tid < GM) rrow[tid] = tid < nr ? gsel<false>(P.rrow)[r0 + tid] : 0;
    if (tid < GN) ecol[tid] = (col0 + tid) < 2 * m ? gsel<false>(P.Be[J & 1])[col0 + tid] : 0;
    __syncthreads();
    fetch(0);
    stash(0);
    __syncthreads();
    for (int ch = 0; ch < nchunks; ++ch) {
      const int buf = ch & 1;
      if (ch + 1 < nchunks) fetch((ch + 1) * KC);
#pragma unroll
      for (int k4 = 0; k4 < KC; k4 += 4) {
        const int kr = k4 + (lane >> 4);
        float hv[2], zv[CT];
#pragma unroll
        for (int a = 0; a < 2; ++a) hv[a] = As[buf][kr * LDA + wm * 32 + a * 16 + (lane & 15)];
#pragma unroll
        for (int c = 0; c < CT; ++c)
          zv[c] = Bs[buf][kr * LDB + wn * (GN / 2) + c * 16 + (lane & 15)];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(zv[c], hv[a], acc[a][c], 0, 0, 0);
      }
      if (ch + 1 < nchunks) stash(buf ^ 1);
      __syncthreads();
    }

    // Epilogue.  acc[a][c][r]: row wm*32 + a*16 + lane%16,
    //                          col wn*GN/2 + c*16 + 4*(lane/16) + r.
    const auto* __restrict__ Rin = gsel<false>(P.Rt[J & 1]);
    auto* __restrict__ Rout = gsel<false>(P.Rt[(J + 1) & 1]);
    const int g = lane >> 4;
    const int rl = lane & 15;
    double rin[CT][4][2], yv[CT][4][2];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int cl = wn * (GN / 2) + c * 16 + 4 * g + r;
        const int k = (col0 + cl) >> 1;
        const bool okc = cl < ncols;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          const int row = wm * 32 + a * 16 + rl;
          const int gi = r0 + row;
          const bool ok = okc && row < nr;
          rin[c][r][a] = ok ? Rin[gi + (int64_t)k * n] : 0.0;
          yv[c][r][a] = (ok && (r & 1)) ? gsel<false>(P.Y)[gi + (int64_t)k * P.ldY] : 0.0;
        }
      }
    double cmaxs[CT][2];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) cmaxs[c][h] = 0.0;
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const bool isz = (r & 1) != 0;
        const int cl = wn * (GN / 2) + c * 16 + 4 * g + r;
        const int ks = cl >> 1;
        const int k = (col0 + cl) >> 1;
        const bool okc = cl < ncols;
        if (okc) {
          const double mu = pmu[ks];
          const double coef = isz ? pZl[ks] : pP[ks];
          const double b0 = isz ? pb0z[ks] : pb0w[ks];
          const int sy = psy[ks], sz = psz[ks];
          const int es = ecol[cl];
#pragma unroll
          for (int a = 0; a < 2; ++a) {
            const int row = wm * 32 + a * 16 + rl;
            const int gi = r0 + row;
            if (row < nr) {
              double val = fma(coef, rin[c][r][a], ldexp((double)acc[a][c][r], es + rrow[row]));
              if (gi == l - 1) val -= mu * b0;
              if (!isz) {
                Rout[gi + (int64_t)k * n] = val;
              } else {
                const double ynew = ldexp(yv[c][r][a], sy) - ldexp(val, sz);
                gsel<false>(P.Y)[gi + (int64_t)k * P.ldY] = ynew;
                cmaxs[c][r >> 1] = fmax(cmaxs[c][r >> 1], fabs(ynew));
              }
            }
          }
        }
      }
    // per-column maxima of |y| over the block's 64 rows (Z columns only)
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const double x = row16_max(cmaxs[c][h]);
        const int cl = wn * (GN / 2) + c * 16 + 4 * g + 2 * h + 1;
        if (rl == 0) red[wm][cl] = x;
      }
    __syncthreads();
    if (tid < GN) {
      const int cl = tid;
      if (cl < ncols && (cl & 1)) {
        const int k = (col0 + cl) >> 1;
        gsel<false>(P.Ynb[J & 1])[(int64_t)it.rb * m + k] = fmax(red[0][cl], red[1][cl]);
      }
    }
    if (!LOOP) break;
    __syncthreads();   // LDS is reused by the block's next item
  }
}

// ---------------------------------------------------------------------------
// K5: back-transform x = Q^T y with tiles brought to the common (smallest)
// scale, three-accumulator 2-norm, convergence test and normalisation
// (shifted_Hessenberg_solve.c:104-245).  The rotation chain
//   tau1_i = s_i tau1_{i-1} + c_i tau2_i,  x_{i-1} = c_i tau1_{i-1} - s_i tau2_i
// is an affine recurrence.  One workgroup per shift walks the vector in
// segments of 1024 elements: y (scaled), c and s are staged coalesced in
// LDS, every thread composes the maps of its 4 elements, a workgroup scan
// (carried across segments) gives each thread its entering value, and the
// outputs leave through LDS as coalesced stores.  HBM traffic per element:
// y, c, s read once, x written once, then read and written once more by the
// normalisation.
// ---------------------------------------------------------------------------
constexpr int BT_PER = 4;              // elements per thread and segment
constexpr int BT_SEG = 256 * BT_PER;   // elements per segment
constexpr int BT_LDS = BT_SEG + BT_SEG / BT_PER;
__device__ __forceinline__ int bt_pad(int e) { return e + e / BT_PER; }   // no bank conflicts

__global__ __launch_bounds__(256) void k_backtransform(const ProbDev* __restrict__ probs,
                                                       const ShiftItem* __restrict__ items) {
  __shared__ double ly[BT_LDS], lc[BT_LDS], ls[BT_LDS], lx[BT_LDS];
  __shared__ int emin_s[4];
  __shared__ double scanA[4], scanB[4];
  __shared__ double accs[3][4];
  __shared__ double ynrm_s;

  const ShiftItem it = items[blockIdx.x];
  const ProbDev& P = probs[it.prob];
  const int k = it.k;
  const int n = P.n, m = P.m, T = P.T;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

  // smallest exponent over the tiles (reduce_scaling_factors, robust.c:58-82)
  int em = 0x7fffffff;
  for (int t = tid; t < T; t += 256) em = min(em, gm(P.Ef)[(int64_t)t * m + k]);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) em = min(em, __shfl_xor(em, off, 64));
  if (lane == 0) emin_s[wave] = em;
  __syncthreads();
  const int emin = min(min(emin_s[0], emin_s[1]), min(emin_s[2], emin_s[3]));

  auto* x = gm(P.Y + (int64_t)k * P.ldY);
  const auto* c = gm(P.C + (int64_t)k * n);
  const auto* s = gm(P.S + (int64_t)k * n);
  const int* Ek = P.Ef + k;

  double carry = 0.0;   // tau1 entering the segment
  Acc3 acc = {0.0, 0.0, 0.0};
  for (int g0 = 0; g0 < n; g0 += BT_SEG) {
    // stage; elements beyond n are identity maps (a = 1, b = 0)
    for (int e = tid; e < BT_SEG; e += 256) {
      const int i = g0 + e;
      double yv = 0.0, cv = 1.0, sv = 1.0;
      if (i < n) {
        const int t = (i < P.b0) ? 0 : 1 + (i - P.b0) / P.b;
        yv = ldexp(x[i], emin - Ek[(int64_t)t * m]);
        cv = c[i];
        sv = s[i];
      }
      ly[bt_pad(e)] = yv;
      lc[bt_pad(e)] = cv;
      ls[bt_pad(e)] = sv;
    }
    __syncthreads();
    // composition of this thread's maps
    const int e0 = tid * BT_PER;
    double A = 1.0, B = 0.0;
#pragma unroll
    for (int j = 0; j < BT_PER; ++j) {
      const int e = bt_pad(e0 + j);
      const bool first = (g0 + e0 + j) == 0;
      const double a = first ? 0.0 : ls[e];
      const double bb = first ? ly[e] : lc[e] * ly[e];
      B = a * B + bb;
      A = a * A;
    }
    double a = A, bb = B;
    wave_affine_scan(a, bb);
    if (lane == 63) {
      scanA[wave] = a;
      scanB[wave] = bb;
    }
    double ea = __shfl_up(a, 1, 64), eb = __shfl_up(bb, 1, 64);
    if (lane == 0) {
      ea = 1.0;
      eb = 0.0;
    }
    __syncthreads();
    double pa = 1.0, pb = 0.0;   // composite of the previous waves
    for (int w = 0; w < wave; ++w) {
      pb = scanA[w] * pb + scanB[w];
      pa = scanA[w] * pa;
    }
    double ta = 1.0, tb = 0.0;   // composite of the whole segment
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      tb = scanA[w] * tb + scanB[w];
      ta = scanA[w] * ta;
    }
    double tau1 = ea * (pa * carry + pb) + eb;   // value entering this thread's chunk
    // the chain itself
#pragma unroll
    for (int j = 0; j < BT_PER; ++j) {
      const int i = g0 + e0 + j;
      if (i < n) {
        const int e = bt_pad(e0 + j);
        const double t2 = ly[e];
        if (i == 0) {
          tau1 = t2;
        } else {
          const double ci = lc[e], si = ls[e];
          const double xi = ci * tau1 - si * t2;
          tau1 = ci * t2 + si * tau1;
          if (e0 + j == 0)
            x[i - 1] = xi;   // last element of the previous segment
          else
            lx[bt_pad(e0 + j - 1)] = xi;
          acc3_add(acc, xi);
        }
        if (i == n - 1) {
          x[n - 1] = tau1;
          acc3_add(acc, tau1);
        }
      }
    }
    carry = ta * carry + tb;
    __syncthreads();
    // x_{g0+e} for e = 0 .. BT_SEG-2 was produced at element e+1
    for (int e = tid; e < BT_SEG - 1; e += 256) {
      const int i = g0 + e;
      if (i + 1 < n) x[i] = lx[bt_pad(e)];
    }
  }
  const double s0 = wave_sum(acc.asml), s1 = wave_sum(acc.amed), s2 = wave_sum(acc.abig);
  if (lane == 0) {
    accs[0][wave] = s0;
    accs[1][wave] = s1;
    accs[2][wave] = s2;
  }
  __syncthreads();
  if (tid == 0) {
    Acc3 t;
    t.asml = (accs[0][0] + accs[0][1]) + (accs[0][2] + accs[0][3]);
    t.amed = (accs[1][0] + accs[1][1]) + (accs[1][2] + accs[1][3]);
    t.abig = (accs[2][0] + accs[2][1]) + (accs[2][2] + accs[2][3]);
    const double ynrm = acc3_norm(t);
    // converged iff the unscaled growth ||x|| 2^-emin exceeds 0.1/sqrt(n)
    const bool conv = isfinite(ynrm) && ynrm > ldexp(0.1 / sqrt((double)n), emin);
    ynrm_s = conv ? ynrm : -1.0;
    // status: bit 0 converged; bits 8-23 the binary orders of downscaling the
    // vector carried into this back-transform (-emin, exponents are <= 0)
    atomicOr(&P.status[k], (conv ? 1 : 0) | (min(-emin, 0xFFFF) << 8));
  }
  __syncthreads();
  const double ynrm = ynrm_s;
  if (ynrm > 0.0)
    for (int i = tid; i < n; i += 256) x[i] = x[i] / ynrm;
}


// ---------------------------------------------------------------------------
// Absolute row sums of H (its 1-norms per row: ||H||_oo and the row spread
// of the FP32 mode, the complex path's eps3).  The residual validator is in
// hsrq_gemm.hip.
// ---------------------------------------------------------------------------
// ||H(i,:)||_1 per row in two passes: partial sums over 16 column slabs x 4
// (64 rows x 4 column groups per workgroup, coalesced down the columns),
// then a fixed-order sum.  Only the Hessenberg part (j >= i - 1) is read.
constexpr int kRowSlabs = 16;
__global__ __launch_bounds__(256) void k_rowsum_part(int n, const double* __restrict__ H,
                                                     int64_t ldH, double* __restrict__ part) {
  const int i = blockIdx.x * 64 + (threadIdx.x & 63);
  const int g = threadIdx.x >> 6;
  const int slab = blockIdx.y;
  const int w = (n + kRowSlabs - 1) / kRowSlabs;
  const int c0 = slab * w, c1 = min(n, c0 + w);
  double a = 0.0;
  if (i < n)
    for (int j = max(c0, i - 1) + g; j < c1; j += 4) a += fabs(H[i + (int64_t)j * ldH]);
  if (i < n) part[(int64_t)(slab * 4 + g) * n + i] = a;
}

__global__ __launch_bounds__(256) void k_rowsum_final(int n, const double* __restrict__ part,
                                                      double* __restrict__ rs) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    double a = 0.0;
    for (int t = 0; t < 4 * kRowSlabs; ++t) a += part[(int64_t)t * n + i];
    rs[i] = a;
  }
}

__device__ double block_max256(double v, double* sh) {
  const int wave = threadIdx.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[wave] = v;
  __syncthreads();
  return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

__device__ double block_sum256(double v, double* sh) {
  const int wave = threadIdx.x >> 6;
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[wave] = v;
  __syncthreads();
  return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Scaled 2-norm of a length-n vector by one workgroup.
__device__ double block_nrm2(int n, const double* __restrict__ v, double* sh) {
  double mx = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) mx = fmax(mx, fabs(v[i]));
  mx = block_max256(mx, sh);
  if (mx == 0.0 || !isfinite(mx)) return mx;
  const int e = ilogb(mx);
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    const double t = ldexp(v[i], -e);
    s += t * t;
  }
  s = block_sum256(s, sh);
  return ldexp(sqrt(s), e);
}

// ---------------------------------------------------------------------------
// MFMA layout self-test: D = A (16x4) * B (4x16), A[i][k] = i + 100 k,
// B[k][j] = (k == j % 4) ? 1 : 0  => D[i][j] = A[i][j % 4].  Operands: lane
// i + 16 k holds A[i][k] and B[k][i]; result: see k_update's epilogue.
// ---------------------------------------------------------------------------
__global__ void k_mfma_selftest(int* __restrict__ bad) {
  const int lane = threadIdx.x;
  const double av = (double)(lane & 15) + 100.0 * (lane >> 4);
  const double bv = ((lane >> 4) == ((lane & 15) & 3)) ? 1.0 : 0.0;
  double4_t acc = (double4_t){0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
  int wrong = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int i = 4 * r + (lane >> 4), j = lane & 15;
    const double want = (double)i + 100.0 * (j & 3);
    if (acc[r] != want) wrong = 1;
  }
  if (wrong) atomicAdd(bad, 1);
}

}  // namespace hsrq

// ---------------------------------------------------------------------------
// Launch wrappers (host side of this translation unit).
// ---------------------------------------------------------------------------
namespace hsrq {

hipError_t launch_norms(const ProbDev* d_probs, const GemmItem* part_items, int part_count,
                        const RowItem* items, int count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_norm_part, dim3(part_count), dim3(256), 0, st, d_probs, part_items);
  hipLaunchKernelGGL(k_norm_final, dim3(count), dim3(256), 0, st, d_probs, items);
  return hipGetLastError();
}

bool norm_block_needed_host(const ProbDev& P, int rb, int cb) { return norm_block_needed(P, rb, cb); }

hipError_t launch_equilibrate(const ProbDev* d_prob, int n, int nrb, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  // the row pass keeps its partial maxima in the norms' part buffer (nrb x n)
  const int chunks = nrb < 16 ? nrb : 16;
  hipLaunchKernelGGL(k_eq_col, dim3((n + 3) / 4), dim3(256), 0, st, d_prob);
  hipLaunchKernelGGL(k_eq_row, dim3((n + 63) / 64, chunks), dim3(256), 0, st, d_prob);
  hipLaunchKernelGGL(k_eq_rowfin, dim3((n + 255) / 256), dim3(256), 0, st, d_prob, chunks);
  hipLaunchKernelGGL(k_eq_conv, dim3((n + 3) / 4), dim3(256), 0, st, d_prob);
  return hipGetLastError();
}

hipError_t launch_init(const ProbDev* d_probs, const RowItem* items, int count, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_init, dim3(count), dim3(256), 0, st, d_probs, items);
  return hipGetLastError();
}

hipError_t launch_diag(int rpl, const ProbDev* d_probs, const K1Item* items, int count,
                       hipStream_t st) {
  if (count <= 0) return hipSuccess;
  switch (rpl) {
    case 1: hipLaunchKernelGGL(k_diag<1>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    case 2: hipLaunchKernelGGL(k_diag<2>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    case 4: hipLaunchKernelGGL(k_diag<4>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_diag16(int rpl, const ProbDev* d_probs, const K1Item* items, int count,
                         hipStream_t st) {
  if (count <= 0) return hipSuccess;
  switch (rpl) {
    case 4: hipLaunchKernelGGL(k_diag16<4>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    case 8: hipLaunchKernelGGL(k_diag16<8>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    case 16: hipLaunchKernelGGL(k_diag16<16>, dim3(count), dim3(256), 0, st, d_probs, items); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

bool step_fusable(int pack, int b) { return pack == 1 && b <= 128; }

hipError_t launch_step(int b, const ProbDev* d_probs, const K1Item* k1, int nk1,
                       const GemmItem* gb, int ngb, int max_blocks, hipStream_t st) {
  const int nk1pad = (nk1 + 7) / 8 * 8;
  const int npass = ngb > 0 ? (max_blocks > 0 && max_blocks < ngb ? max_blocks : ngb) : 0;
  if (nk1pad + npass == 0) return hipSuccess;
  const dim3 grid(nk1pad + npass);
  if (b <= 64)
    hipLaunchKernelGGL(k_step<1>, grid, dim3(256), 0, st, d_probs, k1, nk1, nk1pad, gb, ngb);
  else if (b <= 128)
    hipLaunchKernelGGL(k_step<2>, grid, dim3(256), 0, st, d_probs, k1, nk1, nk1pad, gb, ngb);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_sweep(int pack, int b, const ProbDev* d_probs, const K1Item* items, int count,
                        hipStream_t st) {
  if (pack == 4) return launch_diag16(b / 16, d_probs, items, count, st);
  return launch_diag(b <= 64 ? 1 : (b <= 128 ? 2 : 4), d_probs, items, count, st);
}

hipError_t launch_update(const ProbDev* d_probs, const GemmItem* items, int count,
                         hipStream_t st, int max_blocks) {
  if (count <= 0) return hipSuccess;
  if (max_blocks > 0 && max_blocks < count)
    hipLaunchKernelGGL(k_update<true>, dim3(max_blocks), dim3(256), 0, st, d_probs, items, count);
  else
    hipLaunchKernelGGL(k_update<false>, dim3(count), dim3(256), 0, st, d_probs, items, count);
  return hipGetLastError();
}

hipError_t launch_update32(const ProbDev* d_probs, const GemmItem* items, int count,
                           hipStream_t st, int max_blocks) {
  if (count <= 0) return hipSuccess;
  if (max_blocks > 0 && max_blocks < count)
    hipLaunchKernelGGL(k_update32<true>, dim3(max_blocks), dim3(256), 0, st, d_probs, items,
                       count);
  else
    hipLaunchKernelGGL(k_update32<false>, dim3(count), dim3(256), 0, st, d_probs, items, count);
  return hipGetLastError();
}

hipError_t launch_backtransform(const ProbDev* d_probs, const ShiftItem* items, int count,
                                hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_backtransform, dim3(count), dim3(256), 0, st, d_probs, items);
  return hipGetLastError();
}

hipError_t launch_rowsums(int n, const double* H, int64_t ldH, double* work, double* rs,
                          hipStream_t st) {
  double* part = work;   // 4 kRowSlabs x n
  hipLaunchKernelGGL(k_rowsum_part, dim3((n + 63) / 64, kRowSlabs), dim3(256), 0, st, n, H, ldH,
                     part);
  hipLaunchKernelGGL(k_rowsum_final, dim3((n + 255) / 256), dim3(256), 0, st, n,
                     (const double*)part, rs);
  return hipGetLastError();
}

hipError_t launch_mfma_selftest(int* d_bad, hipStream_t st) {
  hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, st, d_bad);
  return hipGetLastError();
}

}  // namespace hsrq
