// hsrq_gemm.hip -- FP64 MFMA products around the solve, for many problems in
// one launch (item tables, as the solve's kernels):
//
//   mode 0 (back-transformation, the step after the solve):  C = A B with A
//     n x n (Z of A0 = Z H Z^T) and B n x m (the eigenvectors of H);
//   mode 1 (residual validator, validation.c:79-171 at scale):
//     R = H X - X diag(lam) for an upper Hessenberg H, reduced on the fly to
//     per-column Blue accumulators (LAPACK 3.10 dnrm2's three sums of
//     squares) per 128-row panel; a final pass per eigenvector combines the
//     panels and divides by ||H||_oo ||x||_2 eps.  Only the Hessenberg part
//     of H is read (row i from column i-1 on), as the product with the zeros
//     below the subdiagonal that the reference's dense dgemm adds
//     (validation.c:48-77) contributes nothing.
//
// Tiling: a workgroup computes a 128 x 64 block of the product; four waves
// of 64 x 32 (4 x 2 v_mfma_f64_16x16x4_f64 tiles each, 8 independent
// accumulators); K in chunks of 16 staged global -> registers -> LDS, double
// buffered, the next chunk's loads in flight during the current chunk's
// MFMAs.  Blocks are ordered so that the column blocks of one row panel run
// on one XCD at the same time (the 8 XCDs take consecutive blocks round
// robin): the panel's A chunks are then fetched once into that XCD's L2.
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "hsrq_device.h"
#include "hsrq_gemm.h"

namespace hsrq {

typedef double double4_t __attribute__((ext_vector_type(4)));

namespace {

constexpr int BM = kGemmPanelRows;   // 128
constexpr int BK = 16;
// LDS layouts (banks of 4 bytes, a double covers two): A chunk k-major with
// ld 144 (= 16 mod 32 doubles: the 16-lane groups of one MFMA step read
// disjoint bank halves), B chunk column-major with ld 18 (16 columns x 2 k of
// 32 lanes fall on 32 distinct bank pairs).
constexpr int LDA = BM + 16;
constexpr int LDB = BK + 2;

// BN = 32 CT columns per block: waves of 64 rows x 16 CT columns (4 x CT
// MFMA tiles).  Both products take CT = 2 (128 x 64 blocks); 128 x 128
// blocks measured slower for the back-transformation at n = 8000, m = 800
// (3.24 against 2.99 ms: 441 blocks do not fill two per CU).
template <int MODE, int CT>
__global__ __launch_bounds__(256) void k_gemm(const GemmDesc* __restrict__ descs,
                                              const GemmTileItem* __restrict__ items) {
  constexpr int BN = 32 * CT;
  __shared__ double As[2][BK * LDA];
  __shared__ double Bs[2][BN * LDB];
  __shared__ double red[3][2][MODE == 1 ? BN : 1];   // mode 1: the two row halves
  __shared__ double rsh[MODE == 1 ? 256 : 1];         // mode 1: partial row 1-norms

  const GemmTileItem it = items[blockIdx.x];
  const GemmDesc& D = descs[it.prob];
  const int n = D.n, m = D.m;
  const int r0 = it.panel * BM, c0 = it.cb * BN;
  const int nr = min(BM, n - r0), nc = min(BN, m - c0);
  const int kbeg = D.hess ? max(0, r0 - 1) : 0;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const double* __restrict__ A = D.A;
  const double* __restrict__ B = D.B;
  const int64_t ldA = D.ldA, ldB = D.ldB;

  constexpr int NPA = BK * BM / 256, NPB = BK * BN / 256;
  double pa[NPA], pb[NPB];
  // mode 1: ||H(i,:)||_1 of the thread's row (tid % 128: the same for every
  // element the thread fetches), over its half of the columns, in fetch order
  double rsum = 0.0;
  auto fetch = [&](int k0) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NPA; ++i) {
      const int idx = tid + 256 * i;
      const int kk = idx / BM, rr = idx - kk * BM;
      const int k = k0 + kk;
      // (Hessenberg: nothing left of the subdiagonal is read)
      const bool in = rr < nr && k < n && (!D.hess || k >= r0 + rr - 1);
      pa[i] = in ? A[(r0 + rr) + (int64_t)k * ldA] : 0.0;
      if (MODE == 1) rsum += fabs(pa[i]);
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) {
      const int idx = tid + 256 * i;
      const int cc = idx / BK, k2 = idx - cc * BK;
      const int k = k0 + k2;
      pb[i] = (cc < nc && k < n) ? B[k + (int64_t)(c0 + cc) * ldB] : 0.0;
    }
  };
  auto stash = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < NPA; ++i) {
      const int idx = tid + 256 * i;
      const int kk = idx / BM, rr = idx - kk * BM;
      As[buf][kk * LDA + rr] = pa[i];
    }
#pragma unroll
    for (int i = 0; i < NPB; ++i) {
      const int idx = tid + 256 * i;
      const int cc = idx / BK, k2 = idx - cc * BK;
      Bs[buf][cc * LDB + k2] = pb[i];
    }
  };

  // acc[a][c][r]: row wm*64 + a*16 + lane%16, column wn*BN/2 + c*16 + 4r + lane/16
  double4_t acc[4][CT];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < CT; ++c) acc[a][c] = (double4_t){0.0, 0.0, 0.0, 0.0};
  auto mfma_chunk = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int k4 = 0; k4 < BK; k4 += 4) {
      const int kr = k4 + (lane >> 4);
      double hv[4], zv[CT];
#pragma unroll
      for (int a = 0; a < 4; ++a) hv[a] = As[buf][kr * LDA + wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
      for (int c = 0; c < CT; ++c)
        zv[c] = Bs[buf][(wn * (BN / 2) + c * 16 + (lane & 15)) * LDB + kr];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < CT; ++c)
          acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(zv[c], hv[a], acc[a][c], 0, 0, 0);
    }
  };
  const int nchunks = (n - kbeg + BK - 1) / BK;
  fetch(kbeg);
  stash(0);
  __syncthreads();
  for (int ch = 0; ch + 1 < nchunks; ++ch) {
    const int buf = ch & 1;
    fetch(kbeg + (ch + 1) * BK);
    mfma_chunk(buf);
    stash(buf ^ 1);
    __syncthreads();
  }
  mfma_chunk((nchunks - 1) & 1);

  const int rl = lane & 15, g = lane >> 4;
  if (MODE == 0) {
    double* __restrict__ C = D.C;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = wm * 64 + a * 16 + rl;
          const int col = wn * (BN / 2) + c * 16 + 4 * r + g;
          if (row < nr && col < nc) C[(r0 + row) + (int64_t)(c0 + col) * D.ldC] = acc[a][c][r];
        }
    return;
  }
  // mode 1: r = (H x)_i - lam x_i; Blue accumulators per column over the
  // block's rows: per thread over its 4 rows, then over the 16 lanes of its
  // column group (fixed order), then the two row halves through LDS.
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int col = wn * (BN / 2) + c * 16 + 4 * r + g;
      const bool okc = col < nc;
      const double lam = okc ? D.lam[c0 + col] : 0.0;
      Acc3 s = {0.0, 0.0, 0.0};
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int row = wm * 64 + a * 16 + rl;
        if (okc && row < nr) {
          const double x = B[(r0 + row) + (int64_t)(c0 + col) * ldB];
          acc3_add(s, fma(-lam, x, acc[a][c][r]));
        }
      }
#pragma unroll
      for (int off = 8; off >= 1; off >>= 1) {
        s.asml += __shfl_xor(s.asml, off, 16);
        s.amed += __shfl_xor(s.amed, off, 16);
        s.abig += __shfl_xor(s.abig, off, 16);
      }
      if (rl == 0) {
        red[0][wm][col] = s.asml;
        red[1][wm][col] = s.amed;
        red[2][wm][col] = s.abig;
      }
    }
  rsh[tid] = rsum;
  __syncthreads();
  if (tid < nc) {
    double* __restrict__ part = D.part + 3 * ((int64_t)it.panel * m + c0 + tid);
#pragma unroll
    for (int q = 0; q < 3; ++q) part[q] = red[q][0][tid] + red[q][1][tid];
  }
  // the panel's largest row 1-norm (the first column block of the panel
  // writes it; every block reads the same rows of H)
  if (it.cb == 0 && tid < 64) {
    double mx = 0.0;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int row = tid + 64 * h;
      if (row < nr) mx = fmax(mx, rsh[row] + rsh[row + 128]);
    }
    mx = wave_max(mx);
    if (tid == 0) D.rsmax[it.panel] = mx;
  }
}

// One eigenvector per workgroup: ||r||_2 from the panels' accumulators (in
// panel order), ||x||_2 (Blue), ||H||_oo from the row-block maxima.
__global__ __launch_bounds__(256) void k_resid_final(const GemmDesc* __restrict__ descs,
                                                     const GemmTileItem* __restrict__ items) {
  __shared__ double sh[3][4];
  const GemmTileItem it = items[blockIdx.x];
  const GemmDesc& D = descs[it.prob];
  const int n = D.n, k = it.cb;   // cb: the eigenvector
  const int tid = threadIdx.x, wave = tid >> 6;
  const double* __restrict__ x = D.B + (int64_t)k * D.ldB;
  Acc3 s = {0.0, 0.0, 0.0};
  for (int i = tid; i < n; i += 256) acc3_add(s, x[i]);
  s.asml = wave_sum(s.asml);
  s.amed = wave_sum(s.amed);
  s.abig = wave_sum(s.abig);
  if ((tid & 63) == 0) {
    sh[0][wave] = s.asml;
    sh[1][wave] = s.amed;
    sh[2][wave] = s.abig;
  }
  __syncthreads();
  if (tid == 0) {
    Acc3 xs;
    xs.asml = (sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]);
    xs.amed = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    xs.abig = (sh[2][0] + sh[2][1]) + (sh[2][2] + sh[2][3]);
    Acc3 rs = {0.0, 0.0, 0.0};
    const int np = (n + BM - 1) / BM;
    for (int p = 0; p < np; ++p) {
      const double* q = D.part + 3 * ((int64_t)p * D.m + k);
      rs.asml += q[0];
      rs.amed += q[1];
      rs.abig += q[2];
    }
    double hmax = 0.0;
    for (int p = 0; p < np; ++p) hmax = fmax(hmax, D.rsmax[p]);
    D.res[k] = acc3_norm(rs) / (hmax * acc3_norm(xs) * 0x1p-52);
  }
}

}  // namespace

// Block order of one product: the column blocks of row panel p run together
// on XCD p % 8.  Returns the items (invalid slots of a partial last round of
// panels are not emitted: the XCD assignment of the others shifts, which
// costs only locality).
void add_gemm_items(int prob, int n, int m, int bn, std::vector<GemmTileItem>& out) {
  const int np = (n + BM - 1) / BM, ncb = (m + bn - 1) / bn;
  for (int p0 = 0; p0 < np; p0 += 8) {
    const int pe = p0 + 8 < np ? p0 + 8 : np;
    for (int cb = 0; cb < ncb; ++cb)
      for (int p = p0; p < pe; ++p) out.push_back(GemmTileItem{prob, p, cb, 0});
  }
}

void add_resid_final_items(int prob, int m, std::vector<GemmTileItem>& out) {
  for (int k = 0; k < m; ++k) out.push_back(GemmTileItem{prob, 0, k, 0});
}

size_t resid_workspace_doubles(int n, int m) {
  const size_t np = (n + BM - 1) / BM;
  return 3 * np * (size_t)m + np + (size_t)m;
}

void resid_describe(GemmDesc& D, int n, const double* H, int64_t ldH, int m, const double* wr,
                    const double* X, int64_t ldX, double* ws) {
  D = GemmDesc{};
  D.A = H;
  D.ldA = ldH;
  D.B = X;
  D.ldB = ldX;
  D.lam = wr;
  D.n = n;
  D.m = m;
  D.hess = 1;
  const size_t np = (n + BM - 1) / BM;
  D.part = ws;
  D.rsmax = ws + 3 * np * (size_t)m;
  D.res = D.rsmax + np;
}

hipError_t launch_gemm(const GemmDesc* d_descs, const GemmTileItem* items, int count,
                       hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_gemm<0, 2>), dim3(count), dim3(256), 0, st, d_descs, items);
  return hipGetLastError();
}

hipError_t launch_resid(const GemmDesc* d_descs, const GemmTileItem* tiles, int n_tiles,
                        const GemmTileItem* fin, int n_fin, hipStream_t st) {
  if (n_fin <= 0) return hipSuccess;
  hipLaunchKernelGGL((k_gemm<1, 2>), dim3(n_tiles), dim3(256), 0, st, d_descs, tiles);
  hipLaunchKernelGGL(k_resid_final, dim3(n_fin), dim3(256), 0, st, d_descs, fin);
  return hipGetLastError();
}

}  // namespace hsrq
