// hsrq_internal.h -- device problem descriptors and work items shared by the
// kernels and the host planner.
#pragma once
#include <stdint.h>

namespace hsrq {

// Rows handled per GEMM block / granularity of the per-row-block norm
// bookkeeping.  Tile sizes are multiples of it.
constexpr int kRowBlock = 64;
// Shifts per diagonal-sweep workgroup (one wavefront each).
constexpr int kShiftsPerK1 = 4;
// Output columns per GEMM block (kGemmCols / 2 shifts: W and Z columns) and
// row blocks per GEMM block.
constexpr int kGemmCols = 64;
constexpr int kGemmRowBlocks = 1;

// One problem: H (n x n, column-major, ldH), m real shifts, tile size b.
// Rows are tiled with the FIRST tile ragged (b0 = n - (T-1) b), as the
// reference's partition (partition.c:7-17).
struct ProbDev {
  const double* H;
  int64_t ldH;
  int n, m, b, T, b0;
  int nb0blk;   // row blocks of tile 0
  int nrb;      // row blocks in total
  const double* wr;   // m shifts
  double* Y;          // n x m solution (in place), ldY
  int64_t ldY;
  double* Rt[2];      // running cross-over columns, n x m each (ld n), ping-pong
  double* C;          // cosines, n x m (ld n)
  double* S;          // sines, n x m (ld n)
  double* Bm[2];      // GEMM right operand per tile-column parity, b x 2m (ld b):
                      // col 2k = W_k, 2k+1 = Z_k
  double* aux[2];     // per parity, 4 per shift: P_k, Zlast_k (prescaled), ||Z_k||, prescale
  int* Ep[2];         // scale exponent of tile t after the panel pass of tile
                      // column s, in Ep[s & 1][t*m + k]; Ep[T & 1]: initial
  int* Ef;            // final scale exponent of tile t (after its sweep), T x m
  double* Ynb[2];     // per row block max |y| after the panel pass of tile
                      // column s, in Ynb[s & 1] (nrb x m); Ynb[T & 1]: initial
  double* rowinf;     // ||H(I,:)||_oo per tile row, T
  double* diagn;      // ||H(J,J)||_oo per diagonal tile, T
  double* part;       // partial absolute row sums, nrb x n
  double* scal;       // [0] = ||H||_oo, [1] = eps3
  int* status;        // per shift: bit0 converged, bit1 perturbed pivot
  int its;            // -1: start vector eps3*1; >= 0: dlaein restart vector;
                      // -2: the right-hand sides are given in Y
  const int* its_k;   // if set: the start vector of shift k is its_k[k] (-1 or
                      // >= 0 as its), overriding its
  int force_exact;    // 1: always use exact norms in the diagonal sweep
  // FP32 panel pass only (null otherwise): power-of-two equilibration of H,
  // ccol[j] = -ilogb(max_i |H(i, j)|), rrow[i] = ilogb(max_j |H(i, j)| 2^ccol[j])
  // (n each), and per tile-column parity the exponents of the [W | Z]
  // columns scaled by 2^-ccol of their panel columns, Be[s & 1][c] (2m each)
  int* ccol;
  int* rrow;
  int* Be[2];
  // and the operands in float: H(i, j) 2^(ccol[j] - rrow[i]) over the
  // Hessenberg part (ld n), the [W | Z] columns scaled by 2^(-ccol - Be) per
  // parity (b x 2m, ld b)
  float* H32;
  float* B32[2];
};

struct K1Item {
  int prob, J, k0, pad;
};
struct GemmItem {   // panel pass: row blocks rb .. rb+nrbk-1 (<= 2), column block cb
  int prob, J, rb, cb, nrbk, pad;
};
struct RowItem {  // init: (prob, row block, shift group); K0: (prob, tile)
  int prob, rb, k0, pad;
};
struct ShiftItem {  // back-transform: (prob, shift)
  int prob, k, pad0, pad1;
};

// Row block rb -> (tile, first row, rows)
__host__ __device__ inline void rowblock_range(const ProbDev& P, int rb, int& tile, int& r0,
                                               int& nr) {
  if (rb < P.nb0blk) {
    tile = 0;
    r0 = rb * kRowBlock;
    nr = (P.b0 - r0) < kRowBlock ? (P.b0 - r0) : kRowBlock;
  } else {
    const int per = P.b / kRowBlock;
    const int q = rb - P.nb0blk;
    tile = 1 + q / per;
    r0 = P.b0 + q * kRowBlock;
    nr = kRowBlock;
  }
}

__host__ __device__ inline int tile_first_row(const ProbDev& P, int t) {
  return t == 0 ? 0 : P.b0 + (t - 1) * P.b;
}
__host__ __device__ inline int tile_rows(const ProbDev& P, int t) { return t == 0 ? P.b0 : P.b; }
__host__ __device__ inline int tile_first_rowblock(const ProbDev& P, int t) {
  return t == 0 ? 0 : P.nb0blk + (t - 1) * (P.b / kRowBlock);
}
__host__ __device__ inline int tile_num_rowblocks(const ProbDev& P, int t) {
  return t == 0 ? P.nb0blk : P.b / kRowBlock;
}

}  // namespace hsrq
