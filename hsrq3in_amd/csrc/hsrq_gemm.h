// hsrq_gemm.h -- descriptors, item tables and launchers of the FP64 MFMA
// products around the solve (hsrq_gemm.hip): the back-transformation V = Z X
// and the residual validator R = H X - X diag(lam), many problems per launch.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace hsrq {

constexpr int kGemmPanelRows = 128;   // rows of a product block
constexpr int kGemmColsBT = 64;       // columns per block: back-transformation
constexpr int kGemmColsRes = 64;      //   and validator

// One product.  Mode 0: C = A B (A n x n, B n x m).  Mode 1 (validator):
// A = H (upper Hessenberg, hess = 1), B = X, lam = shifts; part (3 x panels
// x m Blue accumulators), rsmax (max row 1-norm per 128-row panel, from the
// panel's first column block) and res (m
// residuals ||(H - lam_k I) x_k||_2 / (||H||_oo ||x_k||_2 2^-52)) in the
// workspace of resid_describe.
struct GemmDesc {
  const double* A;
  int64_t ldA;
  const double* B;
  int64_t ldB;
  double* C;
  int64_t ldC;
  const double* lam;
  double* part;
  double* rsmax;
  double* res;
  int n, m, hess, pad;
};

struct GemmTileItem {   // product block (prob, row panel, column block);
  int prob, panel, cb, pad;   // final pass: cb = column
};

// Items of one product (XCD-aware order; bn: kGemmColsBT or kGemmColsRes) and
// of the per-column final pass of problem `prob`.
void add_gemm_items(int prob, int n, int m, int bn, std::vector<GemmTileItem>& out);
void add_resid_final_items(int prob, int m, std::vector<GemmTileItem>& out);

// Validator workspace of one problem (doubles) and its descriptor.
size_t resid_workspace_doubles(int n, int m);
void resid_describe(GemmDesc& D, int n, const double* H, int64_t ldH, int m, const double* wr,
                    const double* X, int64_t ldX, double* ws);

hipError_t launch_gemm(const GemmDesc* d_descs, const GemmTileItem* items, int count,
                       hipStream_t st);
// The product with its reductions (residual sums and row 1-norms), then the
// final pass: two launches.
hipError_t launch_resid(const GemmDesc* d_descs, const GemmTileItem* tiles, int n_tiles,
                        const GemmTileItem* fin, int n_fin, hipStream_t st);

}  // namespace hsrq
