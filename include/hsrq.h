/*
 * hsrq.h -- extended C API of libhsrq: robust tiled inverse iteration for
 * upper Hessenberg matrices on one MI355X (gfx950).
 *
 * The reference exposes a single host entry point, tiled_inverse_iteration
 * (reference inverse_iteration.h:19-22, redeclared in inverse_iteration.h
 * here).  This header adds what a GPU caller needs around it: device
 * resident problems, batches of independent problems in one schedule, a
 * continuous-batching request server, per-shift status, and an optional
 * FP32-MFMA update path.  All pointers named d* are device pointers on the
 * context's device; matrices are column-major doubles.
 *
 * Error convention: functions return 0 (HSRQ_OK) or a negative code;
 * hsrq_strerror() describes it.  No function throws or aborts.
 */
#ifndef HSRQ_H_
#define HSRQ_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HSRQ_OK = 0,
    HSRQ_EINVAL = -1,      /* invalid argument                            */
    HSRQ_EHIP = -2,        /* HIP runtime error                           */
    HSRQ_ENOMEM = -3,      /* device allocation failed                    */
    HSRQ_EUNSUPPORTED = -4,/* size outside the supported range            */
    HSRQ_ESELFTEST = -5,   /* device self-test failed                     */
    HSRQ_EBUSY = -6,       /* server queue full / capacity exceeded       */
    HSRQ_ENOTFOUND = -7    /* unknown request id                          */
};

/* Per-shift status bits (hsrq_*_status arrays). */
#define HSRQ_CONVERGED 1   /* ||x|| growth test passed, x normalised       */
#define HSRQ_PERTURBED 2   /* a pivot was raised to smin (reference info=1)*/

/* Update precision of the panel pass. */
#define HSRQ_PREC_FP64 0
#define HSRQ_PREC_FP32 1   /* FP32 MFMA panel GEMM + one FP64 correction   */

typedef struct hsrq_context hsrq_context;

/* Context = device + stream + workspace cache.  device < 0: current device. */
int hsrq_context_create(hsrq_context **ctx, int device);
void hsrq_context_destroy(hsrq_context *ctx);
/* Tile size used when a call passes tlsz <= 0 (autotuner default: 0). */
int hsrq_context_set_tile(hsrq_context *ctx, int tlsz);
/* Force exact norms in the diagonal sweep (testing aid; default 0). */
int hsrq_context_set_exact_norms(hsrq_context *ctx, int on);
/* Shifts per wavefront in the diagonal sweep: 1 (64 lanes per shift,
 * shortest step latency), 4 (16 lanes per shift, ~4x fewer instructions per
 * shift step: throughput) or 0 (automatic, default). */
int hsrq_context_set_packing(hsrq_context *ctx, int shifts_per_wave);
/* 1: run the panel pass of the rows above the next diagonal tile on a
 * second stream, overlapping the next diagonal sweep (default 0). */
int hsrq_context_set_lookahead(hsrq_context *ctx, int on);
/* Panel-pass precision, HSRQ_PREC_FP64 (default) or HSRQ_PREC_FP32. */
int hsrq_context_set_precision(hsrq_context *ctx, int prec);
const char *hsrq_strerror(int code);
/* Checks the FP64 MFMA operand layout the kernels assume. */
int hsrq_selftest(hsrq_context *ctx);

/* One problem: H is n x n (ldH >= n), wr has m shifts, vr is n x m
 * (ldvr >= n).  dwr may be reordered (converged first) after restarts,
 * exactly as the reference's wr.  status (host, may be NULL) receives m
 * HSRQ_* bit sets in the final order.  Synchronous. */
int hsrq_solve_device(hsrq_context *ctx, int n, const double *dH, int ldH, int m,
                      double *dwr, double *dvr, int ldvr, int tlsz, int *status);

typedef struct {
    int n;
    const double *H;   /* device */
    int ldH;
    int m;
    double *wr;        /* device, m */
    double *vr;        /* device, ldvr x m */
    int ldvr;
} hsrq_problem;

/* Independent problems in one device schedule (one launch sequence, graph
 * captured when reused through a plan).  status: sum(m) ints or NULL. */
int hsrq_solve_batch_device(hsrq_context *ctx, int count, const hsrq_problem *probs,
                            int tlsz, int *status);

/* Reusable plan: build once, execute many times on the same buffers. */
typedef struct hsrq_plan hsrq_plan;
int hsrq_plan_create(hsrq_context *ctx, int count, const hsrq_problem *probs, int tlsz,
                     hsrq_plan **plan);
/* Asynchronous on the context stream; no restarts (status bits tell). */
int hsrq_plan_execute(hsrq_plan *plan);
/* Launch the plan as a captured HIP graph (captured on first use). */
int hsrq_plan_execute_graph(hsrq_plan *plan);
/* Synchronous execution with HIP events around every launch:
 * kernel_ms[6] / launches[6] receive the summed device time and launch count
 * per kernel kind: 0 norms, 1 init, 2 diagonal sweep, 3 panel update,
 * 4 back-transform, 5 update scaling exponents.  Kernels of the two streams
 * overlap, so the sum can exceed the wall time. */
int hsrq_plan_execute_profiled(hsrq_plan *plan, double *kernel_ms, int *launches);
int hsrq_plan_sync(hsrq_plan *plan);
/* Copies the per-shift status bits (sum(m) ints) to host. */
int hsrq_plan_status(hsrq_plan *plan, int *status);
/* Bytes of device workspace held by the plan. */
size_t hsrq_plan_workspace_bytes(const hsrq_plan *plan);
void hsrq_plan_destroy(hsrq_plan *plan);
/* Stream of the context (hipStream_t), for callers that enqueue around it. */
void *hsrq_context_stream(hsrq_context *ctx);

/* Residual ||(H - wr_k I) x_k||_2 / (||H||_oo ||x_k||_2 eps), eps = 2^-52,
 * per eigenvector, computed on the device (validator).  res: host, m. */
int hsrq_residuals_device(hsrq_context *ctx, int n, const double *dH, int ldH, int m,
                          const double *dwr, const double *dvr, int ldvr, double *res);

#ifdef __cplusplus
}
#endif

#endif
