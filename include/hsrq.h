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
 *
 * Input ordering: the context's stream (hsrq_context_stream) and the
 * server's streams are non-blocking; they do not wait for work on other
 * streams.  Device inputs must be complete when a call is made: a caller
 * whose inputs come from another stream synchronizes it first, or makes the
 * context stream wait for it (hipStreamWaitEvent).  Calls documented as
 * synchronous return with their outputs written; hsrq_plan_execute* is
 * asynchronous on the context stream (hsrq_plan_sync).
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
    HSRQ_ENOTFOUND = -7,   /* unknown request id                          */
    HSRQ_ERANGE = -8,      /* a result would leave the FP64 range         */
    HSRQ_ESCRATCH = -9     /* problem too large for the allocation-free
                              entry point's scratch (see below)            */
};

/* Per-shift status bits (hsrq_*_status arrays). */
#define HSRQ_CONVERGED 1   /* ||x|| growth test passed, x normalised       */
#define HSRQ_PERTURBED 2   /* a pivot was raised to smin (reference info=1)*/
#define HSRQ_VALIDATED 4   /* request server: residual <= 10 n eps (checked
                              online by the device validator)               */
/* Bits 8-23 of a real-shift status word: the binary orders of downscaling
 * the shift's vector carried into the back-transform (minus the smallest
 * power-of-two tile exponent of its last solve, capped at 65535; 0 when no
 * protection triggered).  The reference's counterpart is the smallest tile
 * scaling factor (shifted_Hessenberg_solve.c:123-124). */
#define HSRQ_SCALE_SHIFT 8
#define HSRQ_STATUS_SCALE(st) (((st) >> HSRQ_SCALE_SHIFT) & 0xFFFF)

/* Update precision of the panel pass. */
#define HSRQ_PREC_FP64 0
#define HSRQ_PREC_FP32 1     /* FP32 MFMA panel GEMM (power-of-two
                                equilibrated), FP64-residual check, FP64
                                re-solve of the vectors that fail it */
#define HSRQ_PREC_FP32_RAW 2 /* FP32 MFMA panel GEMM, no correction        */

typedef struct hsrq_context hsrq_context;

/* Context = device + stream + workspace cache (and, made on first use, the
 * descriptor buffer of the back-transformation and the validator).  device < 0: current device.
 * One host thread at a time per context; use one context per thread.  The
 * request server (below) is the thread-safe entry for concurrent callers. */
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
/* Concurrent lanes of a plan (0: automatic, default; 1..8).  A lane is a
 * launch sequence on its own stream; with few problems their shifts are cut
 * into one range per lane, with many problems whole problems are dealt to
 * lanes.  Results do not depend on the number of lanes. */
int hsrq_context_set_lanes(hsrq_context *ctx, int lanes);
/* k in 1..8: run the panel pass of the rows above the next diagonal tile on
 * a second stream, at most k workgroups per CU, overlapping the next tile's
 * panel pass and diagonal sweep; 0: off; -1: automatic (default: on with
 * one shift per wavefront in the sweep, with the most workgroups per CU
 * that leave room for a sweep workgroup; off with four, where the shifts
 * fill the GPU by themselves).  Applies to plans of one lane; plans of
 * several lanes ignore it.  Results do not depend on it. */
int hsrq_context_set_lookahead(hsrq_context *ctx, int blocks_per_cu);
/* Panel-pass precision: HSRQ_PREC_FP64 (default), HSRQ_PREC_FP32 or
 * HSRQ_PREC_FP32_RAW.  In the FP32 modes the panel GEMM runs in FP32 on MFMA
 * with FP32 accumulation, each 64-row block of the panel and the [W | Z]
 * columns equilibrated by powers of two first (column, then row exponents),
 * so graded matrices keep their small terms; the diagonal sweeps (with the
 * next tile's own-rows update), bounds, scale exponents, back-transform and
 * residuals stay FP64.  HSRQ_PREC_FP32 (hsrq_solve_device,
 * hsrq_solve_batch_device): after the FP32 pass the FP64 residual of every
 * vector is computed, and the shifts whose vector did not converge or whose
 * residual exceeds 10 n eps are solved again on the FP64 path.
 * HSRQ_PREC_FP32_RAW and plans: the FP32 pass alone. */
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

/* ---- allocation-free solve ----
 * No counterpart in the reference (its tiled_inverse_iteration allocates its
 * workspace per call, inverse_iteration.c:135-146).  For callers that cannot
 * allocate device memory: tight serving loops and harnesses that forbid
 * hipMalloc.  The same solve as hsrq_solve_device (same tile, same kernels,
 * same restarts; bit-identical results and status), with the workspace, the
 * item tables and the descriptor in a device array of
 * HSRQ_NOALLOC_SCRATCH_BYTES that the library holds on every GPU it runs on
 * until the process ends (set up by hsrq_context_create).  The call makes no
 * allocation; it runs the serial schedule on the context stream.  Sizes:
 * precision HSRQ_PREC_FP64 only (else HSRQ_EUNSUPPORTED), at most
 * HSRQ_NOALLOC_MAX_SHIFTS shifts, and about 32 n m + 8 n^2 / 64 bytes of
 * workspace (n = 4000, m = 400 fits; n = 8000, m = 800 does not):
 * HSRQ_ESCRATCH otherwise.  Calls share the GPU's array and are serialized by
 * a process-wide lock (overlapping calls on one GPU would race on it).
 * Synchronous. */
#define HSRQ_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define HSRQ_NOALLOC_MAX_SHIFTS 65536
int hsrq_solve_device_noalloc(hsrq_context *ctx, int n, const double *dH, int ldH, int m,
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
 * per kernel kind: 0 norms, 1 init, 2 diagonal sweep (with the panel pass of
 * its own tile's rows), 3 panel pass of the rows above, 4 back-transform,
 * 5 reserved (0).  Kernels of the two streams overlap, so the sum can exceed
 * the wall time. */
int hsrq_plan_execute_profiled(hsrq_plan *plan, double *kernel_ms, int *launches);
int hsrq_plan_sync(hsrq_plan *plan);
/* Copies the per-shift status bits (sum(m) ints) to host. */
int hsrq_plan_status(hsrq_plan *plan, int *status);
/* Lanes of the plan: problems sharing one tile size and diagonal-sweep
 * layout, each lane's launch sequence on its own stream (the automatic
 * policy picks the tile per size class and splits large lanes in two).
 * Fills up to max_lanes entries (arrays may be NULL); returns the number of
 * lanes or a negative code. */
int hsrq_plan_lanes(const hsrq_plan *plan, int max_lanes, int *tile, int *packing,
                    int *problems);
/* Autotuner: times the given problems as one plan (HIP-graph replay) for
 * every tile size in {64, 128, 256}, diagonal-sweep layout in {1, 4} shifts
 * per wavefront and lane count in {1, 2, 3}, and sets the fastest as the
 * context's tile, packing and lanes (used by later plans and solves with
 * tlsz <= 0).  The problems' vr are overwritten (no restarts run, wr keeps
 * its order).  tile, packing, lanes, ms (host, may be NULL) receive the
 * choice and its time. */
int hsrq_autotune(hsrq_context *ctx, int count, const hsrq_problem *probs, int *tile,
                  int *packing, int *lanes, double *ms);
/* Bytes of device workspace held by the plan. */
size_t hsrq_plan_workspace_bytes(const hsrq_plan *plan);
void hsrq_plan_destroy(hsrq_plan *plan);
/* Stream of the context (hipStream_t), for callers that enqueue around it. */
void *hsrq_context_stream(hsrq_context *ctx);

/* ------------------------------------------------------------------------
 * Request server: continuous batching.
 *
 * Requests of any n and m are submitted at any time.  A worker thread runs
 * the device schedule in ticks; every tick advances each admitted request
 * by one tile column of its sweep, requests of different sizes and progress
 * sharing each launch.  A request is admitted (first come, first served) as
 * soon as its HBM footprint fits the server's arena: workspace ~32 n m
 * bytes, plus H, wr and X when they arrive in host memory.  H, X and the
 * scale exponents stay in HBM until the result is fetched.  A request's
 * results are bit-identical to those of hsrq_solve_device on the same input
 * (same tile), alone or in any mix.  Every result is validated on the device
 * in the tick that completes it (hsrq_residuals_device's validator): status
 * bit HSRQ_VALIDATED marks residuals <= 10 n eps.
 * ------------------------------------------------------------------------ */
#define HSRQ_ON_HOST 0     /* H, wr, vr are host pointers (copied in/out)  */
#define HSRQ_ON_DEVICE 1   /* H, wr, vr are device pointers, used in place */

typedef struct {
    int n;
    const double *H;   /* n x n, column-major, ldH >= n                   */
    int ldH;
    int m;
    const double *wr;  /* m shifts (device requests: reordered in place)  */
    double *vr;        /* n x m result, ldvr >= n                          */
    int ldvr;
    int location;      /* HSRQ_ON_HOST or HSRQ_ON_DEVICE                   */
} hsrq_request;

typedef struct {
    int64_t submitted, completed, eigvecs, active, queued, ticks;
    double p50_ms, p99_ms;   /* submit-to-completion latency percentiles   */
    double span_s;           /* first submission to last completion        */
    size_t hbm_bytes, hbm_in_use, hbm_peak;
    int64_t validated;       /* eigenvectors with residual <= 10 n eps     */
    double worst_residual;   /* max over converged vectors of residual/(10 n) */
} hsrq_server_stats;

typedef struct hsrq_server hsrq_server;

/* hbm_bytes: arena for all admitted requests; max_active: slots (admitted
 * and not yet fetched); tlsz: tile (0: 128). */
int hsrq_server_create(hsrq_context *ctx, size_t hbm_bytes, int max_active, int tlsz,
                       hsrq_server **srv);
/* Queues a request.  Host buffers are read at admission and vr is written
 * by hsrq_server_wait: keep them valid until it returns.  HSRQ_EBUSY: the
 * request can never fit hbm_bytes. */
int hsrq_server_submit(hsrq_server *srv, const hsrq_request *req, int64_t *id);
/* 1 done, 0 pending, < 0 error. */
int hsrq_server_poll(hsrq_server *srv, int64_t id);
/* Blocks until request id completes; copies vr to host for host requests;
 * wr_out / status (host, m entries, may be NULL) receive the final shift
 * order and per-shift status bits.  Releases the request's HBM. */
int hsrq_server_wait(hsrq_server *srv, int64_t id, double *wr_out, int *status,
                     double *latency_ms);
int hsrq_server_get_stats(hsrq_server *srv, hsrq_server_stats *stats);
/* Completes admitted requests, drops queued ones, frees everything. */
void hsrq_server_destroy(hsrq_server *srv);

/* Residual ||(H - wr_k I) x_k||_2 / (||H||_oo ||x_k||_2 eps), eps = 2^-52,
 * per eigenvector, computed on the device (validator; the reference's
 * validate_real_eigenvectors, validation.c:79-171, at scale): the Hessenberg
 * product H X on FP64 MFMA, reduced per 128-row panel to overflow-safe sums
 * of squares (LAPACK 3.10 dnrm2's three accumulators).  res: host, m.
 * Synchronous. */
int hsrq_residuals_device(hsrq_context *ctx, int n, const double *dH, int ldH, int m,
                          const double *dwr, const double *dvr, int ldvr, double *res);
/* The same for any number of problems in three launches; res (host) receives
 * sum(m) residuals in problem order. */
int hsrq_residuals_batch_device(hsrq_context *ctx, int count, const hsrq_problem *probs,
                                double *res);

/* ---- complex shifts: eigenvectors of complex-conjugate eigenpairs ----
 * sigma_k = wr[k] + i wi[k] (device arrays, m entries).  One inverse-iteration
 * solve (H - sigma_k I) x = eps3 (1, ..., 1) per shift with complex Givens
 * rotations and power-of-two protections.  Shifts that fail the growth test
 * are solved again with dlaein's restart vectors, at most max_restarts times
 * (0: none; n: dlaein's limit); shifts keep their positions.  dvr: n x 2m device
 * column-major, column 2k the real and 2k+1 the imaginary part of x_k
 * (LAPACK's layout for a complex pair), ||x_k||_2 = 1, its first entry of
 * largest modulus real and positive.  status (host, m, may be NULL):
 * HSRQ_CONVERGED (growth test of dlaein), HSRQ_PERTURBED.  Up to n = 8192
 * the vector stays on chip; above, it lives in HBM (slower per row).  Returns
 * after the stream has drained.  The reference has the robust complex
 * primitives (robust.c:639-767) but no complex-shift solve; its driver
 * rejects complex eigenvalues (main.c:95-97). */
int hsrq_solve_complex_device(hsrq_context *ctx, int n, const double *dH, int ldH, int m,
                              const double *dwr, const double *dwi, double *dvr, int ldvr,
                              int max_restarts, int *status);

/* ---- test-problem generation on the device (the step before the solve) ----
 * The reference builds its problem on the host (problemgenerator.c:305-356:
 * random quasi-triangular T, Householder similarity, dgehrd; O(n^3)).  These
 * build the workloads of this library's tests and benchmarks in HBM instead.
 * n <= 65535; dH is column-major with leading dimension ldH >= n; the whole
 * n x n matrix is written (zeros below the subdiagonal).  Enqueued on the
 * context stream; the last two return after the stream has drained. */

/* H(i,j) ~ N(0,1) for i <= j+1: Philox4x32-10 with key = seed and counter
 * (j n + i) / 2, Box-Muller (cosine for even j n + i, sine for odd). */
int hsrq_gaussian_hessenberg_device(hsrq_context *ctx, int n, uint64_t seed, double *dH,
                                    int ldH);
/* H = U B U^-1 with B = diag(lam) + subdiag(nu), U = diag(d) + superdiag(u):
 * an unreduced upper Hessenberg matrix with eigenvalues lam.  lam, d: n
 * entries; nu, u: n-1 (host).  d_i != 0. */
int hsrq_constructed_hessenberg_device(hsrq_context *ctx, int n, const double *lam,
                                       const double *nu, const double *d, const double *u,
                                       double *dH, int ldH);
/* Grading H <- D H D^-1, D = diag(2^k): H(i,j) *= 2^(k_i - k_j) for i <= j+1.
 * k: host, n exponents, |k_i| <= 2^20.  A range check runs first: over the
 * nonzero entries, emax / emin (may be NULL) receive the largest and smallest
 * binary exponent E after scaling (|h| = f 2^E, 1/2 <= f < 1; INT_MIN /
 * INT_MAX if H is zero).  HSRQ_ERANGE, with dH untouched, if an entry would
 * overflow (E > 1024) or, with exact != 0, round (a result below the normal
 * range that drops bits: the similarity would no longer be exact). */
int hsrq_grade_device(hsrq_context *ctx, int n, double *dH, int ldH, const int *k, int exact,
                      int *emax, int *emin);

/* ---- back-transformation (the step after the solve) ----
 * V = Z X: with A = Z H Z^T (Z orthogonal, e.g. the accumulated reflectors of
 * a Hessenberg reduction) the columns of V are the eigenvectors of A for the
 * eigenvectors X of H (unit 2-norm stays unit up to rounding).  Device
 * column-major Z (n x n), X and V (n x m); V must not overlap Z or X.  One
 * FP64 MFMA GEMM of this library enqueued on the context stream
 * (asynchronous).  The reference stops at the eigenvectors of H
 * (inverse_iteration.h:19-22). */
int hsrq_backtransform_device(hsrq_context *ctx, int n, int m, const double *dZ, int ldZ,
                              const double *dX, int ldX, double *dV, int ldV);

#ifdef __cplusplus
}
#endif

#endif
