/*
 * hsrq_oracle.h -- CPU restatement of the reference's robust tiled inverse
 * iteration.  TEST INFRASTRUCTURE ONLY: the oracle is the checker for the HIP
 * library and the CPU baseline of the benchmark; nothing in the product links
 * or calls it.
 *
 * The restatement follows the reference's tiled algorithm step by step
 * (partition.c, reduce.c, solve-task.c, update-task.c, robust.c,
 * shifted_Hessenberg_solve.c, inverse_iteration.c; see the citations on each
 * function in hsrq_oracle.c) with double-valued scaling factors, as the
 * reference's default build.  Where the reference has a defect the intended
 * semantics are implemented; each such place is marked FIX in the source.
 */
#ifndef HSRQ_ORACLE_H
#define HSRQ_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    long long scalings;      /* scaling factors != 1 applied (any protect_*) */
    long long perturbed;     /* tiny pivots replaced by smin (info = 1)       */
    int restarts;            /* restart rounds (new starting vectors)         */
    int nonconverged;        /* shifts still unconverged on return            */
} oracle_stats_t;

/* Same contract as the reference's tiled_inverse_iteration
 * (inverse_iteration.h:19-22).  nthreads <= 0: OpenMP default.
 * Returns 0, or a negative value for invalid arguments.  `stats` may be NULL. */
int oracle_tiled_inverse_iteration(int n, const double *H, int ldH, int nrhs,
                                   double *wr, int tlsz, int rhs_tlsz,
                                   double *vr, int ldvr, int nthreads,
                                   oracle_stats_t *stats);

/* The same, with per-shift outputs in the final shift order (either may be
 * NULL): status[k] bit 0 converged, bit 1 a pivot of the last solve was
 * raised to smin (the reference's info = 1, robust.c:462-465);
 * log2scale[k] = log2 of the shift's smallest tile scaling factor in its last
 * solve (<= 0; the factor its vector carries into the back-transform,
 * shifted_Hessenberg_solve.c:123-124).  max_restarts: restart rounds at
 * most (< 0: n, the reference's limit; 0: the first solve alone). */
int oracle_tiled_inverse_iteration_ex(int n, const double *H, int ldH, int nrhs,
                                      double *wr, int tlsz, int rhs_tlsz,
                                      double *vr, int ldvr, int nthreads,
                                      oracle_stats_t *stats, int *status,
                                      double *log2scale, int max_restarts);

/* Building blocks, exported for the oracle's own pinning tests. */
void oracle_dlartg(double f, double g, double *c, double *s, double *r);
void oracle_givens(double f, double g, double *c, double *s);
double oracle_protect_update(double t, double x, double y);
double oracle_protect_double_update(double t1, double x1, double t2, double x2,
                                    double y);
int oracle_solve_1x1_real_system(double smin, double t, double lambda,
                                 double *x, double *scale);
double oracle_nrm2_blue(int n, const double *x);
double oracle_dlange_inf(int m, int n, const double *A, int ldA);
void oracle_partition(int n, int num_tiles, int tlsz, int *p);

/* Residuals of eigenpairs (wr[k], X(:,k)):
 *   ns[k]  = ||(H - wr[k] I) x||_2 / (||H||_oo ||x||_2 eps), eps = 2^-52
 *   ref[k] = ||H x - wr[k] x|| / ((||H||_F + |wr[k]|) ||x||)   (validation.c:161)
 * Norms are overflow-safe (scaled 2-norms). */
void oracle_residuals(int n, const double *H, int ldH, int m, const double *wr,
                      const double *X, int ldX, double *ns, double *ref,
                      int nthreads);

/* Complex shifts sigma_k = wr[k] + i wi[k]: inverse iteration
 * (H - sigma_k I) x = eps3 (1, ..., 1) per shift, then dlaein's restart
 * vectors while the growth test fails, at most max_restarts times.  vr:
 * n x 2m, column 2k the real and 2k+1 the imaginary part of x_k,
 * ||x_k||_2 = 1 and its first entry of largest modulus real positive.  status
 * (may be NULL): 1 converged (growth test of dlaein), 2 a pivot raised to
 * smin. */
int oracle_complex_inverse_iteration(int n, const double *H, int ldH, int m,
                                     const double *wr, const double *wi, double *vr,
                                     int ldvr, int max_restarts, int *status, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
