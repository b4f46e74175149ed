/*
 * hsrq_oracle.c -- CPU restatement of the reference's robust tiled inverse
 * iteration.  TEST INFRASTRUCTURE ONLY (checker + CPU baseline); the product
 * library never links or calls this file.
 *
 * Algorithm (one step of inverse iteration per real shift mu):
 *   solve (H - mu I) y = eps3 * 1 by an RQ factorisation built from Givens
 *   rotations applied right to left, rows tiled, every update protected
 *   against overflow by per-(shift, tile row) scaling factors; then
 *   x = Q^T y, normalise, test convergence, restart if needed.
 *
 * Order of operations.  The reference runs tiledReduce (all cross-over
 * columns) and then robustTiledSolve as two OpenMP task graphs.  Its task
 * dependences serialise the updates of one (tile row, shift tile) in the
 * order tile column J = T-1 ... 0 (update-task inout dependence,
 * shifted_Hessenberg_solve.c:276-278), and the reduction and solve work on
 * disjoint data, so the serial schedule used here -- for J = T-1 ... 0:
 * diagonal tile J; then for I = J-1 ... 0: update of tile row I, reduction of
 * its cross-over column -- applies the same operations to the same operands.
 * Only one running cross-over column per shift is kept instead of the
 * reference's T copies (Rtildes, shifted_Hessenberg_solve.c:350-367).
 * Shift tiles are independent and run in parallel (OpenMP).
 *
 * FIX markers: places where the reference's code does not do what it
 * intends; the intended semantics are implemented:
 *   Q2/Q3  restart bookkeeping (inverse_iteration.c:80-110, :177-181)
 *   Q4     per-shift rescale factor (update-task.c:218, :223)
 *   Q5/Q19 the rotated copy of X is always formed and scaled whenever the
 *          consolidated Y scale is below the X scale (update-task.c:173-186)
 *   Q7     no H(0,-1) read for the top-left tile (reduce.c:235)
 *   Q6     update bounds include every term of the update
 *          (update-task.c:121-122, :212-213)
 *   Q9/Q21 accumulated scaling factors kept as mantissa and exponent
 *          (robust.c:86-121)
 *   Q10    convergence judged on the unscaled growth
 *          (shifted_Hessenberg_solve.c:232)
 *   Q18    last row scaled once, not twice, in the protected part-1 update
 *          (update-task.c:125-131)
 *   Q20    restart vectors get exact per-tile norm bounds
 *          (inverse_iteration.c:52-56)
 */
#include "hsrq_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Global thresholds, robust.c:14-15. */
static const double g_omega = 1.e+300;
static const double g_omega_inv = 1.e-300;

/* LAPACK 3.10 dnrm2 (Blue's algorithm) thresholds and scalings, as exact
 * powers of two (shifted_Hessenberg_solve.c:15-18). */
static const double tsml = 0x1p-511;
static const double tbig = 0x1p+486;
static const double ssml = 0x1p+537;
static const double sbig = 0x1p-538;

#define HOT __attribute__((optimize("fp-contract=fast")))

/* ------------------------------------------------------------------------ */
/* Tiling: partition.c:7-17 -- T tiles of size tlsz, the FIRST one ragged.  */
/* ------------------------------------------------------------------------ */
void oracle_partition(int n, int num_tiles, int tlsz, int *p)
{
    for (int i = 1; i < num_tiles; i++)
        p[i] = n - (num_tiles - i) * tlsz;
    p[0] = 0;
    p[num_tiles] = n;
}

/* ------------------------------------------------------------------------ */
/* Givens rotations.                                                         */
/* ------------------------------------------------------------------------ */

/* LAPACK 3.10 dlartg (Anderson, "Algorithm 978: Safe scaling in the Level 1
 * BLAS", 2017): c >= 0, r carries the sign of f, c*f + s*g = r,
 * -s*f + c*g = 0.  The reference calls dlartg_ from the linked LAPACK
 * (reduce.c:16-20). */
void oracle_dlartg(double f, double g, double *c, double *s, double *r)
{
    const double safmin = 0x1p-1022;
    const double safmax = 0x1p+1022;
    const double rtmin = sqrt(safmin);
    const double rtmax = sqrt(safmax / 2);
    const double f1 = fabs(f), g1 = fabs(g);
    if (g == 0.0) {
        *c = 1.0; *s = 0.0; *r = f;
    } else if (f == 0.0) {
        *c = 0.0; *s = copysign(1.0, g); *r = g1;
    } else if (f1 > rtmin && f1 < rtmax && g1 > rtmin && g1 < rtmax) {
        const double d = sqrt(f * f + g * g);
        *c = f1 / d;
        *r = copysign(d, f);
        *s = g / *r;
    } else {
        const double u = fmin(safmax, fmax(safmin, fmax(f1, g1)));
        const double fs = f / u, gs = g / u;
        const double d = sqrt(fs * fs + gs * gs);
        *c = fabs(fs) / d;
        const double rr = copysign(d, f);
        *s = gs / rr;
        *r = rr * u;
    }
}

/* reduce.c:14-24: dlartg with the sine negated ("transpose"). */
void oracle_givens(double f, double g, double *c, double *s)
{
    double r;
    oracle_dlartg(f, g, c, s, &r);
    *s = -(*s);
}

/* ------------------------------------------------------------------------ */
/* Robust scalar core, robust.c (default build: scaling_t = double).         */
/* ------------------------------------------------------------------------ */

/* robust.c:144-166 */
static double protect_real_division(double b, double t)
{
    double scale = 1.0;
    if (fabs(t) < g_omega_inv) {
        if (fabs(b) > fabs(t) * g_omega)
            scale = (fabs(t) * g_omega) / fabs(b);
    } else {
        if (1.0 > fabs(t) && fabs(b) > fabs(t) * g_omega)
            scale = 1.0 / fabs(b);
    }
    return scale;
}

/* robust.c:175-185 */
static double protect_sum(double x, double y)
{
    double scale = 1.0;
    if ((x > 0 && y > 0) || (x < 0 && y < 0))
        if (fabs(x) > g_omega - fabs(y))
            scale = 0.5;
    return scale;
}

/* robust.c:244-262: alpha such that (alpha*y) - t*(alpha*x) cannot overflow. */
double oracle_protect_update(double tnorm, double xnorm, double ynorm)
{
    double scale = 1.0;
    if (xnorm <= 1.0) {
        if (tnorm * xnorm > g_omega - ynorm)
            scale = 0.5;
    } else {
        if (tnorm > (g_omega - ynorm) / xnorm)
            scale = 0.5 / xnorm;
    }
    return scale;
}

/* robust.c:270-285 */
double oracle_protect_double_update(double t1, double x1, double t2, double x2,
                                    double y)
{
    return fmin(oracle_protect_update(t1, x1, y), oracle_protect_update(t2, x2, y));
}

/* robust.c:453-477: x := x / (t - lambda) robustly; *scale receives the
 * factor applied to x.  Returns 1 if |t - lambda| was raised to smin. */
int oracle_solve_1x1_real_system(double smin, double t, double lambda,
                                 double *x, double *scale)
{
    int info = 0;
    const double s = protect_sum(t, -lambda);
    double csr = (s * t) - (s * lambda);
    if (fabs(csr) < smin) {
        csr = smin;
        info = 1;
    }
    const double alpha = protect_real_division(x[0], csr);
    x[0] = (alpha * x[0]) / csr;
    scale[0] = alpha / s;
    return info;
}

/* ------------------------------------------------------------------------ */
/* Norms.                                                                    */
/* ------------------------------------------------------------------------ */
static double infnorm(int n, const double *x)
{
    double nrm = 0.0;
    for (int i = 0; i < n; i++) {
        const double a = fabs(x[i]);
        if (a > nrm) nrm = a;
    }
    return nrm;
}

/* dlange('I'): maximum absolute row sum. */
double oracle_dlange_inf(int m, int n, const double *A, int ldA)
{
    double *rows = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
    for (int j = 0; j < n; j++)
        for (int i = 0; i < m; i++)
            rows[i] += fabs(A[i + (size_t)j * ldA]);
    double nrm = 0.0;
    for (int i = 0; i < m; i++)
        if (rows[i] > nrm || isnan(rows[i])) nrm = rows[i];
    free(rows);
    return nrm;
}

/* Three-accumulator 2-norm of LAPACK 3.10 dnrm2 (the accumulation of
 * shifted_Hessenberg_solve.c:170-229). */
typedef struct { double asml, amed, abig; } acc3_t;

static inline void acc3_add(acc3_t *a, double v)
{
    const double x = fabs(v);
    if (x > tbig)
        a->abig += (x * sbig) * (x * sbig);
    else if (x < tsml)
        a->asml += (x * ssml) * (x * ssml);
    else
        a->amed += x * x;
}

static double acc3_norm(acc3_t a)
{
    double ynrm, scale = 1.0;
    if (a.abig > 0.0) {
        ynrm = a.abig + (a.amed * sbig) * sbig;
        scale = 1.0 / sbig;
    } else if (a.asml > 0.0) {
        if (a.amed > 0.0) {
            const double amed = sqrt(a.amed);
            const double asml = sqrt(a.asml) / ssml;
            const double ymin = asml > amed ? amed : asml;
            const double ymax = asml > amed ? asml : amed;
            ynrm = ymax * ymax * (1.0 + (ymin / ymax) * (ymin / ymax));
        } else {
            scale = 1.0 / ssml;
            ynrm = a.asml;
        }
    } else {
        ynrm = a.amed;
    }
    return scale * sqrt(ynrm);
}

double oracle_nrm2_blue(int n, const double *x)
{
    acc3_t a = {0, 0, 0};
    for (int i = 0; i < n; i++) acc3_add(&a, x[i]);
    return acc3_norm(a);
}

/* ------------------------------------------------------------------------ */
/* Plan shared by all shift tiles.                                           */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n, T;
    const int *fr;           /* first row of each tile, T+1 entries */
    const double *H;
    int ldH;
    const double *Hn;        /* Hn[I + J*T] = ||H(I,J)||_oo (tile block) */
    double smlnum;
} plan_t;

#define HH(i, j) P->H[(i) + (size_t)(j) * P->ldH]

static inline void scale_vec(int n, double *x, double a)
{
    if (a != 1.0)
        for (int i = 0; i < n; i++) x[i] = a * x[i];
}

/* utils.c:83-104 */
static inline void scale_except(int n, int skip, double *x, double a)
{
    if (a != 1.0)
        for (int i = 0; i < n; i++)
            if (i != skip) x[i] = a * x[i];
}

typedef struct {
    long long scalings, perturbed;
} tstats_t;

/* Accumulated scaling factor m * 2^e (0.5 <= m < 1, or m = 0).  The vector
 * arithmetic applies the reference's factors unchanged; only the running
 * product is kept as mantissa and exponent.  FIX Q9/Q21: the reference's
 * double-valued product (robust.c:86-93) underflows to 0 after a few
 * protections on graded input, and smin / scale (robust.c:96-109) then
 * yields 0/0; its integer variant (INTSCALING) is inconsistent (Q9). */
typedef struct { double m; int e; } sc_t;

static inline sc_t sc_one(void) { sc_t a = {0.5, 1}; return a; }

static inline void sc_mul(sc_t *a, double phi)
{
    if (phi == 1.0) return;
    int e2;
    const double m2 = frexp(phi, &e2);
    int e3;
    a->m = frexp(a->m * m2, &e3);
    a->e += e2 + e3;
}

static inline int sc_lt(sc_t a, sc_t b)
{
    if (a.m == 0.0 || b.m == 0.0) return a.m < b.m;
    return a.e < b.e || (a.e == b.e && a.m < b.m);
}

/* a / b as a double (a <= b gives a factor <= 1). */
static inline double sc_ratio(sc_t a, sc_t b)
{
    return ldexp(a.m / b.m, a.e - b.e);
}

/*
 * Diagonal tile J for one shift: rotation generation (reduceDiag,
 * reduce.c:27-68; factor_and_solve_R11, solve-task.c:123-225) fused with the
 * robust back substitution (solve, solve-task.c:19-119).  `v` enters as the
 * cross-over column restricted to the tile (shift already applied to its
 * last row when J = T-1).
 */
static void diag_tile(const plan_t *P, int J, double mu, double *v, double *y,
                      double *c, double *s, sc_t *beta, tstats_t *st, int *pert)
{
    const int l = P->fr[J];
    const int b = P->fr[J + 1] - l;
    const double Hnorm = P->Hn[J + (size_t)J * P->T];
    const double smin = fmax(DBL_EPSILON / 2 * fabs(mu), P->smlnum);
#define Ht(i, j) HH(l + (i), l + (j))
    for (int j = b - 1; j >= 1; j--) {
        oracle_givens(v[j], Ht(j, j - 1), &c[j], &s[j]);

        /* y(j) = y(j) / (c(j) v(j) - s(j) H(j,j-1)) */
        double phi = 1.0;
        const int info = oracle_solve_1x1_real_system(
            smin, c[j] * v[j], s[j] * Ht(j, j - 1), &y[j], &phi);
        st->perturbed += info;
        *pert |= info;
        sc_mul(beta, phi);
        if (phi != 1.0) st->scalings++;
        scale_except(b, j, y, phi);

        /* y(j-1) -= y(j) c(j) v(j-1) - y(j) s(j) (H(j-1,j-1) - mu) */
        phi = oracle_protect_double_update(fabs(v[j - 1]), fabs(y[j]),
                                           fabs(Ht(j - 1, j - 1) - mu), fabs(y[j]),
                                           fabs(y[j - 1]));
        if (phi != 1.0) st->scalings++;
        scale_vec(b, y, phi);
        sc_mul(beta, phi);
        y[j - 1] = y[j - 1] - y[j] * c[j] * v[j - 1] + y[j] * s[j] * (Ht(j - 1, j - 1) - mu);

        /* y(0:j-2) -= y(j) c(j) v(0:j-2) - y(j) s(j) H(0:j-2, j-1) */
        const double ynorm = infnorm(j - 1, y);
        const double vnorm = infnorm(j - 1, v);
        phi = oracle_protect_double_update(vnorm, fabs(y[j]), Hnorm, fabs(y[j]), ynorm);
        if (phi != 1.0) st->scalings++;
        scale_vec(b, y, phi);
        sc_mul(beta, phi);
        for (int i = 0; i < j - 1; i++)
            y[i] = y[i] - y[j] * c[j] * v[i] + y[j] * s[j] * Ht(i, j - 1);

        /* next annihilation column */
        v[j - 1] = c[j] * (Ht(j - 1, j - 1) - mu) + s[j] * v[j - 1];
        for (int i = 0; i < j - 1; i++)
            v[i] = c[j] * Ht(i, j - 1) + s[j] * v[i];
    }
    if (J > 0) {
        /* Rotation coupling rows l-1 and l (reduce.c:53-61) applied to
         * R(0,0) (solve-task.c:104-114). */
        const double hleft = HH(l, l - 1);
        if (hleft != 0.0) {
            oracle_givens(v[0], hleft, &c[0], &s[0]);
        } else {
            c[0] = 1.0;
            s[0] = 0.0;
        }
        v[0] = v[0] * c[0] - hleft * s[0];
    }
    /* FIX Q7: the top-left tile has no left neighbour; c[0], s[0] unused. */
    double phi = 1.0;
    const int info = oracle_solve_1x1_real_system(smin, v[0], 0.0, &y[0], &phi);
    st->perturbed += info;
    *pert |= info;
    sc_mul(beta, phi);
    if (phi != 1.0) st->scalings++;
    scale_vec(b - 1, &y[1], phi);
#undef Ht
}

/* Y(m x w) -= A(m x k) * B(k x w); the dgemm of update-task.c:233. */
HOT static void gemm_sub(int m, int w, int k, const double *A, int ldA,
                         const double *B, int ldB, double *Y, int ldY)
{
    for (int r = 0; r < w; r++) {
        double *y = Y + (size_t)r * ldY;
        const double *bb = B + (size_t)r * ldB;
        int j = 0;
        for (; j + 4 <= k; j += 4) {
            const double b0 = bb[j], b1 = bb[j + 1], b2 = bb[j + 2], b3 = bb[j + 3];
            const double *a0 = A + (size_t)j * ldA;
            const double *a1 = a0 + ldA, *a2 = a1 + ldA, *a3 = a2 + ldA;
            for (int i = 0; i < m; i++)
                y[i] -= a0[i] * b0 + a1[i] * b1 + a2[i] * b2 + a3[i] * b3;
        }
        for (; j < k; j++) {
            const double bj = bb[j];
            const double *a0 = A + (size_t)j * ldA;
            for (int i = 0; i < m; i++) y[i] -= a0[i] * bj;
        }
    }
}

/*
 * Update of tile row I by the solved tile J for all w shifts of a shift tile
 * (tile_update, update-task.c:16-262).  Ỹ workspace: b_J x w.
 */
static void tile_update(const plan_t *P, int I, int J, int w, const double *wr,
                        double *vr, int ldvr, double *Ynorms, sc_t *Yscales,
                        const sc_t *Xscales, const double *Rt, int ldRt,
                        const double *Cc, const double *Ss, int ldcs,
                        double *Yt, tstats_t *st)
{
    const int lI = P->fr[I], m = P->fr[I + 1] - lI;
    const int lJ = P->fr[J], k = P->fr[J + 1] - lJ;
    const int applyShift = (I == J - 1);
    const double *h = &HH(lI, lJ - 1);
    const double *Hp = &HH(lI, lJ);
    const double HnIJ = P->Hn[I + (size_t)J * P->T];
#define Y(i, r) vr[lI + (i) + (size_t)(r) * ldvr]
#define X(i, r) vr[lJ + (i) + (size_t)(r) * ldvr]

    /* Consistent scaling (update-task.c:68-88). */
    for (int r = 0; r < w; r++) {
        if (sc_lt(Xscales[r], Yscales[r])) {
            const double a = sc_ratio(Xscales[r], Yscales[r]);
            for (int i = 0; i < m; i++) Y(i, r) = a * Y(i, r);
            Ynorms[r] = a * Ynorms[r];
            Yscales[r] = Xscales[r];
        }
    }

    /* Part 1: y(I) -= (H(I, lJ-1) - mu e_last) * yhat, yhat = -s_0 x_0
     * (update-task.c:103-151). */
    const double hnrm = infnorm(m, h);
    for (int r = 0; r < w; r++) {
        const double mu = wr[r];
        const double *s = Ss + (size_t)r * ldcs;
        double yhat;
        if (sc_lt(Yscales[r], Xscales[r])) {
            const double a = sc_ratio(Yscales[r], Xscales[r]);
            yhat = -s[0] * (a * X(0, r));
        } else {
            yhat = -s[0] * X(0, r);
        }
        /* FIX Q6: the bound of the combined update h*yhat - mu*yhat*e_last is
         * the sum of the two norms (update-task.c:121-122 takes the max). */
        const double tb = applyShift ? hnrm + fabs(mu) : hnrm;
        const double alpha = oracle_protect_update(tb, fabs(yhat), Ynorms[r]);
        if (alpha != 1.0) {
            st->scalings++;
            for (int i = 0; i < m; i++)
                Y(i, r) = (alpha * Y(i, r)) - h[i] * (alpha * yhat);
            /* FIX Q18: row m-1 was scaled in the loop already. */
            if (applyShift) Y(m - 1, r) = Y(m - 1, r) + mu * (alpha * yhat);
        } else {
            for (int i = 0; i < m; i++) Y(i, r) = Y(i, r) - h[i] * yhat;
            if (applyShift) Y(m - 1, r) = Y(m - 1, r) + mu * yhat;
        }
        sc_mul(&Yscales[r], alpha);
        Ynorms[r] = infnorm(m, &Y(0, r));
    }

    /* Part 2: Ỹ = G_{k-1} ... G_1 G_0 x (consistently scaled) (update-task.c:162-198). */
    double *ynrm = (double *)malloc((size_t)w * 2 * sizeof(double));
    double *tmp = ynrm + w;
    for (int r = 0; r < w; r++) {
        const double *c = Cc + (size_t)r * ldcs;
        const double *s = Ss + (size_t)r * ldcs;
        double *yt = Yt + (size_t)r * k;
        /* FIX Q5/Q19: always copy; scale whenever Yscale < Xscale. */
        const double a = sc_lt(Yscales[r], Xscales[r]) ? sc_ratio(Yscales[r], Xscales[r]) : 1.0;
        for (int i = 0; i < k; i++) yt[i] = a * X(i, r);
        yt[0] = c[0] * yt[0];
        for (int j = 1; j < k; j++) {
            const double tau1 = yt[j - 1], tau2 = yt[j];
            yt[j - 1] = c[j] * tau1 - s[j] * tau2;
            yt[j] = s[j] * tau1 + c[j] * tau2;
        }
        ynrm[r] = infnorm(k, yt);
    }
    /* protect_multi_rhs_update_real (robust.c:345-363), applied per shift
     * (FIX Q4).  FIX Q6: the update multiplies [H(I, lJ:lJ+k-2) | R̃(I,J)],
     * so the bound includes ||R̃(I,J)||_oo (update-task.c:212-213 omits the
     * rank-1 term, which overflows on graded input). */
    for (int r = 0; r < w; r++) {
        const double rn = infnorm(m, Rt + lI + (size_t)r * ldRt);
        tmp[r] = oracle_protect_update(HnIJ + rn, ynrm[r], Ynorms[r]);
        if (tmp[r] != 1.0) {
            st->scalings++;
            scale_vec(k, Yt + (size_t)r * k, tmp[r]);
            scale_vec(m, &Y(0, r), tmp[r]);
            sc_mul(&Yscales[r], tmp[r]);
        }
    }
    free(ynrm);

    /* Y(I) -= H(I, lJ : lJ+k-2) * Ỹ(0:k-2)   (update-task.c:233) */
    gemm_sub(m, w, k - 1, Hp, P->ldH, Yt, k, &Y(0, 0), ldvr);

    /* Y(I) -= R̃(I,J) * Ỹ(k-1)   (update-task.c:239-243) */
    for (int r = 0; r < w; r++) {
        const double t = Yt[(size_t)r * k + k - 1];
        const double *rt = Rt + lI + (size_t)r * ldRt;
        for (int i = 0; i < m; i++) Y(i, r) = Y(i, r) - rt[i] * t;
        Ynorms[r] = infnorm(m, &Y(0, r));
    }
#undef Y
#undef X
}

/*
 * Cross-over column of tile row I when leaving tile column J
 * (reduceOffdiagBlocked, reduce.c:73-143): four rotations per pass,
 * r := sum_q w_q H(:, lJ-1+q) + (prod s) r, Horner form from the right.
 */
HOT static void reduce_offdiag(const plan_t *P, int I, int J, int w, const double *wr,
                               double *Rt, int ldRt, const double *Cc,
                               const double *Ss, int ldcs, double *r)
{
    const int lI = P->fr[I], nrows = P->fr[I + 1] - lI;
    const int lJ = P->fr[J], ncols = P->fr[J + 1] - lJ;
    const int applyShift = (I == J - 1);
    const double *Hb = &HH(lI, lJ - 1);
    const size_t ld = (size_t)P->ldH;
#define Hc(i, q) Hb[(i) + (size_t)(q) * ld]
    for (int rhs = 0; rhs < w; rhs++) {
        const double *c = Cc + (size_t)rhs * ldcs;
        const double *s = Ss + (size_t)rhs * ldcs;
        double *Rr = Rt + lI + (size_t)rhs * ldRt;
        memcpy(r, Rr, (size_t)nrows * sizeof(double));
        for (int q = ncols - 1; q >= 4; q -= 4) {
            const double w0 = c[q - 3];
            const double w1 = s[q - 3] * c[q - 2];
            const double w2 = s[q - 3] * s[q - 2] * c[q - 1];
            const double w3 = s[q - 3] * s[q - 2] * s[q - 1] * c[q];
            const double w4 = s[q - 3] * s[q - 2] * s[q - 1] * s[q];
            const double *h0 = &Hc(0, q - 3), *h1 = &Hc(0, q - 2);
            const double *h2 = &Hc(0, q - 1), *h3 = &Hc(0, q);
            for (int i = 0; i < nrows; i++)
                r[i] = w0 * h0[i] + w1 * h1[i] + w2 * h2[i] + w3 * h3[i] + w4 * r[i];
        }
        const double w0 = c[0];
        const double w1 = s[0] * c[1];
        const double w2 = s[0] * s[1] * c[2];
        const double w3 = s[0] * s[1] * s[2] * c[3];
        const double w4 = s[0] * s[1] * s[2] * s[3];
        for (int i = 0; i < nrows - 1; i++)
            Rr[i] = w0 * Hc(i, 0) + w1 * Hc(i, 1) + w2 * Hc(i, 2) + w3 * Hc(i, 3) + w4 * r[i];
        const int i = nrows - 1;
        Rr[i] = w0 * (Hc(i, 0) - applyShift * wr[rhs]) + w1 * Hc(i, 1) + w2 * Hc(i, 2)
              + w3 * Hc(i, 3) + w4 * r[i];
    }
#undef Hc
}

/*
 * One shifted solve for the w shifts of a shift tile:
 * solve_Hessenberg_system_real_shift (shifted_Hessenberg_solve.c:334-394)
 * including the back-transform, norm, convergence test and normalisation
 * of robustTiledSolve (shifted_Hessenberg_solve.c:104-245).
 * vrn: per (shift, tile) norm bounds of the right-hand side, vrn[r + t*w].
 * conv[r]: bit 0 converged, bit 1 a pivot was raised to smin.  lsc[r]:
 * log2 of the smallest tile scale (the factor the vector carries into the
 * back-transform).
 */
static void solve_shift_tile(const plan_t *P, int w, const double *wr, double *vr,
                             int ldvr, double *vrn, int *conv, double *lsc, tstats_t *st)
{
    const int n = P->n, T = P->T;
    const int *fr = P->fr;
    int bmax = 0;
    for (int t = 0; t < T; t++)
        if (fr[t + 1] - fr[t] > bmax) bmax = fr[t + 1] - fr[t];

    double *Rt = (double *)malloc(((size_t)3 * n * w + (size_t)bmax * w
                                   + (size_t)2 * bmax) * sizeof(double));
    double *Cc = Rt + (size_t)n * w;
    double *Ss = Cc + (size_t)n * w;
    sc_t *scales = (sc_t *)malloc((size_t)T * w * sizeof(sc_t));
    double *Yt = Ss + (size_t)n * w;
    double *v = Yt + (size_t)bmax * w;
    double *rr = v + bmax;

    for (size_t i = 0; i < (size_t)T * w; i++) scales[i] = sc_one();
    memset(Cc, 0, (size_t)2 * n * w * sizeof(double));
    int *pert = (int *)calloc((size_t)w, sizeof(int));

    /* Initial cross-over columns H(:, n-1) - mu e_{n-1} (reduce.c:239-246, :266-270). */
    for (int r = 0; r < w; r++) {
        for (int i = 0; i < n; i++) Rt[i + (size_t)r * n] = HH(i, n - 1);
        Rt[n - 1 + (size_t)r * n] -= wr[r];
    }

    for (int J = T - 1; J >= 0; J--) {
        const int l = fr[J], b = fr[J + 1] - l;
        for (int r = 0; r < w; r++) {
            memcpy(v, Rt + l + (size_t)r * n, (size_t)b * sizeof(double));
            diag_tile(P, J, wr[r], v, vr + l + (size_t)r * ldvr,
                      Cc + l + (size_t)r * n, Ss + l + (size_t)r * n,
                      &scales[r + (size_t)J * w], st, &pert[r]);
        }
        for (int I = J - 1; I >= 0; I--) {
            tile_update(P, I, J, w, wr, vr, ldvr, vrn + (size_t)I * w,
                        scales + (size_t)I * w, scales + (size_t)J * w,
                        Rt, n, Cc + l, Ss + l, n, Yt, st);
            reduce_offdiag(P, I, J, w, wr, Rt, n, Cc + l, Ss + l, n, rr);
        }
    }

    /* Back-transform x = Q^T y, 2-norm, convergence, normalisation. */
    for (int r = 0; r < w; r++) {
        sc_t smin = scales[r];
        for (int t = 1; t < T; t++)
            if (sc_lt(scales[r + (size_t)t * w], smin)) smin = scales[r + (size_t)t * w];
        double *x = vr + (size_t)r * ldvr;
        const double *c = Cc + (size_t)r * n;
        const double *s = Ss + (size_t)r * n;
        acc3_t acc = {0, 0, 0};
        double tau1 = 0.0;
        for (int t = 0; t < T; t++) {
            const double alpha = sc_ratio(smin, scales[r + (size_t)t * w]);
            int start = fr[t];
            if (t == 0) {
                start = 1;
                tau1 = alpha * x[0];
            }
            for (int i = start; i < fr[t + 1]; i++) {
                const double tau2 = alpha * x[i];
                x[i - 1] = c[i] * tau1 - s[i] * tau2;
                tau1 = c[i] * tau2 + s[i] * tau1;
                acc3_add(&acc, x[i - 1]);
            }
        }
        x[n - 1] = tau1;
        acc3_add(&acc, x[n - 1]);
        const double ynrm = acc3_norm(acc);
        /* FIX Q10: converged iff the unscaled growth ||x|| / smin exceeds
         * 0.1 / sqrt(n) (shifted_Hessenberg_solve.c:33-40). */
        if (ynrm > ldexp(0.1 / sqrt((double)n) * smin.m, smin.e) && isfinite(ynrm)) {
            for (int i = 0; i < n; i++) x[i] /= ynrm;
            conv[r] = 1;
        } else {
            conv[r] = 0;
        }
        conv[r] |= pert[r] ? 2 : 0;
        lsc[r] = smin.m > 0.0 ? smin.e + log2(smin.m) : -INFINITY;
    }
    free(pert);
    free(Rt);
    free(scales);
}

/* Solve `cnt` contiguous shifts, shift tiles of rhs_tlsz in parallel. */
static void solve_shifts(const plan_t *P, int cnt, const double *wr, double *vr,
                         int ldvr, const double *tile_bound /* per shift x tile */,
                         int rhs_tlsz, int *conv, double *lsc, oracle_stats_t *stats)
{
    if (cnt <= 0) return;
    const int nt = (cnt + rhs_tlsz - 1) / rhs_tlsz;
    int *fc = (int *)malloc((size_t)(nt + 1) * sizeof(int));
    oracle_partition(cnt, nt, rhs_tlsz, fc);
    long long sc = 0, pe = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : sc, pe)
    for (int tk = 0; tk < nt; tk++) {
        const int k0 = fc[tk], w = fc[tk + 1] - k0;
        double *vrn = (double *)malloc((size_t)w * P->T * sizeof(double));
        for (int t = 0; t < P->T; t++)
            for (int r = 0; r < w; r++)
                vrn[r + (size_t)t * w] = tile_bound[(size_t)(k0 + r) * P->T + t];
        tstats_t st = {0, 0};
        solve_shift_tile(P, w, wr + k0, vr + (size_t)k0 * ldvr, ldvr, vrn, conv + k0,
                         lsc + k0, &st);
        sc += st.scalings;
        pe += st.perturbed;
        free(vrn);
    }
    if (stats) {
        stats->scalings += sc;
        stats->perturbed += pe;
    }
    free(fc);
}

/* Starting vectors: eps3 * 1 (inverse_iteration.c:19-34) or the its-th
 * restart vector of LAPACK 3.10 dlaein (inverse_iteration.c:36-57). Per-tile
 * norm bounds are exact (FIX Q20). */
static void start_vectors(const plan_t *P, int cnt, double *vr, int ldvr,
                          double eps3, int its, double *bound)
{
    const int n = P->n;
    for (int r = 0; r < cnt; r++) {
        double *x = vr + (size_t)r * ldvr;
        if (its < 0) {
            for (int i = 0; i < n; i++) x[i] = eps3;
        } else {
            const double tmp = eps3 / (sqrt((double)n) + 1.0);
            x[0] = eps3;
            for (int i = 1; i < n; i++) x[i] = tmp;
            x[n - 1 - its] -= eps3 * sqrt((double)n);
        }
        for (int t = 0; t < P->T; t++)
            bound[(size_t)r * P->T + t] = infnorm(P->fr[t + 1] - P->fr[t], x + P->fr[t]);
    }
}

int oracle_tiled_inverse_iteration(int n, const double *H, int ldH, int nrhs,
                                   double *wr, int tlsz, int rhs_tlsz,
                                   double *vr, int ldvr, int nthreads,
                                   oracle_stats_t *stats)
{
    return oracle_tiled_inverse_iteration_ex(n, H, ldH, nrhs, wr, tlsz, rhs_tlsz, vr, ldvr,
                                             nthreads, stats, NULL, NULL, -1);
}

int oracle_tiled_inverse_iteration_ex(int n, const double *H, int ldH, int nrhs,
                                      double *wr, int tlsz, int rhs_tlsz,
                                      double *vr, int ldvr, int nthreads,
                                      oracle_stats_t *stats, int *status, double *log2scale,
                                      int max_restarts)
{
    if (n <= 0 || ldH < n || nrhs < 0 || tlsz <= 0 || tlsz % 4 != 0 || rhs_tlsz <= 0
        || ldvr < n || !H || !wr || (nrhs > 0 && !vr))
        return -1;
    if (stats) memset(stats, 0, sizeof(*stats));
    if (nrhs == 0) return 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif

    const int T = (n + tlsz - 1) / tlsz;
    int *fr = (int *)malloc((size_t)(T + 1) * sizeof(int));
    oracle_partition(n, T, tlsz, fr);

    /* Tile norms ||H(I,J)||_oo (solve-task.c:30-34 diagonal b x b tiles,
     * update-task.c:56-62 off-diagonal blocks), computed up front instead of
     * lazily. */
    double *Hn = (double *)malloc((size_t)T * T * sizeof(double));
#pragma omp parallel for schedule(dynamic, 1)
    for (int J = 0; J < T; J++)
        for (int I = 0; I <= J; I++)
            Hn[I + (size_t)J * T] = oracle_dlange_inf(fr[I + 1] - fr[I], fr[J + 1] - fr[J],
                                                       H + fr[I] + (size_t)fr[J] * ldH, ldH);

    const double Hnrm = oracle_dlange_inf(n, n, H, ldH);
    const double eps = DBL_EPSILON / 2;
    const double eps3 = Hnrm * eps;   /* inverse_iteration.c:157-158 */
    if (!(eps3 > 0.0) || !isfinite(eps3)) {
        free(fr);
        free(Hn);
        return -2;
    }
    plan_t P = {n, T, fr, H, ldH, Hn, fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps))};

    int *conv = (int *)malloc((size_t)nrhs * sizeof(int));
    double *lsc = (double *)malloc((size_t)nrhs * sizeof(double));
    double *tls = (double *)malloc((size_t)nrhs * sizeof(double));
    double *bound = (double *)malloc((size_t)nrhs * T * sizeof(double));
    double *twr = (double *)malloc((size_t)nrhs * sizeof(double));
    double *tvr = (double *)malloc((size_t)nrhs * ldvr * sizeof(double));
    int *tcv = (int *)malloc((size_t)nrhs * sizeof(int));

    start_vectors(&P, nrhs, vr, ldvr, eps3, -1, bound);
    solve_shifts(&P, nrhs, wr, vr, ldvr, bound, rhs_tlsz, conv, lsc, stats);

    /* FIX Q2/Q3: stable partition [converged | unconverged] of the active
     * segment, then re-solve only the unconverged ones with the its-th
     * restart vector, at most n times (dlaein semantics). */
    int first = 0;
    const int rounds = max_restarts < 0 || max_restarts > n ? n : max_restarts;
    for (int its = 0; its < rounds; its++) {
        int cnt = nrhs - first, nc = 0;
        for (int r = first; r < nrhs; r++) nc += conv[r] & 1;
        if (nc == cnt) break;
        int pc = 0, pn = nc;
        for (int r = first; r < nrhs; r++) {
            const int dst = (conv[r] & 1) ? pc++ : pn++;
            twr[dst] = wr[r];
            tcv[dst] = conv[r];
            tls[dst] = lsc[r];
            memcpy(tvr + (size_t)dst * ldvr, vr + (size_t)r * ldvr, (size_t)n * sizeof(double));
        }
        for (int q = 0; q < cnt; q++) {
            wr[first + q] = twr[q];
            conv[first + q] = tcv[q];
            lsc[first + q] = tls[q];
            memcpy(vr + (size_t)(first + q) * ldvr, tvr + (size_t)q * ldvr, (size_t)n * sizeof(double));
        }
        first += nc;
        start_vectors(&P, nrhs - first, vr + (size_t)first * ldvr, ldvr, eps3, its, bound);
        solve_shifts(&P, nrhs - first, wr + first, vr + (size_t)first * ldvr, ldvr, bound,
                     rhs_tlsz, conv + first, lsc + first, stats);
        if (stats) stats->restarts++;
    }
    if (stats) {
        int nc = 0;
        for (int r = 0; r < nrhs; r++) nc += !(conv[r] & 1);
        stats->nonconverged = nc;
    }
    if (status) memcpy(status, conv, (size_t)nrhs * sizeof(int));
    if (log2scale) memcpy(log2scale, lsc, (size_t)nrhs * sizeof(double));
    free(lsc);
    free(tls);

    free(conv);
    free(bound);
    free(twr);
    free(tvr);
    free(tcv);
    free(fr);
    free(Hn);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Residuals (validation.c:79-171 and the north-star form).                  */
/* ------------------------------------------------------------------------ */

/* Overflow-safe 2-norm: scale by the largest magnitude. */
static double nrm2_scaled(int n, const double *x)
{
    const double amax = infnorm(n, x);
    if (amax == 0.0 || !isfinite(amax)) return amax;
    const int e = ilogb(amax);
    double sum = 0.0;
    for (int i = 0; i < n; i++) {
        const double t = ldexp(x[i], -e);
        sum += t * t;
    }
    return ldexp(sqrt(sum), e);
}

void oracle_residuals(int n, const double *H, int ldH, int m, const double *wr,
                      const double *X, int ldX, double *ns, double *ref, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    /* ||H||_oo (the norm of eps3, inverse_iteration.c:154) for the
     * north-star residual; ||H||_F over the Hessenberg part
     * (validation.c:23-44), scaled, for the reference's form. */
    const double hinf = oracle_dlange_inf(n, n, H, ldH);
    double hmax = 0.0;
    for (int j = 0; j < n; j++) {
        const int len = j + 2 < n ? j + 2 : n;
        const double t = infnorm(len, H + (size_t)j * ldH);
        if (t > hmax) hmax = t;
    }
    double hf = 0.0;
    if (hmax > 0.0) {
        const int e = ilogb(hmax);
        double sum = 0.0;
        for (int j = 0; j < n; j++) {
            const int len = j + 2 < n ? j + 2 : n;
            for (int i = 0; i < len; i++) {
                const double t = ldexp(H[i + (size_t)j * ldH], -e);
                sum += t * t;
            }
        }
        hf = ldexp(sqrt(sum), e);
    }
#pragma omp parallel
    {
        double *r = (double *)malloc((size_t)n * sizeof(double));
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < m; k++) {
            const double *x = X + (size_t)k * ldX;
            for (int i = 0; i < n; i++) r[i] = -wr[k] * x[i];
            for (int j = 0; j < n; j++) {
                const int len = j + 2 < n ? j + 2 : n;
                const double xj = x[j];
                const double *h = H + (size_t)j * ldH;
                for (int i = 0; i < len; i++) r[i] += h[i] * xj;
            }
            const double rn = nrm2_scaled(n, r);
            const double xn = nrm2_scaled(n, x);
            ns[k] = rn / (hinf * xn * DBL_EPSILON);
            ref[k] = rn / (hf * xn + fabs(wr[k]) * xn);
        }
        free(r);
    }
}

/* ------------------------------------------------------------------------ */
/* Complex shifts (the reference's unwired complex capability: its robust    */
/* complex division robust.c:639-671 and solve_1x1_cmplx_system :677-767;    */
/* its driver rejects complex eigenvalues, main.c:95-97).  Restatement of    */
/* this library's complex-shift solve (k_csolve): one unblocked RQ sweep of  */
/* H - sigma I with complex Givens rotations, the column-oriented back       */
/* substitution fused into it, power-of-two protections, then x = U_{n-1}..  */
/* U_1 z.  Written independently of the device code from the same equations. */
/* ------------------------------------------------------------------------ */

/* sqrt(a^2 + b^2 + c^2) with power-of-two scaling (no overflow). */
static double norm3(double a, double b, double c)
{
    const double m = fmax(fabs(a), fmax(fabs(b), fabs(c)));
    if (m == 0.0 || !isfinite(m)) return m;
    const int e = ilogb(m);
    const double as = ldexp(a, -e), bs = ldexp(b, -e), cs = ldexp(c, -e);
    return ldexp(sqrt((as * as + bs * bs) + cs * cs), e);
}

static int pow2_exponent(double alpha)
{
    return alpha >= 1.0 ? 0 : ilogb(alpha);
}

/* Pivot step st (st = 0: the last pivot R(0,0) = v_0).  Unitary
 * M = [[c, -conj(s)], [s, c]] on columns (h, v) = (st-1, st) with
 * c h_st + s v_st = 0 for h_st = hsub real: c = |v|/d, s = -(hsub/d) conj(ph),
 * ph = v/|v| (1 if v = 0), d = sqrt(|v|^2 + hsub^2), R(st,st) = ph d; then
 * q = y_st / R(st,st) with the pivot raised to smin below it (robust.c:639-671
 * replaces a tiny complex pivot by the real smin) and the division protected
 * by 2^e1 (robust.c:144-166).  |v| and d share one power-of-two scaling; the
 * divisions by |v| and d are reciprocals taken once. */
typedef struct { double c, sr, si, qr, qi; int e1, info; } cpivot_t;

static cpivot_t cpivot(int st, double vr, double vi, double hsub, double yr, double yi,
                       double smin)
{
    cpivot_t o = {1.0, 0.0, 0.0, 0.0, 0.0, 0, 0};
    double pr = 1.0, pi = 0.0, av, d;
    const double h = st >= 1 ? hsub : 0.0;
    const double mm = fmax(fabs(vr), fmax(fabs(vi), fabs(h)));
    if (mm == 0.0 || !isfinite(mm)) {
        av = fmax(fabs(vr), fabs(vi));
        d = mm;
    } else {
        const int e = ilogb(mm);
        const double a = ldexp(vr, -e), b = ldexp(vi, -e), g = ldexp(h, -e);
        const double s2v = a * a + b * b;
        av = ldexp(sqrt(s2v), e);
        d = ldexp(sqrt(s2v + g * g), e);
    }
    const int tiny = d < smin;
    const double deff = tiny ? smin : d;
    const double id = 1.0 / deff;
    if (av > 0.0) {
        const double ia = 1.0 / av;
        pr = vr * ia;
        pi = vi * ia;
    }
    if (st >= 1 && d > 0.0) {
        const double idr = tiny ? 1.0 / d : id;
        o.c = av * idr;
        const double t = hsub * idr;
        o.sr = -(t * pr);
        o.si = t * pi;
    }
    const double qpr = tiny ? 1.0 : pr, qpi = tiny ? 0.0 : pi;
    double wr = yr * qpr + yi * qpi;
    double wi = yi * qpr - yr * qpi;
    o.e1 = pow2_exponent(protect_real_division(fabs(wr) + fabs(wi), deff));
    if (o.e1 != 0) {
        wr = ldexp(wr, o.e1);
        wi = ldexp(wi, o.e1);
    }
    o.qr = wr * id;
    o.qi = wi * id;
    o.info = tiny;
    return o;
}

/* One complex shift; x (2n: re, im) unit 2-norm with its largest entry real
 * positive.  Returns status bits (1 converged, 2 perturbed). */
static int csolve_one(int n, const double *H, int ldH, double hinf, double sr0, double si0,
                      int its, double *xr, double *xi, double *C, double *SR, double *SI)
{
    const double eps = 0x1p-53;
    const double smlnum = fmax(2 * DBL_MIN, DBL_MIN * ((double)n / eps));
    const double asig = norm3(sr0, si0, 0.0);
    const double smin = fmax(eps * asig, smlnum);
    const double eps3 = hinf * 0x1p-53;
    const double slack = 1.0 + 0x1p-40;
    const double hm = hinf + asig;
    /* |v_i| <= hm sqrt(j + 1) after j rotations, |r_i| <= sqrt(|v_i|^2 + |h_i|^2),
     * |Re z| + |Im z| <= sqrt(2) |z|: a bound of |Re r_i| + |Im r_i| in every step */
    const double br = M_SQRT2 * hm * sqrt((double)n + 1.0) * slack;
    double *vr = (double *)malloc(sizeof(double) * (size_t)n);
    double *vi = (double *)malloc(sizeof(double) * (size_t)n);
    double *yr = xr, *yi = xi;   /* y lives in x until the back-transform */
    /* start vector (inverse_iteration.c:19-57): eps3 (1, ..., 1), or dlaein's
     * its-th restart vector */
    const double sq = sqrt((double)n);
    for (int i = 0; i < n; i++) {
        vr[i] = H[i + (size_t)(n - 1) * ldH];
        vi[i] = 0.0;
        if (its < 0) {
            yr[i] = eps3;
        } else {
            yr[i] = (i == 0) ? eps3 : eps3 / (sq + 1.0);
            if (i == n - 1 - its) yr[i] -= eps3 * sq;
        }
        yi[i] = 0.0;
    }
    vr[n - 1] -= sr0;
    vi[n - 1] = -si0;
    int E = 0, info = 0;
    double by = its < 0 ? eps3 : eps3 * (1.0 + sq);
    for (int st = n - 1; st >= 1; st--) {
        const cpivot_t P = cpivot(st, vr[st], vi[st], H[st + (size_t)(st - 1) * ldH], yr[st],
                                  yi[st], smin);
        const double c = P.c, sr = P.sr, si = P.si;
        double qr = P.qr, qi = P.qi;
        C[st] = c; SR[st] = sr; SI[st] = si;
        info |= P.info;
        if (P.e1 != 0) {
            for (int i = 0; i < n; i++) { yr[i] = ldexp(yr[i], P.e1); yi[i] = ldexp(yi[i], P.e1); }
            by = ldexp(by, P.e1);
            E += P.e1;
        }
        const double nq = fabs(qr) + fabs(qi);
        double byn = (by + nq * br) * slack;
        const double *hc = H + (size_t)(st - 1) * ldH;
        if (!(byn <= g_omega)) {
            /* exact bound a = max_i N(y_i) + N(q) N(r_i) of the update, in the
             * scaled form a 2^-kk (no overflow), then one power-of-two
             * protection */
            const int kk = nq > 0.0 ? ilogb(nq) : 0;
            const double nqs = ldexp(nq, -kk);
            double a = 0.0;
            for (int i = 0; i < st; i++) {
                const double hr = (i == st - 1) ? hc[i] - sr0 : hc[i];
                const double hi = (i == st - 1) ? -si0 : 0.0;
                const double rr = c * vr[i] - (sr * hr + si * hi);
                const double ri = c * vi[i] - (sr * hi - si * hr);
                a = fmax(a, nqs * (fabs(rr) + fabs(ri)) + ldexp(fabs(yr[i]) + fabs(yi[i]), -kk));
            }
            int e2 = 0;
            if (a > ldexp(g_omega, -kk)) e2 = -(ilogb(a) + kk) - 1;
            if (e2 != 0) {
                for (int i = 0; i < n; i++) { yr[i] = ldexp(yr[i], e2); yi[i] = ldexp(yi[i], e2); }
                qr = ldexp(qr, e2);
                qi = ldexp(qi, e2);
                E += e2;
            }
            byn = ldexp(a, kk + e2) * slack;
        }
        /* r_i = c v_i - conj(s) h_i, v_i <- c h_i + s v_i, y_i <- y_i - q r_i
         * (fused multiply-adds) */
        for (int i = 0; i < st; i++) {
            const double hr = (i == st - 1) ? hc[i] - sr0 : hc[i];
            const double hi = (i == st - 1) ? -si0 : 0.0;
            const double rr = fma(c, vr[i], -fma(sr, hr, si * hi));
            const double ri = fma(c, vi[i], fma(si, hr, -(sr * hi)));
            const double nvr = fma(c, hr, fma(sr, vr[i], -(si * vi[i])));
            const double nvi = fma(c, hi, fma(sr, vi[i], si * vr[i]));
            yr[i] = fma(-qr, rr, fma(qi, ri, yr[i]));
            yi[i] = fma(-qr, ri, fma(-qi, rr, yi[i]));
            vr[i] = nvr;
            vi[i] = nvi;
        }
        yr[st] = qr;
        yi[st] = qi;
        by = byn;
    }
    {   /* R(0,0) = v_0 */
        const cpivot_t P = cpivot(0, vr[0], vi[0], 0.0, yr[0], yi[0], smin);
        info |= P.info;
        if (P.e1 != 0) {
            for (int i = 1; i < n; i++) { yr[i] = ldexp(yr[i], P.e1); yi[i] = ldexp(yi[i], P.e1); }
            E += P.e1;
        }
        yr[0] = P.qr;
        yi[0] = P.qi;
    }
    free(vr);
    free(vi);
    /* x = U_{n-1} ... U_1 z, U_i = [[c, -conj(s)], [s, c]] on entries (i-1, i) */
    double tr = yr[0], ti = yi[0];
    for (int i = 1; i < n; i++) {
        const double c = C[i], sr = SR[i], si = SI[i];
        const double zr = yr[i], zi = yi[i];
        xr[i - 1] = c * tr - (sr * zr + si * zi);
        xi[i - 1] = c * ti - (sr * zi - si * zr);
        const double ntr = (sr * tr - si * ti) + c * zr;
        const double nti = (sr * ti + si * tr) + c * zi;
        tr = ntr;
        ti = nti;
    }
    xr[n - 1] = tr;
    xi[n - 1] = ti;
    /* ||x||_2, the first entry of largest modulus, normalisation */
    double mx = 0.0;
    for (int i = 0; i < n; i++) mx = fmax(mx, fmax(fabs(xr[i]), fabs(xi[i])));
    double nrm = mx;
    int p = 0;
    if (mx > 0.0 && isfinite(mx)) {
        const int e = ilogb(mx);
        double sum = 0.0, best = -1.0;
        for (int i = 0; i < n; i++) {
            const double a = ldexp(xr[i], -e), b = ldexp(xi[i], -e);
            const double q = a * a + b * b;
            sum += q;
            if (q > best) { best = q; p = i; }
        }
        nrm = ldexp(sqrt(sum), e);
    }
    const int conv = isfinite(nrm) && nrm > ldexp(0.1 / sqrt((double)n), E);
    if (nrm > 0.0 && isfinite(nrm)) {
        const double ap = norm3(xr[p], xi[p], 0.0);
        const double pr = xr[p] / ap, pi = xi[p] / ap;
        for (int i = 0; i < n; i++) {
            const double a = xr[i], b = xi[i];
            xr[i] = (a * pr + b * pi) / nrm;
            xi[i] = (b * pr - a * pi) / nrm;
        }
    }
    return (conv ? 1 : 0) | (info ? 2 : 0);
}

int oracle_complex_inverse_iteration(int n, const double *H, int ldH, int m, const double *wr,
                                     const double *wi, double *vr, int ldvr, int max_restarts,
                                     int *status, int nthreads)
{
    if (n <= 0 || ldH < n || m < 0 || ldvr < n || !H || max_restarts < 0 ||
        (m > 0 && (!wr || !wi || !vr)))
        return -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const double hinf = oracle_dlange_inf(n, n, H, ldH);
#pragma omp parallel
    {
        double *w = (double *)malloc(sizeof(double) * 3 * (size_t)n);
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < m; k++) {
            /* dlaein: restart vectors until the growth test passes, at most
             * max_restarts times (and n) */
            const int lim = max_restarts < n ? max_restarts : n;
            int st = 0;
            for (int its = -1; its < lim; its++) {
                st = csolve_one(n, H, ldH, hinf, wr[k], wi[k], its, vr + (size_t)2 * k * ldvr,
                                vr + (size_t)(2 * k + 1) * ldvr, w, w + n, w + 2 * n);
                if (st & 1) break;
            }
            if (status) status[k] = st;
        }
        free(w);
    }
    return 0;
}
