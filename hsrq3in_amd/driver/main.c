/*
 * eigvecs -- command-line driver of libhsrq with the reference's interface
 * (reference main.c:18-116):
 *
 *     eigvecs n tlsz rhs-tlsz select-frac seed
 *
 * It builds an upper Hessenberg matrix with the well separated real spectrum
 * 1, 2, ..., n (reference problemgenerator.c:305-311 chooses the same
 * eigenvalues), times one call of tiled_inverse_iteration on the first
 * n * select-frac eigenvalues as shifts, and validates every eigenvector.
 *
 * Differences from the reference driver, each deliberate:
 *  - the seed argument is honoured (the reference overwrites it with 0,
 *    main.c:45);
 *  - the test matrix is H = U B U^-1 with B lower bidiagonal (diagonal
 *    1..n) and U upper bidiagonal, built in O(n^2) from the seed.  The
 *    reference applies a random orthogonal similarity to a quasi-triangular
 *    matrix and reduces it with LAPACK dgehrd (O(n^3)); both give an
 *    unreduced Hessenberg matrix with the same prescribed spectrum;
 *  - validation prints the reference's relative error
 *    ||Hy - l y||_2 / ((||H||_F + |l|) ||y||_2) against 1e-12 (validation.c:
 *    153-168) with ||H||_F computed once, and also the residual
 *    ||Hy - l y||_2 / (||H||_oo ||y||_2 2^-52) against 10 n.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "inverse_iteration.h"

static double wall_time(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return (double)tv.tv_sec + 1e-6 * (double)tv.tv_usec;
}

static double *alloc_doubles(size_t count)
{
    void *p = NULL;
    return posix_memalign(&p, 64, sizeof(double) * (count ? count : 1)) == 0 ? (double *)p : NULL;
}

/* Leading dimension: n rounded up to a multiple of 8 doubles (64 bytes). */
static int padded(int n) { return 8 * ((n + 7) / 8); }

/* splitmix64: a small counter-based generator, uniform in [0, 1). */
static double uniform01(uint64_t *state)
{
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53;
}

/*
 * H = U B U^-1, column by column from H U = U B:
 *   B = diag(lam) + subdiag(nu),  U = diag(d) + superdiag(u).
 * U B is tridiagonal, so column j of H U involves columns j-1, j of H:
 *   H(:, j) d_j = (U B)(:, j) - u_{j-1} H(:, j-1).
 * |u_{j-1} / d_j| <= 0.9 keeps U well conditioned; nu <= 0.5 keeps the
 * eigenvectors of B (entries ~ nu^p / p!) well conditioned.
 */
static void generate_hessenberg(int n, double *H, int ld, double *lam, unsigned int seed)
{
    uint64_t st = 0x243F6A8885A308D3ull ^ ((uint64_t)seed << 17);
    double *d = malloc(sizeof(double) * (size_t)n);
    double *u = malloc(sizeof(double) * (size_t)n);
    double *nu = malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < n; j++) {
        lam[j] = j + 1.0;
        d[j] = 1.0 + uniform01(&st);
    }
    for (int j = 0; j + 1 < n; j++) {
        nu[j] = 0.25 + 0.25 * uniform01(&st);
        u[j] = (1.8 * uniform01(&st) - 0.9) * d[j + 1];
    }
    for (int j = 0; j < n; j++) {
        double *c = H + (size_t)ld * j;
        memset(c, 0, sizeof(double) * (size_t)n);
        /* (U B)(:, j): rows j-1, j, j+1 */
        c[j] = d[j] * lam[j] + (j + 1 < n ? u[j] * nu[j] : 0.0);
        if (j + 1 < n) c[j + 1] = d[j + 1] * nu[j];
        if (j >= 1) {
            const double *p = H + (size_t)ld * (j - 1);
            c[j - 1] = u[j - 1] * lam[j];
            for (int i = 0; i <= j; i++) c[i] -= u[j - 1] * p[i];
        }
        const int rows = j + 2 < n ? j + 2 : n;
        for (int i = 0; i < rows; i++) c[i] /= d[j];
    }
    free(d);
    free(u);
    free(nu);
}

/* ||H||_F over the Hessenberg part and ||H||_oo (row sums). */
static void hessenberg_norms(int n, const double *H, int ld, double *fro, double *inf)
{
    double s2 = 0.0;
    double *row = calloc((size_t)n, sizeof(double));
    for (int j = 0; j < n; j++) {
        const double *c = H + (size_t)ld * j;
        const int rows = j + 2 < n ? j + 2 : n;
        for (int i = 0; i < rows; i++) {
            s2 += c[i] * c[i];
            row[i] += fabs(c[i]);
        }
    }
    double m = 0.0;
    for (int i = 0; i < n; i++) m = row[i] > m ? row[i] : m;
    free(row);
    *fro = sqrt(s2);
    *inf = m;
}

/* Scaled 2-norm (no overflow for finite input). */
static double norm2(int n, const double *x)
{
    double s = 0.0, ssq = 1.0;
    for (int i = 0; i < n; i++) {
        const double a = fabs(x[i]);
        if (a == 0.0) continue;
        if (s < a) {
            ssq = 1.0 + ssq * (s / a) * (s / a);
            s = a;
        } else {
            ssq += (a / s) * (a / s);
        }
    }
    return s * sqrt(ssq);
}

static int validate(int n, const double *H, int ld, int m, const double *wr, const double *Y,
                    int ldY, double tol)
{
    double fro, inf;
    hessenberg_norms(n, H, ld, &fro, &inf);
    const double bound = 10.0 * n;
    double worst_ref = 0.0, worst_ns = 0.0;
    int bad_ref = 0, bad_ns = 0;
#pragma omp parallel
    {
        double *r = malloc(sizeof(double) * (size_t)n);
        double wr_loc = 0.0, wn_loc = 0.0;
        int br = 0, bn = 0;
#pragma omp for schedule(dynamic, 4)
        for (int k = 0; k < m; k++) {
            const double *y = Y + (size_t)ldY * k;
            for (int i = 0; i < n; i++) r[i] = -wr[k] * y[i];
            for (int j = 0; j < n; j++) {
                const double *c = H + (size_t)ld * j;
                const double yj = y[j];
                const int rows = j + 2 < n ? j + 2 : n;
                for (int i = 0; i < rows; i++) r[i] += c[i] * yj;
            }
            const double ny = norm2(n, y), nr = norm2(n, r);
            const double err = nr / (fro * ny + fabs(wr[k]) * ny);
            const double ns = nr / (inf * ny * 0x1.0p-52);
            if (err > tol || err != err) {
#pragma omp critical
                printf("Eigenvector stored in column %d has relative error %.6e.\n", k, err);
                br++;
            }
            if (!(ns <= bound)) bn++;
            if (err > wr_loc || err != err) wr_loc = err;
            if (ns > wn_loc || ns != ns) wn_loc = ns;
        }
#pragma omp critical
        {
            bad_ref += br;
            bad_ns += bn;
            if (wr_loc > worst_ref || wr_loc != wr_loc) worst_ref = wr_loc;
            if (wn_loc > worst_ns || wn_loc != wn_loc) worst_ns = wn_loc;
        }
        free(r);
    }
    printf("Max relative error ||Hy - ly|| / ((||H||_F + |l|) ||y||) = %.3e (tol %.0e, %d above)\n",
           worst_ref, tol, bad_ref);
    printf("Max residual ||Hy - ly|| / (||H||_oo ||y|| eps) = %.3e (bound 10n = %.0f, %d above)\n",
           worst_ns, bound, bad_ns);
    return bad_ns;
}

int main(int argc, char **argv)
{
    if (argc != 6) {
        printf("Usage %s n tlsz rhs-tlsz select-frac seed\n", argv[0]);
        printf("n:             Dimension of square matrix H\n");
        printf("tlsz:          Tile size\n");
        printf("rhs-tlsz:      Tile size for column partitioning of the right-hand sides\n");
        printf("select-frac:   Fraction of the selected eigenvalues\n");
        printf("seed:          Seed for the random number generator\n");
        return EXIT_FAILURE;
    }
    const int n = atoi(argv[1]);
    int tlsz = atoi(argv[2]);
    const int rhs_tlsz = atoi(argv[3]);
    const double select_frac = atof(argv[4]);
    const unsigned int seed = (unsigned int)strtoul(argv[5], NULL, 10);
    if (n <= 0 || tlsz <= 0 || rhs_tlsz <= 0 || !(select_frac >= 0.0 && select_frac <= 1.0)) {
        printf("Invalid arguments: need n > 0, tlsz > 0, rhs-tlsz > 0, 0 <= select-frac <= 1\n");
        return EXIT_FAILURE;
    }
    if (tlsz % 4 != 0) {
        tlsz = 4 * ((tlsz + 3) / 4);
        printf("INFO: tlsz must be divisible by 4. Set tlsz to %d.\n", tlsz);
    }

    printf("Configuration:\n");
    printf("  OpenMP threads = %d (validation)\n", omp_get_max_threads());
    printf("  n = %d\n", n);
    printf("  tlsz = %d\n", tlsz);
    printf("  rhs-tlsz = %d\n", rhs_tlsz);
    printf("  select-frac = %.2f\n", select_frac);
    printf("  Seed = %u\n", seed);

    const int nrhs = (int)(n * select_frac);
    const int ldH = padded(n), ldY = padded(n);
    const double tol = 1.e-12;
    double *H = alloc_doubles((size_t)ldH * n);
    double *wr = alloc_doubles((size_t)padded(n));
    double *Y = alloc_doubles((size_t)ldY * (nrhs > 0 ? nrhs : 1));
    if (!H || !wr || !Y) {
        printf("Out of host memory\n");
        return EXIT_FAILURE;
    }
    memset(H, 0, sizeof(double) * (size_t)ldH * n);
    memset(Y, 0, sizeof(double) * (size_t)ldY * (nrhs > 0 ? nrhs : 1));

    printf("Setting up the problem...\n");
    generate_hessenberg(n, H, ldH, wr, seed);

    printf("Solve...\n");
    const double t0 = wall_time();
    tiled_inverse_iteration(n, H, ldH, nrhs, wr, tlsz, rhs_tlsz, Y, ldY);
    const double t1 = wall_time();
    printf("Execution time [s]= %.2f\n", t1 - t0);

    printf("Validate...\n");
    const int bad = validate(n, H, ldH, nrhs, wr, Y, ldY, tol);

    free(H);
    free(Y);
    free(wr);
    return bad ? EXIT_FAILURE : EXIT_SUCCESS;
}
