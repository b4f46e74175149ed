/*
 * inverse_iteration.h -- drop-in replacement of the reference's entry point
 * (reference inverse_iteration.h:19-22), served by libhsrq on one MI355X.
 *
 * Same prototype and semantics as the reference:
 *   computes one eigenvector per real shift wr[k] of the upper Hessenberg H
 *   by one step of robust inverse iteration (Givens RQ shifted solves with
 *   overflow protection), with restarts for unconverged shifts.
 *
 * @param n        Order of H, n > 0.
 * @param H        Host pointer, column-major, ldH >= n.  Only the upper
 *                 Hessenberg part is used (and copied to the device, in
 *                 column panels); entries below the subdiagonal do not enter
 *                 the result.
 * @param nrhs     Number of shifts.
 * @param wr       Host pointer, nrhs shifts; may be reordered on return
 *                 (converged first) when restarts were needed.
 * @param tlsz     Row tile size, a positive multiple of 4.  The GPU uses the
 *                 nearest supported tile of 64, 128, 192 or 256 rows.
 * @param rhs_tlsz Shift tile size, > 0 (accepted for compatibility; shifts
 *                 are processed all at once on the GPU).
 * @param vr       Host pointer, n x nrhs (ldvr >= n): unit 2-norm
 *                 eigenvectors on return.
 *
 * Synchronous.  H2D/D2H copies are included; the device-resident API is in
 * hsrq.h.  Invalid arguments print a message to stderr and return.
 */
#ifndef INVERSE_ITERATION_H_
#define INVERSE_ITERATION_H_

#ifdef __cplusplus
extern "C" {
#endif

void tiled_inverse_iteration(
    int n, const double *H, int ldH,
    int nrhs, double *wr,
    int tlsz, int rhs_tlsz, double *vr, int ldvr);

#ifdef __cplusplus
}
#endif

#endif
