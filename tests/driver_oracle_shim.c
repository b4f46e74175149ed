/* Test-only: binds the CLI driver's solver call to the CPU oracle so that the
 * driver's generator and validation can be checked without a GPU. */
#include "../oracle/hsrq_oracle.h"

void tiled_inverse_iteration(int n, const double *H, int ldH, int nrhs, double *wr, int tlsz,
                             int rhs_tlsz, double *vr, int ldvr)
{
    (void)oracle_tiled_inverse_iteration(n, H, ldH, nrhs, wr, tlsz, rhs_tlsz, vr, ldvr, 0, 0);
}
