/* cpu_panel.h -- TEST INFRASTRUCTURE ONLY: the OpenMP panel-blocked CPU baseline timed by bench.py (cpu_panel.c). */
#ifndef CPU_PANEL_H
#define CPU_PANEL_H
#include <stdint.h>
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif
/* B = A[row_perm, :][:, col_perm], Ruiz, the blocked supernodal LU (relax_rows, pivot_rel, panel columns per panel),
 * then x = A^-1 b for the nrhs columns of b (n x nrhs column-major, overwritten), on nthreads threads.
 * out[6]: seconds of permutation + Ruiz, of the factorization, of the solves; factor flops c * sum_k l_k u_k;
 * supernodes; threads used. */
int cpb_factor_solve(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx, const int64_t *row_perm,
                     const int64_t *col_perm, int64_t relax_rows, double pivot_rel, int64_t panel, int nthreads, double *b,
                     int64_t nrhs, double *out, char *err, size_t errlen);
#ifdef __cplusplus
}
#endif
#endif
