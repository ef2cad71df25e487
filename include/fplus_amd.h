/* fplus_amd.h -- entry points of the MI355X build beyond the reference interface (include/fplus.h).
 *
 * None of these has a counterpart in the reference; each serves a caller of this build named in its comment.
 */
#ifndef FPLUS_AMD_H
#define FPLUS_AMD_H

#include "fplus.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A symbolic analysis whose orders are given (callers that hold their own orders, e.g. the identity orders of the
 * Python SparseLU with colamd=False): A[row_perm, :][:, col_perm] is what fplus_factor factors. */
int fplus_amd_symbolic_from_orders(int64_t n, const int64_t *Ap, const int64_t *Ai, const int64_t *row_perm,
                                   const int64_t *col_perm, fplus_symbolic **sym, char *err, size_t errlen);

/* Statistics of a factorization: nnz(L), nnz(U), supernodes, device blocks, panels, dependency levels, factor flops
 * c * sum_k l_k u_k (c = 2 real, 8 complex), update GEMM flops, and the seconds spent in Ruiz, in the factorization,
 * on the host side of it, and waiting for the device.  For benchmarks and logs. */
typedef struct {
    int64_t nnzL, nnzU, nsupernodes, nblocks, panels, levels;
    double flops, gemm_flops;
    double t_ruiz, t_factor, t_host, t_wait, t_total;
    int ruiz_iters;
} fplus_amd_stats;
int fplus_amd_numeric_stats(const fplus_numeric *num, fplus_amd_stats *st);

#ifdef __cplusplus
}
#endif
#endif /* FPLUS_AMD_H */
