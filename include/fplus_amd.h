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
 * on the host side of it, and waiting for the device, the Ruiz iterations, and the block updates' algorithmic bytes
 * (each update reads its block's rows in inverse form and its segment once, writes the solved segment once, and reads
 * and writes its target rows once).  For benchmarks and logs. */
typedef struct {
    int64_t nnzL, nnzU, nsupernodes, nblocks, panels, levels;
    double flops, gemm_flops;
    double t_ruiz, t_factor, t_host, t_wait, t_total;
    int ruiz_iters;
    double gemm_bytes;
    int64_t shared_levels;   /* dependency levels whose update tasks can write the same rows */
    int64_t max_block_w;     /* the widest device block (columns) */
    int coop_retried;        /* 1: the cooperative panel LU gave up at a barrier and the factorization ran again without it */
} fplus_amd_stats;
int fplus_amd_numeric_stats(const fplus_numeric *num, fplus_amd_stats *st);

/* Device-resident inputs, for callers that keep the values and right-hand sides in HBM (refactorization loops,
 * benchmarks): fplus_factor with Ax a device pointer (the analysed pattern Ap, Ai on the host), and fplus_solve on a
 * device array b (column-major, leading dimension ldb), in place. */
int fplus_amd_factor_device(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *d_Ax, int is_complex,
                            const fplus_options *opt, fplus_numeric **num, char *err, size_t errlen);
int fplus_amd_solve_device(const fplus_numeric *num, void *d_b, int64_t nrhs, int64_t ldb, int trans, char *err, size_t errlen);

/* Device memory and the library's stream for those callers: allocate, free, copy (kind 0 host to device, 1 device to
 * host, 2 device to device), wait for the library's stream. */
int fplus_amd_dev_alloc(size_t bytes, void **ptr);
int fplus_amd_dev_free(void *ptr);
int fplus_amd_dev_copy(void *dst, const void *src, size_t bytes, int kind);
int fplus_amd_dev_sync(void);

/* Tuning knobs (process-wide; tests and benchmarks).  "panel_lu_max_workgroups": the most workgroups of the
 * cooperative panel LU (at most 512, the default); 0 runs every sub-panel on the one-workgroup kernel.  Unknown names
 * are FPLUS_ERR_BADARG. */
int fplus_amd_tune(const char *name, int64_t value);

/* Per-kernel timing with HIP events on the library's stream (benchmarks): on = 1 starts recording and clears the sums,
 * on = 0 stops.  fplus_amd_profile_get returns, for a kernel name, the summed milliseconds, the launches and the summed
 * algorithmic work (flops, or bytes for the streaming kernels) of the launches recorded so far. */
int fplus_amd_profile(int on);
int fplus_amd_profile_get(const char *kernel, double *ms, int64_t *launches, double *work);

#ifdef __cplusplus
}
#endif
#endif /* FPLUS_AMD_H */
