/* fplus.h -- the FPLUS sparse LU C interface, as exported by libfplus_amd.so (the MI355X build).
 *
 * Same names, signatures, struct layouts, enum values and error convention as the reference interface
 * (reference include/fplus.h:1-205), so a program or binding written against the reference links against this library
 * unchanged.  Each declaration cites the reference line it replaces.
 *
 * Index arrays are int64; values are double or double complex (interleaved re, im), chosen by is_complex.  Every entry
 * point returns FPLUS_OK or a negative FPLUS_ERR_* code and writes a message into err[errlen] when err is not NULL.
 *
 * What runs where: the symbolic analysis (dense-line pinning, the fill-reducing order, the column-etree postorder) and
 * the symbolic half of the supernodal factorization (reach, relaxed supernode joins, pruning) run on the host; Ruiz
 * equilibration, the permutation of A, the Schur-complement updates, the pivot search, the L / U stores and the
 * triangular solves run as HIP kernels on one MI355X (gfx950).  use_gpu = 1 in fplus_options (and in the low-level
 * calls) selects the device path; a missing or failing device then reports FPLUS_ERR_GPU: there is no silent host
 * fallback.  With use_gpu = 0 the same host-side code drives the identical device kernels (the library has no CPU
 * numeric path of its own).
 */
#ifndef FPLUS_H
#define FPLUS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (reference include/fplus.h:17-24) */
enum {
    FPLUS_OK = 0,
    FPLUS_ERR_NOMEM = -1,
    FPLUS_ERR_SINGULAR = -2,
    FPLUS_ERR_BADARG = -3,
    FPLUS_ERR_GPU = -4,
    FPLUS_ERR_INTERNAL = -5
};

/* pivot strategies (reference include/fplus.h:27-29) */
#define FPLUS_PIVOT_FULL    0
#define FPLUS_PIVOT_PARTIAL 1
#define FPLUS_PIVOT_ROOK    2

/* sqrt(DBL_EPSILON) (reference include/fplus.h:32) */
#define FPLUS_PIVOT_REL_DEFAULT 1.4901161193847656e-08

/* P A Q = L U, CSC factors, malloc'd by the library (reference include/fplus.h:37-46) */
typedef struct {
    int64_t n;
    int is_complex;
    int csc;
    int64_t *Lp, *Li; void *Lx;
    int64_t *Up, *Ui; void *Ux;
    int64_t *p, *q;
    int64_t nsupernodes;
    int64_t n_pivots_replaced;
} fplus_factors;

/* reference include/fplus.h:48 */
void fplus_factors_free(fplus_factors *f);

/* supernodal left-looking LU with partial pivoting of a CSC matrix (reference include/fplus.h:54-56) */
int fplus_factor_sn(int64_t n, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex, int verbose,
                    int64_t relax_rows, int64_t panel, int use_gpu, double gpu_min_flops, double pivot_rel,
                    fplus_factors *out, char *err, size_t errlen);

/* row-kernel LU, rook / full / partial pivoting (reference include/fplus.h:60-62).  Not part of the device path:
 * returns FPLUS_ERR_BADARG with a message naming the supernodal entry point. */
int fplus_factor_row(int64_t n, const int64_t *indptr, const int64_t *indices, const void *data, double drop_tol,
                     int is_complex, int verbose, int pivot_strategy, double pivot_rel,
                     fplus_factors *out, char *err, size_t errlen);

/* Ruiz L-inf equilibration in place, CSR or CSC (reference include/fplus.h:66-67) */
int fplus_ruiz(int64_t n, const int64_t *indptr, const int64_t *indices, void *data, int is_complex, int is_csc,
               int max_iter, double tol, int verbose, double *Dr, double *Dc, char *err, size_t errlen);

/* triangular solves on CSC factors, in place (reference include/fplus.h:72-73) */
int fplus_tri_solve(int64_t n, const int64_t *Lp, const int64_t *Li, const void *Lx, const int64_t *Up, const int64_t *Ui,
                    const void *Ux, void *x, int trans, int is_complex);

/* solve plans (reference include/fplus.h:81-88) */
typedef struct fplus_solve_plan fplus_solve_plan;
int fplus_solve_plan_create(int64_t n, const int64_t *Lp, const int64_t *Li, const int64_t *Up, const int64_t *Ui,
                            fplus_solve_plan **plan, char *err, size_t errlen);
int fplus_tri_solve_plan(const fplus_solve_plan *plan, int64_t n, const int64_t *Lp, const int64_t *Li, const void *Lx,
                         const int64_t *Up, const int64_t *Ui, const void *Ux, void *x, int64_t nrhs, int64_t ldx,
                         int trans, int is_complex);
int64_t fplus_solve_plan_nsupernodes(const fplus_solve_plan *plan);
void fplus_solve_plan_free(fplus_solve_plan *plan);

/* column-etree postorder (reference include/fplus.h:92) */
int fplus_coletree_postorder(int64_t m, int64_t n, const int64_t *Ap, const int64_t *Ai, int64_t *post);

/* symbolic analysis (reference include/fplus.h:95-120) */
typedef struct fplus_symbolic fplus_symbolic;
int fplus_analyze(int64_t n, const int64_t *Ap, const int64_t *Ai, int pin_dense, double tau_density, int64_t tau_min_nnz,
                  const unsigned char *tau_row_mask, const unsigned char *tau_col_mask, int postorder,
                  fplus_symbolic **sym, char *err, size_t errlen);
#define FPLUS_ORDER_COLAMD  0
#define FPLUS_ORDER_AMD_ATA 1
#define FPLUS_ORDER_AUTO    2
int fplus_analyze_ord(int64_t n, const int64_t *Ap, const int64_t *Ai, int pin_dense, double tau_density, int64_t tau_min_nnz,
                      const unsigned char *tau_row_mask, const unsigned char *tau_col_mask, int postorder, int ordering,
                      fplus_symbolic **sym, char *err, size_t errlen);
void fplus_symbolic_free(fplus_symbolic *sym);
int64_t fplus_symbolic_n(const fplus_symbolic *sym);
int64_t fplus_symbolic_nnz(const fplus_symbolic *sym);
const int64_t *fplus_symbolic_row_perm(const fplus_symbolic *sym);
const int64_t *fplus_symbolic_col_perm(const fplus_symbolic *sym);
int64_t fplus_symbolic_n_dense_rows(const fplus_symbolic *sym);
int64_t fplus_symbolic_n_dense_cols(const fplus_symbolic *sym);

/* the supernodal pipeline: permute, Ruiz, factor (reference include/fplus.h:125-134) */
int fplus_factor_sn_perm(int64_t n, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex,
                         const int64_t *row_perm, const int64_t *col_perm,
                         int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                         int64_t relax_rows, int64_t panel, int use_gpu, double gpu_min_flops, double pivot_rel,
                         fplus_factors *out, double *Dr, double *Dc, char *err, size_t errlen);
int fplus_factor_sn_sym(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex,
                        int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                        int64_t relax_rows, int64_t panel, int use_gpu, double gpu_min_flops, double pivot_rel,
                        fplus_factors *out, double *Dr, double *Dc, char *err, size_t errlen);

/* the row-kernel pipeline (reference include/fplus.h:139-147): FPLUS_ERR_BADARG, as fplus_factor_row */
int fplus_factor_row_perm(int64_t n, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex,
                          const int64_t *row_perm, const int64_t *col_perm,
                          int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                          double drop_tol, int pivot_strategy, double pivot_rel,
                          fplus_factors *out, double *Dr, double *Dc, char *err, size_t errlen);
int fplus_factor_row_sym(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex,
                         int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                         double drop_tol, int pivot_strategy, double pivot_rel,
                         fplus_factors *out, double *Dr, double *Dc, char *err, size_t errlen);

/* options, 14 fields in this order (reference include/fplus.h:150-166) */
typedef struct {
    int     pin_dense;
    double  tau_density;
    int64_t tau_min_nnz;
    int     postorder;
    int     equilibrate;
    int     ruiz_max_iter;
    double  ruiz_tol;
    double  pivot_rel;
    int     pivot_strategy;
    double  drop_tol;
    int64_t relax_rows;
    int64_t panel;
    int     use_gpu;
    double  gpu_min_flops;
    int     verbose;
} fplus_options;

/* reference include/fplus.h:168-200 */
void fplus_default_options(fplus_options *opt);
int fplus_analyze_opts(int64_t n, const int64_t *Ap, const int64_t *Ai, const fplus_options *opt,
                       fplus_symbolic **sym, char *err, size_t errlen);
typedef struct fplus_numeric fplus_numeric;
int fplus_factor(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *Ax, int is_complex,
                 const fplus_options *opt, fplus_numeric **num, char *err, size_t errlen);
int fplus_solve(const fplus_numeric *num, void *b, int64_t nrhs, int64_t ldb, int trans, char *err, size_t errlen);
const fplus_factors *fplus_numeric_factors(const fplus_numeric *num);
const int64_t *fplus_numeric_p(const fplus_numeric *num);
const int64_t *fplus_numeric_q(const fplus_numeric *num);
const double *fplus_numeric_Dr(const fplus_numeric *num);
const double *fplus_numeric_Dc(const fplus_numeric *num);
void fplus_numeric_free(fplus_numeric *num);

#ifdef __cplusplus
}
#endif
#endif /* FPLUS_H */
