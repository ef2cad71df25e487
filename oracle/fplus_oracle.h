/* fplus_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement, in plain C99, of the FPLUS numeric path: the permutation of A, Ruiz equilibration, the supernodal
 * left-looking column kernel for partial pivoting (the loop form the reference compiles to without a CBLAS), the
 * assembly of L, the triangular sweeps and the composed solve, and the row kernel (full / partial / rook pivoting and
 * the incomplete factorization).  It is the checker for the HIP product in
 * fplus_amd/: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and never as the thing
 * measured or shipped.
 *
 * Every function cites the reference file:line whose behaviour it restates.  The arithmetic is ordered as in the
 * reference so that, compiled without floating-point contraction, the results are bit-identical to it.
 */
#ifndef FPLUS_ORACLE_H
#define FPLUS_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_OK = 0, ORC_NOMEM = -1, ORC_SINGULAR = -2, ORC_BADARG = -3, ORC_INTERNAL = -5 };

/* factors of the permuted, equilibrated matrix: CSC L (unit diagonal first), CSC U (diagonal last), p[j] = pivot row of
 * column j (rows of the matrix given to the kernel), sn_first[s] = first column of supernode s (nsn + 1 entries) */
typedef struct {
    int64_t n;
    int32_t cplx;
    int32_t pad;
    int64_t nsn;
    int64_t *Lp, *Li; double *Lx;
    int64_t *Up, *Ui; double *Ux;
    int64_t *p;
    int64_t *sn_first;
} orc_lu;

int orc_check_csc(int64_t n, const int64_t *ptr, const int64_t *idx, int canonical, char *err, size_t errlen);
int orc_permute(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                const int64_t *row_perm, const int64_t *col_perm, int64_t *Bp, int64_t *Bi, double *Bx);
int orc_ruiz(int64_t n, const int64_t *ptr, const int64_t *idx, double *x, int cplx, int is_csc, int max_iter,
             double tol, double *Dr, double *Dc, int *iters, char *err, size_t errlen);
int orc_factor_sn(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx, int64_t relax_rows,
                  double pivot_rel, orc_lu **out, char *err, size_t errlen);
void orc_lu_free(orc_lu *f);
int orc_tri_solve(int64_t n, const int64_t *Lp, const int64_t *Li, const double *Lx, const int64_t *Up,
                  const int64_t *Ui, const double *Ux, double *x, int trans, int cplx);
/* the supernode count of a solve plan (src/solve_plan.c:97-118 with the layout tests of :101-102, :133-160, :189-199):
   -1 when L lacks the diagonal first in some column, or when the blocked layout meets a row repeated above the diagonal
   of a U column */
int64_t orc_plan_nsupernodes(int64_t n, const int64_t *Lp, const int64_t *Li, const int64_t *Up, const int64_t *Ui);
/* the composed numeric object: permute, Ruiz, factor, composition of p, q, Dr, Dc into the original coordinates */
typedef struct {
    orc_lu *lu;
    int64_t *p, *q;
    double *Dr, *Dc;
    int ruiz_iters;
} orc_num;
int orc_factor_perm(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                    const int64_t *row_perm, const int64_t *col_perm, int equilibrate, int ruiz_max_iter,
                    double ruiz_tol, int64_t relax_rows, double pivot_rel, orc_num **out, char *err, size_t errlen);
int orc_solve(const orc_num *num, double *b, int64_t nrhs, int64_t ldb, int trans);
void orc_num_free(orc_num *num);
/* the row kernel (src/factor_row.c:1191-1685): CSR in, CSR factors out -- L with its unit diagonal first in each row,
 * U from the diagonal on; p[k] / q[k] = the input row / column at position k; pivot_strategy 0 full, 1 partial, 2 rook;
 * drop_tol > 0 the incomplete factorization.  On ORC_SINGULAR *out still holds fail_step (the step that found no
 * pivot), p and q as they were then, and no factors. */
typedef struct {
    int64_t n;
    int32_t cplx;
    int32_t pad;
    int64_t *Lp, *Li; double *Lx;
    int64_t *Up, *Ui; double *Ux;
    int64_t *p, *q;
    int64_t n_replaced;
    int64_t fail_step;
} orc_rowlu;
int orc_factor_row(int64_t n, const int64_t *Rp, const int64_t *Ri, const double *Rx, int cplx, double drop_tol, int strategy,
                   double pivot_rel, orc_rowlu **out, char *err, size_t errlen);
void orc_rowlu_free(orc_rowlu *f);

/* ||op(A) x - b|| / (||A||_1 ||x|| + ||b||) */
double orc_backward_error(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                          const double *x, const double *b, int trans);

#ifdef __cplusplus
}
#endif
#endif
