/* fplus_oracle.c -- TEST INFRASTRUCTURE ONLY (see fplus_oracle.h): the CPU restatement of the FPLUS partial-pivoting
 * numeric path used as the checker of the HIP product.  Nothing in the product links or loads this file.
 *
 * Build: oracle/Makefile (gcc -std=c99 -O2 -ffp-contract=off): no contraction, so every operation rounds as written
 * and the results are bit-identical to the reference compiled the same way.
 *
 * Restated from (paths relative to the reference tree):
 *   input checks          src/fplus_internal.h:147-182
 *   permutation           src/pipeline.c:14-82 (the counting-sort route; the qsort route gives the same arrays)
 *   Ruiz                  src/equil.c:60-183 (CSC iteration), :186-441 (driver, range guard, CSR route)
 *   supernodal LU         src/factor_sn.c:97-210 (search, relaxed joins), :246-451 (column kernel), :609-808 (driver, L)
 *   triangular sweeps     src/solve.c:11-74
 *   composition / solve   src/numeric.c:39-131
 */
#define _GNU_SOURCE
#include "fplus_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void orc_msg(char *err, size_t errlen, const char *fmt, ...)
{
    if (!err || errlen == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

/* ---------------------------------------------------------------- input checks (src/fplus_internal.h:147-182) */
int orc_check_csc(int64_t n, const int64_t *ptr, const int64_t *idx, int canonical, char *err, size_t errlen)
{
    if (n < 0) { orc_msg(err, errlen, "n must be >= 0"); return ORC_BADARG; }
    if (ptr[0] != 0) { orc_msg(err, errlen, "indptr[0] must be 0"); return ORC_BADARG; }
    for (int64_t j = 0; j < n; j++) {
        if (ptr[j + 1] < ptr[j]) { orc_msg(err, errlen, "indptr decreases at %ld", (long)j); return ORC_BADARG; }
        for (int64_t q = ptr[j]; q < ptr[j + 1]; q++) {
            if (idx[q] < 0 || idx[q] >= n) { orc_msg(err, errlen, "index out of range in line %ld", (long)j); return ORC_BADARG; }
            if (canonical && q > ptr[j] && idx[q] <= idx[q - 1]) {
                orc_msg(err, errlen, "indices not strictly increasing in line %ld", (long)j);
                return ORC_BADARG;
            }
        }
    }
    return ORC_OK;
}

static int orc_check_finite(int64_t nnz, const double *x, int cplx, char *err, size_t errlen)
{
    const int64_t m = cplx ? 2 * nnz : nnz;
    for (int64_t k = 0; k < m; k++)
        if (!isfinite(x[k])) { orc_msg(err, errlen, "non-finite value in entry %ld", (long)(cplx ? k / 2 : k)); return ORC_BADARG; }
    return ORC_OK;
}

/* ---------------------------------------------------------------- permutation (src/pipeline.c:14-82)
 * B = A[row_perm, :][:, col_perm], canonical: the columns in col_perm order, their rows renumbered by the inverse row
 * order, then a transpose and a transpose back leave every column's rows ascending. */
int orc_permute(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                const int64_t *row_perm, const int64_t *col_perm, int64_t *Bp, int64_t *Bi, double *Bx)
{
    const int64_t nnz = Ap[n], w = cplx ? 2 : 1;
    int64_t *inv = malloc((size_t)(n + 1) * sizeof(int64_t));
    int64_t *rp = calloc((size_t)(n + 2), sizeof(int64_t));
    int64_t *tc = malloc((size_t)(nnz + 1) * sizeof(int64_t)), *ts = malloc((size_t)(nnz + 1) * sizeof(int64_t));
    int64_t *next = malloc((size_t)(n + 1) * sizeof(int64_t));
    if (!inv || !rp || !tc || !ts || !next) { free(inv); free(rp); free(tc); free(ts); free(next); return ORC_NOMEM; }
    for (int64_t k = 0; k < n; k++) inv[row_perm[k]] = k;
    Bp[0] = 0;
    for (int64_t k = 0; k < n; k++) Bp[k + 1] = Bp[k] + Ap[col_perm[k] + 1] - Ap[col_perm[k]];
    for (int64_t q = 0; q < nnz; q++) rp[inv[Ai[q]] + 1]++;
    for (int64_t r = 0; r < n; r++) rp[r + 1] += rp[r];
    memcpy(next, rp, (size_t)n * sizeof(int64_t));
    for (int64_t k = 0; k < n; k++)                       /* row-wise: (new column, source entry), columns ascending */
        for (int64_t q = Ap[col_perm[k]]; q < Ap[col_perm[k] + 1]; q++) {
            const int64_t t = next[inv[Ai[q]]]++;
            tc[t] = k; ts[t] = q;
        }
    memcpy(next, Bp, (size_t)n * sizeof(int64_t));
    for (int64_t r = 0; r < n; r++)                       /* back to columns: rows ascending */
        for (int64_t t = rp[r]; t < rp[r + 1]; t++) {
            const int64_t o = next[tc[t]]++;
            Bi[o] = r;
            memcpy(Bx + o * w, Ax + ts[t] * w, (size_t)w * sizeof(double));
        }
    free(inv); free(rp); free(tc); free(ts); free(next);
    return ORC_OK;
}

/* ---------------------------------------------------------------- Ruiz equilibration (src/equil.c)
 * Squared moduli, read as the reference reads them: re*re + im*im. */
static inline double orc_sq(const double *x, int64_t q, int cplx)
{
    if (cplx) { const double re = x[2 * q], im = x[2 * q + 1]; return re * re + im * im; }
    return x[q] * x[q];
}

static inline void orc_scale(double *x, int64_t q, int cplx, double s)
{
    if (cplx) { x[2 * q] *= s; x[2 * q + 1] *= s; } else x[q] *= s;
}

/* One iteration on line-compressed storage (lines = columns when is_csc, rows otherwise).  The maxima are taken over
 * rows first, then the rows are scaled and the column maxima of the scaled entries gathered, then the columns scaled
 * (src/equil.c:60-183 for CSC, :288-403 for CSR; the two give the same bits).  Returns 0, 1+i for an all-zero row i,
 * -(1+j) for an all-zero column j. */
static int64_t orc_ruiz_iter(int64_t n, const int64_t *ptr, const int64_t *idx, double *x, int cplx, int is_csc,
                             double *Dr, double *Dc, double *rmax, double *cmax, double *res_out)
{
    for (int64_t i = 0; i < n; i++) { rmax[i] = 0.0; cmax[i] = 0.0; }
    for (int64_t l = 0; l < n; l++)
        for (int64_t q = ptr[l]; q < ptr[l + 1]; q++) {
            const int64_t r = is_csc ? idx[q] : l;
            const double a = orc_sq(x, q, cplx);
            if (a > rmax[r]) rmax[r] = a;
        }
    for (int64_t i = 0; i < n; i++) if (rmax[i] == 0.0) return 1 + i;
    double res = 0.0;
    for (int64_t i = 0; i < n; i++) { const double v = 0.5 * fabs(log(rmax[i])); if (v > res) res = v; }
    for (int64_t i = 0; i < n; i++) { const double d = 1.0 / sqrt(sqrt(rmax[i])); Dr[i] *= d; rmax[i] = d; }
    for (int64_t l = 0; l < n; l++)
        for (int64_t q = ptr[l]; q < ptr[l + 1]; q++) {
            const int64_t r = is_csc ? idx[q] : l, c = is_csc ? l : idx[q];
            orc_scale(x, q, cplx, rmax[r]);
            const double a = orc_sq(x, q, cplx);
            if (a > cmax[c]) cmax[c] = a;
        }
    for (int64_t j = 0; j < n; j++) if (cmax[j] == 0.0) return -(1 + j);
    for (int64_t j = 0; j < n; j++) {
        const double v = 0.5 * fabs(log(cmax[j]));
        if (v > res) res = v;
        cmax[j] = 1.0 / sqrt(sqrt(cmax[j]));
        Dc[j] *= cmax[j];
    }
    for (int64_t l = 0; l < n; l++)
        for (int64_t q = ptr[l]; q < ptr[l + 1]; q++) orc_scale(x, q, cplx, cmax[is_csc ? l : idx[q]]);
    *res_out = res;
    return 0;
}

int orc_ruiz(int64_t n, const int64_t *ptr, const int64_t *idx, double *x, int cplx, int is_csc, int max_iter,
             double tol, double *Dr, double *Dc, int *iters, char *err, size_t errlen)
{
    int rc = orc_check_csc(n, ptr, idx, 0, err, errlen);
    if (rc == ORC_OK) rc = orc_check_finite(ptr[n], x, cplx, err, errlen);
    if (rc != ORC_OK) return rc;
    const int64_t nnz = ptr[n], m = cplx ? 2 * nnz : nnz;
    /* range guard (src/equil.c:193-241): an exact power-of-two prescale 2^k4, k4 a multiple of 4 */
    int k4 = 0;
    double amax = 0.0, amin = INFINITY;
    for (int64_t q = 0; q < m; q++) { const double a = fabs(x[q]); if (a > amax) amax = a; if (a > 0.0 && a < amin) amin = a; }
    if (amax > 0.0 && (amax > 0x1p500 || amin < 0x1p-500)) {
        int emax = ilogb(amax), emin = ilogb(amin);
        if (emax - emin > 990) {
            /* the span that matters: from the smallest row or column maximum (entries by max(|re|, |im|)) to amax */
            double *lmax = calloc((size_t)(2 * (n > 0 ? n : 1)), sizeof(double));
            if (!lmax) return ORC_NOMEM;
            for (int64_t l = 0; l < n; l++)
                for (int64_t q = ptr[l]; q < ptr[l + 1]; q++) {
                    const double a = cplx ? fmax(fabs(x[2 * q]), fabs(x[2 * q + 1])) : fabs(x[q]);
                    const int64_t r = is_csc ? idx[q] : l, c = is_csc ? l : idx[q];
                    if (a > lmax[r]) lmax[r] = a;
                    if (a > lmax[n + c]) lmax[n + c] = a;
                }
            double lmin = INFINITY;
            for (int64_t q = 0; q < 2 * n; q++) if (lmax[q] > 0.0 && lmax[q] < lmin) lmin = lmax[q];
            free(lmax);
            const int elmin = ilogb(lmin);
            if (emax - elmin > 990) {
                orc_msg(err, errlen, "line maxima span 2^%d .. 2^%d: wider than Ruiz equilibration can handle", elmin, emax);
                return ORC_BADARG;
            }
            emin = elmin;
        }
        k4 = -((emax + emin) / 2 / 4) * 4;
        const double sc = ldexp(1.0, k4);
        for (int64_t q = 0; q < m; q++) x[q] *= sc;
    }
    double *rmax = malloc((size_t)(n > 0 ? n : 1) * sizeof(double)), *cmax = malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    if (!rmax || !cmax) { free(rmax); free(cmax); return ORC_NOMEM; }
    for (int64_t i = 0; i < n; i++) { Dr[i] = 1.0; Dc[i] = 1.0; }
    int used = 0;
    for (int it = 0; it < max_iter; it++) {
        double res = 0.0;
        const int64_t bad = orc_ruiz_iter(n, ptr, idx, x, cplx, is_csc, Dr, Dc, rmax, cmax, &res);
        if (bad != 0) {
            free(rmax); free(cmax);
            if (bad > 0) orc_msg(err, errlen, "Row %ld of the matrix is entirely zero", (long)(bad - 1));
            else orc_msg(err, errlen, "Column %ld of the matrix is entirely zero", (long)(-bad - 1));
            return ORC_SINGULAR;
        }
        used = it + 1;
        if (res < tol) break;
    }
    free(rmax); free(cmax);
    if (k4 != 0) {   /* fold the prescale back (src/equil.c:436-439) */
        const double h = ldexp(1.0, k4 / 2);
        for (int64_t i = 0; i < n; i++) { Dr[i] *= h; Dc[i] *= h; }
    }
    if (iters) *iters = used;
    return ORC_OK;
}

/* ---------------------------------------------------------------- supernodal LU (src/factor_sn.c)
 * A supernode is a dense column-major block over its rows: the pivot rows of its columns first, in column order, then
 * the other rows.  It is open while columns may still join it. */
typedef struct {
    int64_t c0, w, nr, ld, cap;   /* first column, columns, rows, row capacity (= leading dimension), column capacity */
    int64_t *ri;                  /* row of each block row */
    void *v;                      /* values, column-major, ld x cap */
    int64_t *srch, srch_end;      /* closed: the search list (its rows past the pivot rows); end of its pivotal part,
                                     -1 before pruning */
    double nzero;                 /* explicit zeros stored by relaxed joins */
} OSn;

typedef struct {
    int64_t n;
    int cplx;
    size_t vs;
    OSn *sn;
    int64_t nsn;
    int64_t *col_sn, *pinv, *p;
    int64_t *bpos;                /* block row of a row in the open supernode, -1 elsewhere */
    int64_t *rmark, *smark;       /* visit stamps of rows and supernodes */
    int64_t *seg0;                /* first column reached per supernode, this column */
    int64_t *topo;                /* reached supernodes, topological order, at topo[top..n) */
    int64_t *cand, *miss;         /* non-pivotal rows of the column in discovery order; rows missing from the open one */
    int64_t *stack, *pos;         /* search stack and per-supernode resume position */
    int64_t relax;                /* fixed rule: at most this many rows of pattern mismatch */
    int relax_auto;
    int64_t rz_w[3], rz_cap;
    double rz_z[3];
    double pivot_rel;
    int64_t *Up, *Ui, nU, capU;
    void *Ux;
    void *acc, *tmp;
} OWork;

#define ORC_T double
#define ORC_ABS fabs
#define ORC_SUF d
#include "orc_numeric.inc"
#undef ORC_T
#undef ORC_ABS
#undef ORC_SUF
#define ORC_T double complex
#define ORC_ABS cabs
#define ORC_SUF z
#include "orc_numeric.inc"
#undef ORC_T
#undef ORC_ABS
#undef ORC_SUF

/* one row of the row kernel: sorted columns, values of the scalar type in v */
typedef struct {
    int64_t *c;
    void *v;
    int64_t nz, cap;
} orc_rrow;

/* position of column col in a sorted row, -1 when absent (src/factor_row.c:89-105) */
static int64_t orc_rrow_find(const orc_rrow *r, int64_t col)
{
    int64_t lo = 0, hi = r->nz - 1;
    while (lo <= hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (r->c[mid] == col) return mid;
        if (r->c[mid] < col) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

#define ORC_T double
#define ORC_ABS fabs
#define ORC_SUF d
#include "orc_row.inc"
#undef ORC_T
#undef ORC_ABS
#undef ORC_SUF
#define ORC_T double complex
#define ORC_ABS cabs
#define ORC_SUF z
#include "orc_row.inc"
#undef ORC_T
#undef ORC_ABS
#undef ORC_SUF

void orc_rowlu_free(orc_rowlu *f)
{
    if (!f) return;
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux); free(f->p); free(f->q); free(f);
}

/* the row kernel's entry point (src/factor_row.c:1191-1330: argument checks, rows loaded, the elimination) */
int orc_factor_row(int64_t n, const int64_t *Rp, const int64_t *Ri, const double *Rx, int cplx, double drop_tol, int strategy,
                   double pivot_rel, orc_rowlu **out, char *err, size_t errlen)
{
    *out = NULL;
    if (strategy < 0 || strategy > 2) { orc_msg(err, errlen, "pivot_strategy must be 0 (full), 1 (partial), or 2 (rook); got %d", strategy); return ORC_BADARG; }
    if (!(pivot_rel >= 0.0)) { orc_msg(err, errlen, "pivot_rel must be >= 0"); return ORC_BADARG; }
    if (!(drop_tol >= 0.0)) { orc_msg(err, errlen, "drop_tol must be >= 0 (got %g)", drop_tol); return ORC_BADARG; }
    int rc = orc_check_csc(n, Rp, Ri, 1, err, errlen);
    if (rc == ORC_OK) rc = orc_check_finite(Rp[n], Rx, cplx, err, errlen);
    if (rc != ORC_OK) return rc;
    const size_t vs = cplx ? sizeof(double complex) : sizeof(double);
    orc_rowlu *f = (orc_rowlu *)calloc(1, sizeof(orc_rowlu));
    orc_rrow *M = (orc_rrow *)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_rrow));
    orc_rrow *Lm = (orc_rrow *)calloc((size_t)(n > 0 ? n : 1), sizeof(orc_rrow));
    if (f) { f->p = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t)); f->q = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t)); }
    rc = (!f || !M || !Lm || !f->p || !f->q) ? ORC_NOMEM : ORC_OK;
    if (rc == ORC_OK) {
        f->n = n; f->cplx = cplx; f->fail_step = -1;
        for (int64_t i = 0; i < n; i++) { f->p[i] = i; f->q[i] = i; }
        for (int64_t i = 0; i < n && rc == ORC_OK; i++) {
            const int64_t nz = Rp[i + 1] - Rp[i];
            M[i].c = (int64_t *)malloc((size_t)(nz > 0 ? nz : 1) * sizeof(int64_t));
            M[i].v = malloc((size_t)(nz > 0 ? nz : 1) * vs);
            if (!M[i].c || !M[i].v) { rc = ORC_NOMEM; break; }
            M[i].cap = nz > 0 ? nz : 1;
            M[i].nz = nz;
            memcpy(M[i].c, Ri + Rp[i], (size_t)nz * sizeof(int64_t));
            memcpy(M[i].v, (const char *)Rx + (size_t)Rp[i] * vs, (size_t)nz * vs);
        }
    }
    if (rc == ORC_OK)
        rc = cplx ? row_kernel_z(M, Lm, n, Rp, (const double complex *)Rx, drop_tol, strategy, pivot_rel, f, err, errlen)
                  : row_kernel_d(M, Lm, n, Rp, Rx, drop_tol, strategy, pivot_rel, f, err, errlen);
    if (rc == ORC_NOMEM) orc_msg(err, errlen, "out of memory");
    for (int64_t i = 0; M && Lm && i < n; i++) { free(M[i].c); free(M[i].v); free(Lm[i].c); free(Lm[i].v); }
    free(M); free(Lm);
    if (rc != ORC_OK) {
        if (f && rc == ORC_SINGULAR) { f->Lp = f->Li = f->Up = f->Ui = NULL; f->Lx = f->Ux = NULL; *out = f; return rc; }
        orc_rowlu_free(f);
        return rc;
    }
    *out = f;
    return ORC_OK;
}

/* the rows a search reads from supernode s: the open one's rows past its pivot rows, a closed one's list (only its
 * pivotal part once pruned) (src/factor_sn.c:97-108) */
static inline const int64_t *orc_srch_rows(const OWork *W, int64_t s, int64_t open, int64_t *end)
{
    const OSn *S = &W->sn[s];
    if (s == open) { *end = S->nr - S->w; return S->ri + S->w; }
    *end = S->srch_end >= 0 ? S->srch_end : S->nr - S->w;
    return S->srch;
}

/* depth-first search from supernode `root`, reached at column `col`: finished supernodes go to topo[--top],
 * non-pivotal rows met on the way to cand in discovery order (src/factor_sn.c:110-154) */
static int64_t orc_dfs(OWork *W, int64_t root, int64_t col, int64_t top, int64_t stamp, int64_t open, int64_t *ncand)
{
    int64_t sp = 0;
    W->stack[0] = root; W->smark[root] = stamp; W->seg0[root] = col; W->pos[root] = 0;
    while (sp >= 0) {
        const int64_t s = W->stack[sp];
        int64_t end;
        const int64_t *rows = orc_srch_rows(W, s, open, &end);
        int64_t q = W->pos[s], t = -1;
        for (; q < end; q++) {
            const int64_t i = rows[q], k = W->pinv[i];
            if (k < 0) {
                if (W->rmark[i] != stamp) { W->rmark[i] = stamp; W->cand[(*ncand)++] = i; }
                continue;
            }
            const int64_t u = W->col_sn[k];
            if (W->smark[u] == stamp) { if (k < W->seg0[u]) W->seg0[u] = k; continue; }
            t = u;
            W->smark[u] = stamp; W->seg0[u] = k;
            break;
        }
        if (t >= 0) { W->pos[s] = q + 1; W->pos[t] = 0; W->stack[++sp] = t; }
        else { sp--; W->topo[--top] = s; }
    }
    return top;
}

/* the reach of column j: its pattern rows, non-pivotal ones listed, pivotal ones searched from (src/factor_sn.c:261-271) */
static int64_t orc_reach(OWork *W, const int64_t *Ap, const int64_t *Ai, int64_t j, int64_t open, int64_t *ncand)
{
    const int64_t stamp = j + 1;
    int64_t top = W->n;
    *ncand = 0;
    for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
        const int64_t i = Ai[q], k = W->pinv[i];
        if (k < 0) {
            if (W->rmark[i] != stamp) { W->rmark[i] = stamp; W->cand[(*ncand)++] = i; }
            continue;
        }
        const int64_t s = W->col_sn[k];
        if (W->smark[s] == stamp) { if (k < W->seg0[s]) W->seg0[s] = k; continue; }
        top = orc_dfs(W, s, k, top, stamp, open, ncand);
    }
    return top;
}

/* relaxed join test (src/factor_sn.c:163-173): fixed mismatch bound, or the automatic zero-share rule */
static int orc_relax_ok(const OWork *W, const OSn *S, int64_t n_miss, int64_t extra)
{
    const int64_t mm = n_miss + extra;
    if (!W->relax_auto) return mm <= W->relax;
    if (mm > W->rz_cap) return 0;
    const int64_t w = S->w + 1, r = S->nr + n_miss;
    if (w <= W->rz_w[0]) return 1;
    const double z = S->nzero + (double)n_miss * (double)S->w + (double)extra, tot = (double)w * (double)r;
    const double zmax = w <= W->rz_w[1] ? W->rz_z[0] : (w <= W->rz_w[2] ? W->rz_z[1] : W->rz_z[2]);
    return z <= zmax * tot;
}

/* rows appended to the open supernode as zeros in its stored columns (src/factor_sn.c:176-198) */
static int orc_add_rows(OSn *S, const int64_t *rows, int64_t k, int64_t *bpos, size_t vs)
{
    const int64_t r0 = S->nr, nr = r0 + k;
    if (nr > S->ld) {
        const int64_t ld = nr + nr / 4 + 4;
        int64_t *ri = realloc(S->ri, (size_t)ld * sizeof(int64_t));
        if (!ri) return -1;
        S->ri = ri;
        char *v = malloc((size_t)S->cap * (size_t)ld * vs);
        if (!v) return -1;
        for (int64_t c = 0; c < S->w; c++) memcpy(v + (size_t)c * ld * vs, (char *)S->v + (size_t)c * S->ld * vs, (size_t)r0 * vs);
        free(S->v); S->v = v; S->ld = ld;
    }
    for (int64_t c = 0; c < S->w; c++) memset((char *)S->v + ((size_t)c * S->ld + r0) * vs, 0, (size_t)k * vs);
    for (int64_t t = 0; t < k; t++) { S->ri[r0 + t] = rows[t]; bpos[rows[t]] = r0 + t; }
    S->nr = nr;
    return 0;
}

static int orc_grow_cols(OSn *S, int64_t need, size_t vs)
{
    if (need <= S->cap) return 0;
    const int64_t cap = S->cap * 2 > need ? S->cap * 2 : need;
    void *v = realloc(S->v, (size_t)cap * (size_t)S->ld * vs);
    if (!v) return -1;
    S->v = v; S->cap = cap;
    return 0;
}

/* Eisenstat-Liu pruning of the closed supernodes the column reached whose search list holds its pivot row: the list is
 * partitioned, pivotal rows first (src/factor_sn.c:427-443) */
static void orc_prune(OWork *W, int64_t top, int64_t open, int64_t pivrow)
{
    for (int64_t t = top; t < W->n; t++) {
        const int64_t s = W->topo[t];
        OSn *S = &W->sn[s];
        if (s == open || S->srch_end >= 0) continue;
        const int64_t len = S->nr - S->w;
        int64_t *rows = S->srch, hit = 0;
        for (int64_t r = 0; r < len; r++) if (rows[r] == pivrow) { hit = 1; break; }
        if (!hit) continue;
        int64_t a = 0, b = len;
        while (a < b) {
            if (W->pinv[rows[a]] >= 0) { a++; continue; }
            b--;
            const int64_t t2 = rows[a]; rows[a] = rows[b]; rows[b] = t2;
        }
        S->srch_end = b;
    }
}

typedef struct { int64_t key, rr; } OKey;
static int okey_cmp(const void *a, const void *b)
{
    const int64_t x = ((const OKey *)a)->key, y = ((const OKey *)b)->key;
    return (x > y) - (x < y);
}

void orc_lu_free(orc_lu *f)
{
    if (!f) return;
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux); free(f->p); free(f->sn_first); free(f);
}

static void orc_work_free(OWork *W)
{
    if (W->sn) for (int64_t s = 0; s < W->nsn; s++) { free(W->sn[s].ri); free(W->sn[s].v); free(W->sn[s].srch); }
    free(W->sn); free(W->col_sn); free(W->pinv); free(W->p); free(W->bpos); free(W->rmark); free(W->smark);
    free(W->seg0); free(W->topo); free(W->cand); free(W->miss); free(W->stack); free(W->pos);
    free(W->Up); free(W->Ui); free(W->Ux); free(W->acc); free(W->tmp);
}

/* the column loop: reach, updates, pivot, U, L join or new supernode, clear, prune (src/factor_sn.c:246-451) */
static int orc_columns(OWork *W, const int64_t *Ap, const int64_t *Ai, const double *Ax, char *err, size_t errlen)
{
    const int64_t n = W->n;
    const size_t vs = W->vs;
    int64_t open = -1;
    for (int64_t j = 0; j < n; j++) {
        int64_t ncand;
        const int64_t top = orc_reach(W, Ap, Ai, j, open, &ncand);
        double colmax, best;
        int64_t pivrow;
        if (W->cplx) {
            double complex *x = W->acc;
            colmax = scatter_column_z(x, Ap, Ai, (const double complex *)Ax, j);
            for (int64_t t = top; t < n; t++) { const OSn *S = &W->sn[W->topo[t]]; apply_supernode_z(S, W->seg0[W->topo[t]] - S->c0, x, W->tmp); }
            pivrow = pick_pivot_z(x, W->cand, ncand, &best);
        } else {
            double *x = W->acc;
            colmax = scatter_column_d(x, Ap, Ai, Ax, j);
            for (int64_t t = top; t < n; t++) { const OSn *S = &W->sn[W->topo[t]]; apply_supernode_d(S, W->seg0[W->topo[t]] - S->c0, x, W->tmp); }
            pivrow = pick_pivot_d(x, W->cand, ncand, &best);
        }
        const double thr = W->pivot_rel * colmax;
        if (pivrow < 0 || best <= 0.0 || best < thr) {
            orc_msg(err, errlen, "Near-zero or zero pivot %.6e at step %ld (threshold %.6e). Matrix is singular.",
                    best > 0 ? best : 0.0, (long)j, thr);
            return ORC_SINGULAR;
        }
        /* U(:, j): every reached segment in topological order, the diagonal last (src/factor_sn.c:329-345) */
        {
            int64_t need = 1;
            for (int64_t t = top; t < n; t++) { const OSn *S = &W->sn[W->topo[t]]; need += S->c0 + S->w - W->seg0[W->topo[t]]; }
            if (W->nU + need > W->capU) {
                int64_t cap = W->capU * 2 > W->nU + need ? W->capU * 2 : W->nU + need;
                int64_t *ui = realloc(W->Ui, (size_t)cap * sizeof(int64_t));
                if (!ui) return ORC_NOMEM;
                W->Ui = ui;
                void *ux = realloc(W->Ux, (size_t)cap * vs);
                if (!ux) return ORC_NOMEM;
                W->Ux = ux; W->capU = cap;
            }
            int64_t u = W->nU;
            for (int64_t t = top; t < n; t++) {
                const OSn *S = &W->sn[W->topo[t]];
                for (int64_t cc = W->seg0[W->topo[t]] - S->c0; cc < S->w; cc++) {
                    W->Ui[u] = S->c0 + cc;
                    memcpy((char *)W->Ux + (size_t)u * vs, (char *)W->acc + (size_t)S->ri[cc] * vs, vs);
                    u++;
                }
            }
            W->Ui[u] = j;
            memcpy((char *)W->Ux + (size_t)u * vs, (char *)W->acc + (size_t)pivrow * vs, vs);
            u++;
            W->nU = u; W->Up[j + 1] = u;
        }
        /* L(:, j): join the open supernode (exactly, or relaxed), else close it and open one (src/factor_sn.c:347-416) */
        int join = 0;
        int64_t n_miss = 0;
        if (open >= 0) {
            OSn *S = &W->sn[open];
            int64_t in_S = 0;
            for (int64_t s = 0; s < ncand; s++) {
                if (W->bpos[W->cand[s]] >= S->w) in_S++;
                else W->miss[n_miss++] = W->cand[s];
            }
            const int64_t extra = (S->nr - S->w) - in_S;
            if (n_miss + extra == 0) join = 1;
            else if (orc_relax_ok(W, S, n_miss, extra)) join = 2;
            if (join == 2) {
                S->nzero += (double)n_miss * (double)S->w + (double)extra;
                if (n_miss > 0 && orc_add_rows(S, W->miss, n_miss, W->bpos, vs)) return ORC_NOMEM;
            }
        }
        if (join) {
            OSn *S = &W->sn[open];
            const int64_t w = S->w;
            if (orc_grow_cols(S, w + 1, vs)) return ORC_NOMEM;
            const int64_t pp = W->bpos[pivrow];
            if (pp != w) {
                const int64_t other = S->ri[w];
                S->ri[w] = pivrow; S->ri[pp] = other;
                W->bpos[pivrow] = w; W->bpos[other] = pp;
                if (W->cplx) swap_rows_z(S, pp, w); else swap_rows_d(S, pp, w);
            }
            if (W->cplx) { const double complex *x = W->acc; fill_join_z(S, x, x[pivrow]); }
            else { const double *x = W->acc; fill_join_d(S, x, x[pivrow]); }
            S->w = w + 1;
            W->col_sn[j] = open;
        } else {
            if (open >= 0) {   /* close: its rows past the pivot rows become its search list */
                OSn *S = &W->sn[open];
                const int64_t nrest = S->nr - S->w;
                S->srch = malloc((size_t)(nrest > 0 ? nrest : 1) * sizeof(int64_t));
                if (!S->srch) return ORC_NOMEM;
                memcpy(S->srch, S->ri + S->w, (size_t)nrest * sizeof(int64_t));
                for (int64_t rr = 0; rr < S->nr; rr++) W->bpos[S->ri[rr]] = -1;
            }
            const int64_t s = W->nsn++;
            OSn *S = &W->sn[s];
            S->c0 = j; S->w = 1; S->nr = ncand; S->ld = ncand; S->cap = 4; S->nzero = 0;
            S->srch = NULL; S->srch_end = -1;
            S->ri = malloc((size_t)(ncand > 0 ? ncand : 1) * sizeof(int64_t));
            S->v = malloc((size_t)S->cap * (size_t)(ncand > 0 ? ncand : 1) * vs);
            if (!S->ri || !S->v) return ORC_NOMEM;
            if (W->cplx) { const double complex *x = W->acc; fill_new_z(S, x, x[pivrow], pivrow, W->cand, ncand, W->bpos); }
            else { const double *x = W->acc; fill_new_d(S, x, x[pivrow], pivrow, W->cand, ncand, W->bpos); }
            W->col_sn[j] = s;
            open = s;
        }
        W->pinv[pivrow] = j; W->p[j] = pivrow;
        /* clear the accumulator on the pattern and the segments (src/factor_sn.c:419-425) */
        for (int64_t t = 0; t < ncand; t++) memset((char *)W->acc + (size_t)W->cand[t] * vs, 0, vs);
        for (int64_t t = top; t < n; t++) {
            const OSn *S = &W->sn[W->topo[t]];
            for (int64_t cc = W->seg0[W->topo[t]] - S->c0; cc < S->w; cc++) memset((char *)W->acc + (size_t)S->ri[cc] * vs, 0, vs);
        }
        memset((char *)W->acc + (size_t)pivrow * vs, 0, vs);
        orc_prune(W, top, open, pivrow);
    }
    return ORC_OK;
}

int orc_factor_sn(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx, int64_t relax_rows,
                  double pivot_rel, orc_lu **out, char *err, size_t errlen)
{
    *out = NULL;
    if (!(pivot_rel >= 0.0)) { orc_msg(err, errlen, "pivot_rel must be >= 0"); return ORC_BADARG; }
    int rc = orc_check_csc(n, Ap, Ai, 0, err, errlen);
    if (rc == ORC_OK) rc = orc_check_finite(Ap[n], Ax, cplx, err, errlen);
    if (rc != ORC_OK) return rc;
    const size_t nn = (size_t)(n > 0 ? n : 1), vs = (cplx ? 2 : 1) * sizeof(double);
    const int64_t nnzA = Ap[n];
    OWork W;
    memset(&W, 0, sizeof(W));
    W.n = n; W.cplx = cplx; W.vs = vs; W.pivot_rel = pivot_rel;
    /* relax rule and its automatic form (src/factor_sn.c:636-649) */
    W.relax = relax_rows > 0 ? relax_rows : 0;
    W.relax_auto = relax_rows < 0;
    const int small_sparse = n <= 4096 && nnzA <= 24 * n;
    if (small_sparse) { W.rz_w[0] = 2; W.rz_w[1] = 8; W.rz_w[2] = 32; W.rz_z[0] = 0.3; W.rz_z[1] = 0.03; W.rz_z[2] = 0.01; W.rz_cap = 8; }
    else { W.rz_w[0] = 4; W.rz_w[1] = 16; W.rz_w[2] = 48; W.rz_z[0] = 0.8; W.rz_z[1] = 0.1; W.rz_z[2] = 0.05; W.rz_cap = 32; }
    W.sn = calloc(nn, sizeof(OSn));
    W.col_sn = malloc(nn * sizeof(int64_t)); W.pinv = malloc(nn * sizeof(int64_t)); W.p = malloc(nn * sizeof(int64_t));
    W.bpos = malloc(nn * sizeof(int64_t)); W.rmark = calloc(nn, sizeof(int64_t)); W.smark = calloc(nn, sizeof(int64_t));
    W.seg0 = malloc(nn * sizeof(int64_t)); W.topo = malloc(nn * sizeof(int64_t)); W.cand = malloc(nn * sizeof(int64_t));
    W.miss = malloc(nn * sizeof(int64_t)); W.stack = malloc(nn * sizeof(int64_t)); W.pos = malloc(nn * sizeof(int64_t));
    W.capU = 4 * nnzA + n + 16;
    W.Up = malloc((nn + 1) * sizeof(int64_t)); W.Ui = malloc((size_t)W.capU * sizeof(int64_t)); W.Ux = malloc((size_t)W.capU * vs);
    W.acc = calloc(nn, vs); W.tmp = malloc(2 * nn * vs);
    orc_lu *f = calloc(1, sizeof(orc_lu));
    if (!f || !W.sn || !W.col_sn || !W.pinv || !W.p || !W.bpos || !W.rmark || !W.smark || !W.seg0 || !W.topo || !W.cand
        || !W.miss || !W.stack || !W.pos || !W.Up || !W.Ui || !W.Ux || !W.acc || !W.tmp) {
        free(f); orc_work_free(&W); orc_msg(err, errlen, "out of memory"); return ORC_NOMEM;
    }
    for (int64_t i = 0; i < n; i++) { W.pinv[i] = -1; W.bpos[i] = -1; }
    W.Up[0] = 0;
    rc = orc_columns(&W, Ap, Ai, Ax, err, errlen);
    if (rc != ORC_OK) { if (rc == ORC_NOMEM) orc_msg(err, errlen, "out of memory"); free(f); orc_work_free(&W); return rc; }
    /* L in CSC, rows in pivot order: unit diagonal, the supernode's later pivot rows, then its other rows ascending by
     * pivot index (src/factor_sn.c:741-788) */
    int64_t nnzL = 0, maxrest = 0;
    for (int64_t s = 0; s < W.nsn; s++) {
        const OSn *S = &W.sn[s];
        for (int64_t cc = 0; cc < S->w; cc++) nnzL += S->nr - cc;
        if (S->nr - S->w > maxrest) maxrest = S->nr - S->w;
    }
    f->n = n; f->cplx = cplx; f->nsn = W.nsn;
    f->Lp = malloc((nn + 1) * sizeof(int64_t)); f->Li = malloc((size_t)(nnzL > 0 ? nnzL : 1) * sizeof(int64_t));
    f->Lx = malloc((size_t)(nnzL > 0 ? nnzL : 1) * vs); f->sn_first = malloc((size_t)(W.nsn + 1) * sizeof(int64_t));
    OKey *ord = malloc((size_t)(maxrest > 0 ? maxrest : 1) * sizeof(OKey));
    if (!f->Lp || !f->Li || !f->Lx || !f->sn_first || !ord) {
        free(ord); orc_lu_free(f); orc_work_free(&W); orc_msg(err, errlen, "out of memory"); return ORC_NOMEM;
    }
    int64_t o = 0;
    for (int64_t s = 0; s < W.nsn; s++) {
        const OSn *S = &W.sn[s];
        const int64_t nrest = S->nr - S->w;
        f->sn_first[s] = S->c0;
        for (int64_t t = 0; t < nrest; t++) { ord[t].key = W.pinv[S->ri[S->w + t]]; ord[t].rr = S->w + t; }
        qsort(ord, (size_t)nrest, sizeof(OKey), okey_cmp);
        for (int64_t cc = 0; cc < S->w; cc++) {
            const int64_t j = S->c0 + cc;
            f->Lp[j] = o;
            f->Li[o] = j;
            if (cplx) { f->Lx[2 * o] = 1.0; f->Lx[2 * o + 1] = 0.0; } else f->Lx[o] = 1.0;
            o++;
            for (int64_t rr = cc + 1; rr < S->w; rr++, o++) {
                f->Li[o] = W.pinv[S->ri[rr]];
                memcpy((char *)f->Lx + (size_t)o * vs, (char *)S->v + ((size_t)cc * S->ld + rr) * vs, vs);
            }
            for (int64_t t = 0; t < nrest; t++, o++) {
                f->Li[o] = ord[t].key;
                memcpy((char *)f->Lx + (size_t)o * vs, (char *)S->v + ((size_t)cc * S->ld + ord[t].rr) * vs, vs);
            }
        }
    }
    free(ord);
    f->Lp[n] = o;
    f->sn_first[W.nsn] = n;
    f->Up = W.Up; f->Ui = W.Ui; f->Ux = W.Ux; f->p = W.p;
    W.Up = NULL; W.Ui = NULL; W.Ux = NULL; W.p = NULL;
    orc_work_free(&W);
    *out = f;
    return ORC_OK;
}

/* ---------------------------------------------------------------- triangular sweeps (src/solve.c:11-74) */
#define ORC_SWEEP(NAME, T, CONJ)                                                                                    \
static int NAME(int64_t n, const int64_t *Lp, const int64_t *Li, const T *Lx, const int64_t *Up, const int64_t *Ui, \
                const T *Ux, T *x, int trans)                                                                       \
{                                                                                                                   \
    if (trans == 0) {                                                                                               \
        for (int64_t j = 0; j < n; j++) {              /* L y = b, column by column */                              \
            const T xj = x[j];                                                                                      \
            if (xj == 0) continue;                                                                                  \
            for (int64_t q = Lp[j]; q < Lp[j + 1]; q++) if (Li[q] != j) x[Li[q]] -= xj * Lx[q];                     \
        }                                                                                                           \
        for (int64_t j = n - 1; j >= 0; j--) {         /* U x = y */                                                \
            T d = 0; int have = 0;                                                                                  \
            for (int64_t q = Up[j]; q < Up[j + 1]; q++) if (Ui[q] == j) { d = Ux[q]; have = 1; break; }             \
            if (!have || d == 0) return -1;                                                                         \
            const T xj = x[j] / d;                                                                                  \
            x[j] = xj;                                                                                              \
            if (xj == 0) continue;                                                                                  \
            for (int64_t q = Up[j]; q < Up[j + 1]; q++) if (Ui[q] < j) x[Ui[q]] -= xj * Ux[q];                      \
        }                                                                                                           \
        return 0;                                                                                                   \
    }                                                                                                               \
    const int cj = trans == 2;                                                                                      \
    for (int64_t j = 0; j < n; j++) {                  /* U^T y = b: dot products */                                \
        T s = x[j], d = 0; int have = 0;                                                                            \
        for (int64_t q = Up[j]; q < Up[j + 1]; q++) {                                                               \
            const T v = cj ? CONJ(Ux[q]) : Ux[q];                                                                   \
            if (Ui[q] < j) s -= v * x[Ui[q]];                                                                       \
            else if (Ui[q] == j) { d = v; have = 1; }                                                               \
        }                                                                                                           \
        if (!have || d == 0) return -1;                                                                             \
        x[j] = s / d;                                                                                               \
    }                                                                                                               \
    for (int64_t j = n - 1; j >= 0; j--) {             /* L^T x = y */                                              \
        T s = x[j];                                                                                                 \
        for (int64_t q = Lp[j]; q < Lp[j + 1]; q++) if (Li[q] > j) s -= (cj ? CONJ(Lx[q]) : Lx[q]) * x[Li[q]];      \
        x[j] = s;                                                                                                   \
    }                                                                                                               \
    return 0;                                                                                                       \
}
static inline double orc_id(double v) { return v; }
ORC_SWEEP(orc_sweep_d, double, orc_id)
ORC_SWEEP(orc_sweep_z, double complex, conj)

int orc_tri_solve(int64_t n, const int64_t *Lp, const int64_t *Li, const double *Lx, const int64_t *Up,
                  const int64_t *Ui, const double *Ux, double *x, int trans, int cplx)
{
    if (trans < 0 || trans > 2 || n < 0) return ORC_BADARG;
    const int rc = cplx ? orc_sweep_z(n, Lp, Li, (const double complex *)Lx, Up, Ui, (const double complex *)Ux, (double complex *)x, trans)
                        : orc_sweep_d(n, Lp, Li, Lx, Up, Ui, Ux, x, trans);
    return rc == 0 ? ORC_OK : ORC_SINGULAR;
}

/* ---------------------------------------------------------------- solve-plan supernodes (src/solve_plan.c:97-199) */
int64_t orc_plan_nsupernodes(int64_t n, const int64_t *Lp, const int64_t *Li, const int64_t *Up, const int64_t *Ui)
{
    int64_t j, q, ns = 0, nseg = 0, odd = 0;
    for (j = 0; j < n; j++)                       /* the layout: L(j, j) first in column j (:101-102) */
        if (Lp[j] >= Lp[j + 1] || Li[Lp[j]] != j) return -1;
    for (j = 0; j < n; j++) {                     /* column j+1 continues j: (j, j+1, R) then (j+1, R) (:108-117) */
        int join = 0;
        if (j + 1 < n) {
            const int64_t a = Lp[j], la = Lp[j + 1] - a, b = Lp[j + 1], lb = Lp[j + 2] - b;
            join = la >= 2 && lb == la - 1 && Li[a + 1] == j + 1;
            for (q = 2; join && q < la; q++) join = Li[a + q] == Li[b + q - 1];
        }
        if (!join) ns++;
    }
    for (j = 0; j < n; j++) {                     /* U segments and odd entries (:137-150) */
        int64_t prev = -2; int diag = 0;
        for (q = Up[j]; q < Up[j + 1]; q++) {
            const int64_t i = Ui[q];
            if (i > j || (i == j && diag)) odd++;
            if (i == j) diag = 1;
            if (i >= j) { prev = -2; continue; }
            if (i != prev + 1) nseg++;
            prev = i;
        }
    }
    for (j = 0; j < n && !odd; j++)               /* L entries on or above the diagonal after the first (:154-155) */
        for (q = Lp[j] + 1; q < Lp[j + 1]; q++) if (Li[q] <= j) { odd++; break; }
    /* plain sweeps (:156, :162): no segment lists, no overlap test */
    if (!odd && Lp[n] + Up[n] < 4000000 && (nseg == 0 || Up[n] - n < 35 * nseg) && n <= INT32_MAX) return ns;
    {   /* blocked: a row repeated above the diagonal of a U column makes segments overlap (:189-199) */
        int64_t *last = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
        if (!last) return -2;
        for (j = 0; j < n; j++) last[j] = -1;
        for (j = 0; j < n; j++)
            for (q = Up[j]; q < Up[j + 1]; q++) {
                const int64_t i = Ui[q];
                if (i >= j) continue;
                if (last[i] == j) { free(last); return -1; }
                last[i] = j;
            }
        free(last);
    }
    return ns;
}

/* ---------------------------------------------------------------- composition and solve (src/numeric.c:39-131) */
void orc_num_free(orc_num *num)
{
    if (!num) return;
    orc_lu_free(num->lu); free(num->p); free(num->q); free(num->Dr); free(num->Dc); free(num);
}

int orc_factor_perm(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                    const int64_t *row_perm, const int64_t *col_perm, int equilibrate, int ruiz_max_iter,
                    double ruiz_tol, int64_t relax_rows, double pivot_rel, orc_num **out, char *err, size_t errlen)
{
    *out = NULL;
    int rc = orc_check_csc(n, Ap, Ai, 1, err, errlen);
    if (rc != ORC_OK) return rc;
    const size_t nn = (size_t)(n > 0 ? n : 1), w = cplx ? 2 : 1;
    const int64_t nnz = Ap[n];
    int64_t *Bp = malloc((nn + 1) * sizeof(int64_t)), *Bi = malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int64_t));
    double *Bx = malloc((size_t)(nnz > 0 ? nnz : 1) * w * sizeof(double));
    double *Drp = malloc(nn * sizeof(double)), *Dcp = malloc(nn * sizeof(double));
    orc_num *num = calloc(1, sizeof(orc_num));
    if (!Bp || !Bi || !Bx || !Drp || !Dcp || !num) { rc = ORC_NOMEM; goto done; }
    rc = orc_permute(n, Ap, Ai, Ax, cplx, row_perm, col_perm, Bp, Bi, Bx);
    if (rc != ORC_OK) goto done;
    if (equilibrate) rc = orc_ruiz(n, Bp, Bi, Bx, cplx, 1, ruiz_max_iter, ruiz_tol, Drp, Dcp, &num->ruiz_iters, err, errlen);
    else for (int64_t i = 0; i < n; i++) { Drp[i] = 1.0; Dcp[i] = 1.0; }
    if (rc != ORC_OK) goto done;
    rc = orc_factor_sn(n, Bp, Bi, Bx, cplx, relax_rows, pivot_rel, &num->lu, err, errlen);
    if (rc != ORC_OK) goto done;
    num->p = malloc(nn * sizeof(int64_t)); num->q = malloc(nn * sizeof(int64_t));
    num->Dr = malloc(nn * sizeof(double)); num->Dc = malloc(nn * sizeof(double));
    if (!num->p || !num->q || !num->Dr || !num->Dc) { rc = ORC_NOMEM; goto done; }
    for (int64_t k = 0; k < n; k++) {
        num->p[k] = row_perm[num->lu->p[k]];     /* p = row_perm[p_lu]; q = col_perm (the kernel's q is the identity) */
        num->q[k] = col_perm[k];
        num->Dr[row_perm[k]] = Drp[k];
        num->Dc[col_perm[k]] = Dcp[k];
    }
done:
    free(Bp); free(Bi); free(Bx); free(Drp); free(Dcp);
    if (rc != ORC_OK) { if (rc == ORC_NOMEM) orc_msg(err, errlen, "out of memory"); orc_num_free(num); return rc; }
    *out = num;
    return ORC_OK;
}

/* gather with scaling, the sweeps, scatter with scaling, per right-hand side (src/numeric.c:100-126) */
int orc_solve(const orc_num *num, double *b, int64_t nrhs, int64_t ldb, int trans)
{
    const orc_lu *f = num->lu;
    const int64_t n = f->n;
    if (trans < 0 || trans > 2 || nrhs < 0 || ldb < n) return ORC_BADARG;
    const int cplx = f->cplx;
    const int64_t *p = num->p, *q = num->q;
    const double *Dr = num->Dr, *Dc = num->Dc;
    const int tr = cplx ? trans : (trans ? 1 : 0);
    double *w = malloc((size_t)(n > 0 ? n : 1) * (cplx ? 2 : 1) * sizeof(double));
    if (!w) return ORC_NOMEM;
    for (int64_t r = 0; r < nrhs; r++) {
        if (cplx) {
            double complex *x = (double complex *)b + r * ldb, *ww = (double complex *)w;
            if (tr == 0) for (int64_t k = 0; k < n; k++) ww[k] = Dr[p[k]] * x[p[k]];
            else         for (int64_t k = 0; k < n; k++) ww[k] = Dc[q[k]] * x[q[k]];
        } else {
            double *x = b + r * ldb;
            if (tr == 0) for (int64_t k = 0; k < n; k++) w[k] = Dr[p[k]] * x[p[k]];
            else         for (int64_t k = 0; k < n; k++) w[k] = Dc[q[k]] * x[q[k]];
        }
        const int rc = orc_tri_solve(n, f->Lp, f->Li, f->Lx, f->Up, f->Ui, f->Ux, w, tr, cplx);
        if (rc != ORC_OK) { free(w); return rc; }
        if (cplx) {
            double complex *x = (double complex *)b + r * ldb; const double complex *ww = (const double complex *)w;
            if (tr == 0) for (int64_t k = 0; k < n; k++) x[q[k]] = Dc[q[k]] * ww[k];
            else         for (int64_t k = 0; k < n; k++) x[p[k]] = Dr[p[k]] * ww[k];
        } else {
            double *x = b + r * ldb;
            if (tr == 0) for (int64_t k = 0; k < n; k++) x[q[k]] = Dc[q[k]] * w[k];
            else         for (int64_t k = 0; k < n; k++) x[p[k]] = Dr[p[k]] * w[k];
        }
    }
    free(w);
    return ORC_OK;
}

/* the backward error of the reference's C tests (tests/c/test_fplus.c:74-89) */
double orc_backward_error(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx,
                          const double *x, const double *b, int trans)
{
    double nA = 0;
    for (int64_t j = 0; j < n; j++) {
        double c = 0;
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) c += cplx ? hypot(Ax[2 * q], Ax[2 * q + 1]) : fabs(Ax[q]);
        if (c > nA) nA = c;
    }
    const int64_t w = cplx ? 2 : 1;
    double *r = calloc((size_t)(n > 0 ? n : 1) * (size_t)w, sizeof(double));
    if (!r) return NAN;
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) {
            const int64_t i = Ai[q];
            if (cplx) {
                const double ar = Ax[2 * q], ai = (trans == 2 ? -1.0 : 1.0) * Ax[2 * q + 1];
                const int64_t src = trans == 0 ? j : i, dst = trans == 0 ? i : j;
                const double xr = x[2 * src], xi = x[2 * src + 1];
                r[2 * dst] += ar * xr - ai * xi;
                r[2 * dst + 1] += ar * xi + ai * xr;
            } else {
                if (trans == 0) r[i] += Ax[q] * x[j]; else r[j] += Ax[q] * x[i];
            }
        }
    double nr = 0, nx = 0, nb = 0;
    for (int64_t k = 0; k < n * w; k++) { const double d = r[k] - b[k]; nr += d * d; nx += x[k] * x[k]; nb += b[k] * b[k]; }
    free(r);
    return sqrt(nr) / (nA * sqrt(nx) + sqrt(nb));
}
