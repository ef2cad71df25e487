/* cpu_panel.c -- TEST INFRASTRUCTURE ONLY: the CPU baseline that bench.py times beside the GPU path (its
 * cpu_baseline leg); nothing in the product links or loads it.
 *
 * The same numeric path as fplus_oracle.c -- the permutation into the analysed order, Ruiz equilibration, the
 * supernodal left-looking LU with threshold partial pivoting and relaxed supernodes, the solves -- blocked and run on
 * all host threads with OpenMP, the way the reference's panel kernel with a BLAS runs it (src/factor_sn.c:453-606, the
 * panel form of the column kernel at :246-451):
 *   - columns go in panels of `panel` columns; supernodes close at a panel's end;
 *   - phase A (parallel, chunks of CW columns per task): each chunk's reach over the supernodes of earlier panels
 *     (a depth-first search, one for the chunk's union), then those supernodes applied to the chunk's CW columns at
 *     once -- the triangular solve on the segment and the product with the rows below, dense over the chunk's columns,
 *     so every value of L read serves CW columns;
 *   - phase B (in column order): the panel's own supernodes applied to each column (their row-block products split
 *     over the threads when large), the pivot (the candidate of largest modulus, the first in discovery order on
 *     ties; a pivot below pivot_rel times the column's largest entry fails the factorization), U's column, and the
 *     column joined to the open supernode (its pattern equal up to relax_rows rows) or a new supernode opened.
 * The results are those of the column kernel up to rounding (the updates of a value add up in another order) and
 * supernode boundaries at panel ends.  Timing only: the checker of the product is fplus_oracle.c.
 * Build: oracle/Makefile (gcc -O3 -march=x86-64-v3 -fopenmp). */
#include "cpu_panel.h"
#include "fplus_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CW 8   /* columns per phase-A task */

typedef struct {
    int64_t c0, w, nr, ld, cap;   /* first column, columns, rows, leading dimension, column capacity */
    int64_t *ri;                  /* rows: the pivot rows of its columns first, then the others */
    double *v;                    /* values, column-major ld x cap scalars (vs doubles each) */
    int64_t *srch, nsrch;         /* closed: the rows past its pivot rows (its search list) */
} PSn;

typedef struct {
    int64_t n, nsn, capsn;
    int vs;                       /* doubles per scalar */
    PSn *sn;
    int64_t *col_sn, *pinv, *p;
    int64_t *Up, *Ui, nU, capU;
    double *Ux;
    double flops;
} PLU;

static void cp_msg(char *err, size_t errlen, const char *fmt, double a, long b, double c)
{
    if (err && errlen) snprintf(err, errlen, fmt, a, b, c);
}

/* ---- phase A: a chunk of up to CW columns against the supernodes of earlier panels */
typedef struct {
    int64_t nT;
    int64_t *T;      /* the rows touched, in discovery order */
    double *X;       /* their values, nT x CW scalars, row-major */
} Chunk;

typedef struct {
    double *X;       /* n x CW scalars, zero outside a task */
    double *Y;       /* the row-block product's rows below the segment */
    int64_t capY;
    int64_t *rmark, *smark, *seg0, *stack, *pos, *topo, *T;
    int64_t stamp;
} Ws;

static int ws_init(Ws *w, int64_t n, int vs)
{
    memset(w, 0, sizeof(*w));
    const size_t nn = (size_t)(n > 0 ? n : 1);
    w->X = calloc(nn * CW * vs, sizeof(double));
    w->rmark = calloc(nn, sizeof(int64_t)); w->smark = calloc(nn, sizeof(int64_t));
    w->seg0 = malloc(nn * sizeof(int64_t)); w->stack = malloc(nn * sizeof(int64_t)); w->pos = malloc(nn * sizeof(int64_t));
    w->topo = malloc(nn * sizeof(int64_t)); w->T = malloc(nn * sizeof(int64_t));
    return w->X && w->rmark && w->smark && w->seg0 && w->stack && w->pos && w->topo && w->T ? 0 : -1;
}
static void ws_free(Ws *w)
{
    free(w->X); free(w->Y); free(w->rmark); free(w->smark); free(w->seg0); free(w->stack); free(w->pos); free(w->topo); free(w->T);
}

/* touch row i: listed once per stamp */
static inline void touch(Ws *w, int64_t i, int64_t *nT)
{
    if (w->rmark[i] != w->stamp) { w->rmark[i] = w->stamp; w->T[(*nT)++] = i; }
}

/* the depth-first search over supernodes of earlier panels (pinv < j0) from supernode s reached at column k: finished
 * supernodes to topo[--top] */
static int64_t dfs_out(const PLU *F, Ws *w, int64_t s, int64_t k, int64_t top, int64_t j0, int64_t *nT)
{
    int64_t sp = 0;
    w->stack[0] = s; w->smark[s] = w->stamp; w->seg0[s] = k; w->pos[s] = 0;
    while (sp >= 0) {
        const int64_t u = w->stack[sp];
        const PSn *S = &F->sn[u];
        int64_t q = w->pos[u], nxt = -1;
        for (; q < S->nsrch; q++) {
            const int64_t i = S->srch[q], kk = F->pinv[i];
            touch(w, i, nT);
            if (kk < 0 || kk >= j0) continue;
            const int64_t t = F->col_sn[kk];
            if (w->smark[t] == w->stamp) { if (kk < w->seg0[t]) w->seg0[t] = kk; continue; }
            nxt = t; w->smark[t] = w->stamp; w->seg0[t] = kk;
            break;
        }
        if (nxt >= 0) { w->pos[u] = q + 1; w->pos[nxt] = 0; w->stack[++sp] = nxt; }
        else { sp--; w->topo[--top] = u; }
    }
    return top;
}

/* supernode S from its column kf on applied to the CW columns of X (rows of X: scalars of CW columns each) */
static void apply_block(const PSn *S, int64_t kf, double *X, double *Y, int vs)
{
    const int64_t w = S->w, ld = S->ld, nrest = S->nr - S->w;
    const int64_t *ri = S->ri;
    if (nrest > 0) memset(Y, 0, (size_t)nrest * CW * vs * sizeof(double));
    for (int64_t cc = kf; cc < w; cc++) {
        double xr[CW], xi[CW];
        int any = 0;
        const double *xs = X + ri[cc] * CW * vs;
        if (vs == 1) {
            for (int c = 0; c < CW; c++) { xr[c] = xs[c]; any |= xr[c] != 0.0; }
        } else {
            for (int c = 0; c < CW; c++) { xr[c] = xs[2 * c]; xi[c] = xs[2 * c + 1]; any |= (xr[c] != 0.0) | (xi[c] != 0.0); }
        }
        if (!any) continue;
        const double *col = S->v + (size_t)cc * ld * vs;
        if (vs == 1) {
            for (int64_t rr = cc + 1; rr < w; rr++) {
                double *xt = X + ri[rr] * CW;
                const double l = col[rr];
                for (int c = 0; c < CW; c++) xt[c] -= l * xr[c];
            }
            const double *below = col + w;
            for (int64_t r = 0; r < nrest; r++) {
                double *y = Y + r * CW;
                const double l = below[r];
                for (int c = 0; c < CW; c++) y[c] += l * xr[c];
            }
        } else {
            for (int64_t rr = cc + 1; rr < w; rr++) {
                double *xt = X + ri[rr] * CW * 2;
                const double lr = col[2 * rr], li = col[2 * rr + 1];
                for (int c = 0; c < CW; c++) {
                    xt[2 * c] -= lr * xr[c] - li * xi[c];
                    xt[2 * c + 1] -= lr * xi[c] + li * xr[c];
                }
            }
            const double *below = col + 2 * w;
            for (int64_t r = 0; r < nrest; r++) {
                double *y = Y + r * CW * 2;
                const double lr = below[2 * r], li = below[2 * r + 1];
                for (int c = 0; c < CW; c++) {
                    y[2 * c] += lr * xr[c] - li * xi[c];
                    y[2 * c + 1] += lr * xi[c] + li * xr[c];
                }
            }
        }
    }
    for (int64_t r = 0; r < nrest; r++) {
        double *xt = X + ri[w + r] * CW * vs;
        const double *y = Y + r * CW * vs;
        for (int c = 0; c < CW * vs; c++) xt[c] -= y[c];
    }
}

static int phase_a(const PLU *F, Ws *w, const int64_t *Bp, const int64_t *Bi, const double *Bx, int64_t c0, int64_t nc, int64_t j0,
                   Chunk *out)
{
    const int vs = F->vs;
    w->stamp++;
    int64_t nT = 0, top = F->n;
    for (int64_t c = 0; c < nc; c++) {
        const int64_t j = c0 + c;
        for (int64_t q = Bp[j]; q < Bp[j + 1]; q++) {
            const int64_t i = Bi[q];
            touch(w, i, &nT);
            double *x = w->X + (i * CW + c) * vs;
            x[0] += Bx[q * vs];
            if (vs == 2) x[1] += Bx[q * vs + 1];
            const int64_t k = F->pinv[i];
            if (k < 0 || k >= j0) continue;
            const int64_t s = F->col_sn[k];
            if (w->smark[s] == w->stamp) { if (k < w->seg0[s]) w->seg0[s] = k; continue; }
            top = dfs_out(F, w, s, k, top, j0, &nT);
        }
    }
    for (int64_t t = top; t < F->n; t++) {
        const PSn *S = &F->sn[w->topo[t]];
        const int64_t kf = w->seg0[w->topo[t]] - S->c0, nrest = S->nr - S->w;
        for (int64_t rr = kf; rr < S->nr; rr++) touch(w, S->ri[rr], &nT);
        if (nrest * CW * vs > w->capY) {
            free(w->Y);
            w->capY = nrest * CW * vs * 2;
            w->Y = malloc((size_t)w->capY * sizeof(double));
            if (!w->Y) return -1;
        }
        apply_block(S, kf, w->X, w->Y, vs);
    }
    out->nT = nT;
    out->T = malloc((size_t)(nT > 0 ? nT : 1) * sizeof(int64_t));
    out->X = malloc((size_t)(nT > 0 ? nT : 1) * CW * vs * sizeof(double));
    if (!out->T || !out->X) return -1;
    memcpy(out->T, w->T, (size_t)nT * sizeof(int64_t));
    for (int64_t t = 0; t < nT; t++) {
        double *x = w->X + w->T[t] * CW * vs;
        memcpy(out->X + t * CW * vs, x, (size_t)CW * vs * sizeof(double));
        memset(x, 0, (size_t)CW * vs * sizeof(double));
    }
    return 0;
}

/* ---- phase B: one column against the panel's own supernodes, pivot, U, L */
typedef struct {
    double *x;                    /* n scalars, zero outside a column */
    double *y;
    int64_t capy;
    int64_t *mark, *smark, *seg0, *stack, *pos, *topo, *pat, *cand, *miss, *bpos;
} Bw;

static int64_t dfs_in(const PLU *F, Bw *b, int64_t s, int64_t k, int64_t top, int64_t j0, int64_t open, int64_t stamp, int64_t *np)
{
    int64_t sp = 0;
    b->stack[0] = s; b->smark[s] = stamp; b->seg0[s] = k; b->pos[s] = 0;
    while (sp >= 0) {
        const int64_t u = b->stack[sp];
        const PSn *S = &F->sn[u];
        const int64_t *rows = u == open ? S->ri + S->w : S->srch;
        const int64_t end = u == open ? S->nr - S->w : S->nsrch;
        int64_t q = b->pos[u], nxt = -1;
        for (; q < end; q++) {
            const int64_t i = rows[q], kk = F->pinv[i];
            if (b->mark[i] != stamp) { b->mark[i] = stamp; b->pat[(*np)++] = i; }
            if (kk < j0) continue;
            const int64_t t = F->col_sn[kk];
            if (b->smark[t] == stamp) { if (kk < b->seg0[t]) b->seg0[t] = kk; continue; }
            nxt = t; b->smark[t] = stamp; b->seg0[t] = kk;
            break;
        }
        if (nxt >= 0) { b->pos[u] = q + 1; b->pos[nxt] = 0; b->stack[++sp] = nxt; }
        else { sp--; b->topo[--top] = u; }
    }
    return top;
}

/* supernode S from column kf applied to the dense column x; the row-block product split over the threads when large */
static void apply_col(const PSn *S, int64_t kf, double *x, double *y, int vs)
{
    const int64_t w = S->w, ld = S->ld, nrest = S->nr - S->w;
    const int64_t *ri = S->ri;
    for (int64_t cc = kf; cc < w; cc++) {
        const double *col = S->v + (size_t)cc * ld * vs;
        if (vs == 1) {
            const double xc = x[ri[cc]];
            if (xc == 0.0) continue;
            for (int64_t rr = cc + 1; rr < w; rr++) x[ri[rr]] -= col[rr] * xc;
        } else {
            const double xr = x[2 * ri[cc]], xi = x[2 * ri[cc] + 1];
            if (xr == 0.0 && xi == 0.0) continue;
            for (int64_t rr = cc + 1; rr < w; rr++) {
                const double lr = col[2 * rr], li = col[2 * rr + 1];
                x[2 * ri[rr]] -= lr * xr - li * xi;
                x[2 * ri[rr] + 1] -= lr * xi + li * xr;
            }
        }
    }
    if (nrest == 0) return;
    /* y = L21[:, kf:] x_seg, rows split over the threads when the block is large */
#pragma omp parallel for schedule(static) if (nrest * (w - kf) > 262144)
    for (int64_t r0 = 0; r0 < nrest; r0 += 256) {
        const int64_t r1 = r0 + 256 < nrest ? r0 + 256 : nrest;
        for (int64_t r = r0; r < r1; r++) { y[r * vs] = 0.0; if (vs == 2) y[r * 2 + 1] = 0.0; }
        for (int64_t cc = kf; cc < w; cc++) {
            const double *below = S->v + ((size_t)cc * ld + w) * vs;
            if (vs == 1) {
                const double xc = x[ri[cc]];
                if (xc == 0.0) continue;
                for (int64_t r = r0; r < r1; r++) y[r] += below[r] * xc;
            } else {
                const double xr = x[2 * ri[cc]], xi = x[2 * ri[cc] + 1];
                if (xr == 0.0 && xi == 0.0) continue;
                for (int64_t r = r0; r < r1; r++) {
                    const double lr = below[2 * r], li = below[2 * r + 1];
                    y[2 * r] += lr * xr - li * xi;
                    y[2 * r + 1] += lr * xi + li * xr;
                }
            }
        }
        for (int64_t r = r0; r < r1; r++) {
            x[ri[w + r] * vs] -= y[r * vs];
            if (vs == 2) x[ri[w + r] * 2 + 1] -= y[r * 2 + 1];
        }
    }
}

static inline double mag(const double *x, int64_t i, int vs)
{
    return vs == 1 ? fabs(x[i]) : hypot(x[2 * i], x[2 * i + 1]);
}

static int grow_u(PLU *F, int64_t need)
{
    if (F->nU + need <= F->capU) return 0;
    int64_t cap = F->capU * 2 > F->nU + need ? F->capU * 2 : F->nU + need;
    int64_t *ui = realloc(F->Ui, (size_t)cap * sizeof(int64_t));
    if (!ui) return -1;
    F->Ui = ui;
    double *ux = realloc(F->Ux, (size_t)cap * F->vs * sizeof(double));
    if (!ux) return -1;
    F->Ux = ux; F->capU = cap;
    return 0;
}

static int sn_new(PLU *F, int64_t j, int64_t nr, int64_t cap)
{
    if (F->nsn == F->capsn) {
        const int64_t c = F->capsn * 2 + 16;
        PSn *sn = realloc(F->sn, (size_t)c * sizeof(PSn));
        if (!sn) return -1;
        F->sn = sn; F->capsn = c;
    }
    PSn *S = &F->sn[F->nsn++];
    memset(S, 0, sizeof(*S));
    S->c0 = j; S->w = 0; S->nr = nr; S->ld = nr > 0 ? nr : 1; S->cap = cap;
    S->ri = malloc((size_t)S->ld * sizeof(int64_t));
    S->v = calloc((size_t)S->ld * cap * F->vs, sizeof(double));
    return S->ri && S->v ? 0 : -1;
}

static void sn_close(PSn *S, int64_t *bpos)
{
    const int64_t nrest = S->nr - S->w;
    S->srch = malloc((size_t)(nrest > 0 ? nrest : 1) * sizeof(int64_t));
    memcpy(S->srch, S->ri + S->w, (size_t)nrest * sizeof(int64_t));
    S->nsrch = nrest;
    for (int64_t r = 0; r < S->nr; r++) bpos[S->ri[r]] = -1;
}

/* the factorization of B (CSC, the analysed order, equilibrated) */
static int factor(PLU *F, const int64_t *Bp, const int64_t *Bi, const double *Bx, int64_t relax, double pivot_rel, int64_t panel,
                  char *err, size_t errlen)
{
    const int64_t n = F->n;
    const int vs = F->vs;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const int nth = omp_get_max_threads();
    Ws *ws = calloc((size_t)nth, sizeof(Ws));
    Bw b;
    memset(&b, 0, sizeof(b));
    b.x = calloc(nn * vs, sizeof(double));
    b.mark = calloc(nn, sizeof(int64_t)); b.smark = calloc(nn, sizeof(int64_t)); b.seg0 = malloc(nn * sizeof(int64_t));
    b.stack = malloc(nn * sizeof(int64_t)); b.pos = malloc(nn * sizeof(int64_t)); b.topo = malloc(nn * sizeof(int64_t));
    b.pat = malloc(nn * sizeof(int64_t)); b.cand = malloc(nn * sizeof(int64_t)); b.miss = malloc(nn * sizeof(int64_t));
    b.bpos = malloc(nn * sizeof(int64_t));
    b.capy = n * vs + 1; b.y = malloc((size_t)b.capy * sizeof(double));
    int rc = (ws && b.x && b.mark && b.smark && b.seg0 && b.stack && b.pos && b.topo && b.pat && b.cand && b.miss && b.bpos && b.y) ? 0 : ORC_NOMEM;
    for (int t = 0; t < nth && rc == 0; t++) if (ws_init(&ws[t], n, vs)) rc = ORC_NOMEM;
    for (int64_t i = 0; i < n; i++) { F->pinv[i] = -1; b.bpos[i] = -1; }
    const int64_t nchunk_max = (panel + CW - 1) / CW;
    Chunk *ch = calloc((size_t)nchunk_max, sizeof(Chunk));
    if (!ch) rc = ORC_NOMEM;
    F->Up[0] = 0;
    for (int64_t j0 = 0; j0 < n && rc == 0; j0 += panel) {
        const int64_t j1 = j0 + panel < n ? j0 + panel : n;
        const int64_t nch = (j1 - j0 + CW - 1) / CW;
        int fail = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : fail)
        for (int64_t k = 0; k < nch; k++) {
            const int64_t c0 = j0 + k * CW, nc = c0 + CW < j1 ? CW : j1 - c0;
            if (phase_a(F, &ws[omp_get_thread_num()], Bp, Bi, Bx, c0, nc, j0, &ch[k])) fail = 1;
        }
        if (fail) { rc = ORC_NOMEM; break; }
        int64_t open = -1;
        for (int64_t j = j0; j < j1 && rc == 0; j++) {
            const Chunk *C = &ch[(j - j0) / CW];
            const int c = (int)((j - j0) % CW);
            const int64_t stamp = j + 1;
            /* the column's pattern: its rows of B, then the rows the chunk's updates left nonzero in this column (the
             * chunk's other columns touch rows this one does not have: those stay zero here and are left out) */
            int64_t np = 0;
            for (int64_t q = Bp[j]; q < Bp[j + 1]; q++)
                if (b.mark[Bi[q]] != stamp) { b.mark[Bi[q]] = stamp; b.pat[np++] = Bi[q]; }
            for (int64_t t = 0; t < C->nT; t++) {
                const int64_t i = C->T[t];
                const double *xv = C->X + (t * CW + c) * vs;
                if (xv[0] == 0.0 && (vs == 1 || xv[1] == 0.0)) continue;
                memcpy(b.x + i * vs, xv, (size_t)vs * sizeof(double));
                if (b.mark[i] != stamp) { b.mark[i] = stamp; b.pat[np++] = i; }
            }
            /* the panel's own supernodes reached from the pattern */
            int64_t top = n;
            const int64_t np0 = np;
            for (int64_t t = 0; t < np0; t++) {
                const int64_t i = b.pat[t], k = F->pinv[i];
                if (k < j0) continue;
                const int64_t s = F->col_sn[k];
                if (b.smark[s] == stamp) { if (k < b.seg0[s]) b.seg0[s] = k; continue; }
                top = dfs_in(F, &b, s, k, top, j0, open, stamp, &np);
            }
            for (int64_t t = top; t < n; t++) {
                const PSn *S = &F->sn[b.topo[t]];
                const int64_t kf = b.seg0[b.topo[t]] - S->c0;
                /* its segment's pivot rows join the pattern (U's entries), as do the rows below */
                for (int64_t rr = kf; rr < S->nr; rr++)
                    if (b.mark[S->ri[rr]] != stamp) { b.mark[S->ri[rr]] = stamp; b.pat[np++] = S->ri[rr]; }
                apply_col(S, kf, b.x, b.y, vs);
            }
            /* pivot among the non-pivotal rows, in discovery order */
            double colmax = 0.0;
            for (int64_t q = Bp[j]; q < Bp[j + 1]; q++) { const double a = mag(Bx, q, vs); if (a > colmax) colmax = a; }
            int64_t ncand = 0, piv = -1;
            double best = -1.0;
            for (int64_t t = 0; t < np; t++) {
                const int64_t i = b.pat[t];
                if (F->pinv[i] >= 0) continue;
                b.cand[ncand++] = i;
                const double a = mag(b.x, i, vs);
                if (a > best) { best = a; piv = i; }
            }
            if (piv < 0 || best <= 0.0 || best < pivot_rel * colmax) {
                cp_msg(err, errlen, "Near-zero or zero pivot %.6e at step %ld (threshold %.6e). Matrix is singular.", best > 0 ? best : 0.0,
                       (long)j, pivot_rel * colmax);
                rc = ORC_SINGULAR;
                break;
            }
            /* U(:, j): the pivotal rows of the pattern, the diagonal last */
            if (grow_u(F, np + 1)) { rc = ORC_NOMEM; break; }
            int64_t u = F->nU;
            for (int64_t t = 0; t < np; t++) {
                const int64_t i = b.pat[t], k = F->pinv[i];
                if (k < 0) continue;
                F->Ui[u] = k;
                memcpy(F->Ux + u * vs, b.x + i * vs, (size_t)vs * sizeof(double));
                u++;
            }
            F->Ui[u] = j;
            memcpy(F->Ux + u * vs, b.x + piv * vs, (size_t)vs * sizeof(double));
            u++;
            F->nU = u; F->Up[j + 1] = u;
            /* L(:, j): join the open supernode or open a new one */
            int join = 0;
            int64_t n_miss = 0;
            if (open >= 0) {
                PSn *S = &F->sn[open];
                int64_t in_S = 0;
                for (int64_t s = 0; s < ncand; s++) {
                    if (b.bpos[b.cand[s]] >= S->w) in_S++;
                    else b.miss[n_miss++] = b.cand[s];
                }
                const int64_t extra = (S->nr - S->w) - in_S;
                if (n_miss + extra <= relax) join = 1;
                if (join && n_miss > 0) {
                    /* rows appended to the open supernode, zero in its stored columns */
                    const int64_t nr = S->nr + n_miss;
                    if (nr > S->ld) {
                        const int64_t ld = nr + nr / 4 + 4;
                        int64_t *ri = realloc(S->ri, (size_t)ld * sizeof(int64_t));
                        double *v = calloc((size_t)ld * S->cap * vs, sizeof(double));
                        if (!ri || !v) { free(v); rc = ORC_NOMEM; break; }
                        S->ri = ri;
                        for (int64_t cc = 0; cc < S->w; cc++) memcpy(v + cc * ld * vs, S->v + cc * S->ld * vs, (size_t)S->nr * vs * sizeof(double));
                        free(S->v); S->v = v; S->ld = ld;
                    }
                    for (int64_t t = 0; t < n_miss; t++) { S->ri[S->nr + t] = b.miss[t]; b.bpos[b.miss[t]] = S->nr + t; }
                    S->nr = nr;
                }
            }
            if (!join) {
                if (open >= 0) sn_close(&F->sn[open], b.bpos);
                if (sn_new(F, j, ncand, 4)) { rc = ORC_NOMEM; break; }
                open = F->nsn - 1;
                PSn *S = &F->sn[open];
                S->ri[0] = piv; b.bpos[piv] = 0;
                int64_t rr = 1;
                for (int64_t t = 0; t < ncand; t++) if (b.cand[t] != piv) { S->ri[rr] = b.cand[t]; b.bpos[b.cand[t]] = rr; rr++; }
            }
            PSn *S = &F->sn[open];
            const int64_t w = S->w;
            if (w + 1 > S->cap) {
                const int64_t cap = S->cap * 2;
                double *v = realloc(S->v, (size_t)S->ld * cap * vs * sizeof(double));
                if (!v) { rc = ORC_NOMEM; break; }
                memset(v + (size_t)S->ld * S->cap * vs, 0, (size_t)S->ld * (cap - S->cap) * vs * sizeof(double));
                S->v = v; S->cap = cap;
            }
            const int64_t pp = b.bpos[piv];
            if (pp != w) {   /* the pivot row to block row w */
                const int64_t other = S->ri[w];
                S->ri[w] = piv; S->ri[pp] = other; b.bpos[piv] = w; b.bpos[other] = pp;
                for (int64_t cc = 0; cc < w; cc++) {
                    double tmp[2];
                    memcpy(tmp, S->v + (cc * S->ld + w) * vs, vs * sizeof(double));
                    memcpy(S->v + (cc * S->ld + w) * vs, S->v + (cc * S->ld + pp) * vs, vs * sizeof(double));
                    memcpy(S->v + (cc * S->ld + pp) * vs, tmp, vs * sizeof(double));
                }
            }
            double *col = S->v + (size_t)w * S->ld * vs;
            for (int64_t rr = 0; rr < w; rr++) memcpy(col + rr * vs, b.x + S->ri[rr] * vs, vs * sizeof(double));
            const double pr = b.x[piv * vs], pi = vs == 2 ? b.x[piv * vs + 1] : 0.0, den = pr * pr + pi * pi;
            col[w * vs] = 1.0;
            if (vs == 2) col[w * vs + 1] = 0.0;
            for (int64_t rr = w + 1; rr < S->nr; rr++) {
                const double *xv = b.x + S->ri[rr] * vs;
                if (vs == 1) col[rr] = xv[0] / pr;
                else {
                    col[2 * rr] = (xv[0] * pr + xv[1] * pi) / den;
                    col[2 * rr + 1] = (xv[1] * pr - xv[0] * pi) / den;
                }
            }
            S->w = w + 1;
            F->col_sn[j] = open;
            F->pinv[piv] = j; F->p[j] = piv;
            for (int64_t t = 0; t < np; t++) memset(b.x + b.pat[t] * vs, 0, (size_t)vs * sizeof(double));
        }
        if (open >= 0 && rc == 0) sn_close(&F->sn[open], b.bpos);
        for (int64_t k = 0; k < nch; k++) { free(ch[k].T); free(ch[k].X); ch[k].T = NULL; ch[k].X = NULL; }
    }
    for (int t = 0; ws && t < nth; t++) ws_free(&ws[t]);
    free(ws); free(ch);
    free(b.x); free(b.y); free(b.mark); free(b.smark); free(b.seg0); free(b.stack); free(b.pos); free(b.topo); free(b.pat);
    free(b.cand); free(b.miss); free(b.bpos);
    if (rc == ORC_NOMEM) cp_msg(err, errlen, "out of memory%.0f%ld%.0f", 0.0, 0L, 0.0);
    return rc;
}

/* factor flops c * sum_k l_k u_k (l_k: entries of L below the diagonal in column k, u_k: of U right of it in row k) */
static double count_flops(const PLU *F)
{
    const int64_t n = F->n;
    int64_t *ucnt = calloc((size_t)(n > 0 ? n : 1), sizeof(int64_t));
    if (!ucnt) return -1;
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = F->Up[j]; q < F->Up[j + 1]; q++) if (F->Ui[q] != j) ucnt[F->Ui[q]]++;
    double fl = 0;
    for (int64_t s = 0; s < F->nsn; s++) {
        const PSn *S = &F->sn[s];
        for (int64_t cc = 0; cc < S->w; cc++) fl += (double)(S->nr - cc - 1) * (double)ucnt[S->c0 + cc];
    }
    free(ucnt);
    return (F->vs == 1 ? 2.0 : 8.0) * fl;
}

/* (L U) x = t for one right-hand side: t in the row space of B (overwritten), x in the column space of B */
static void solve_one(const PLU *F, double *t, double *x)
{
    const int64_t n = F->n;
    const int vs = F->vs;
    /* L y = t in row space: column k's value sits at its pivot row ri[cc] */
    for (int64_t s = 0; s < F->nsn; s++) {
        const PSn *S = &F->sn[s];
        for (int64_t cc = 0; cc < S->w; cc++) {
            const double *col = S->v + (size_t)cc * S->ld * vs;
            const double *yc = t + S->ri[cc] * vs;
            if (vs == 1) {
                const double y0 = yc[0];
                if (y0 == 0.0) continue;
                for (int64_t rr = cc + 1; rr < S->nr; rr++) t[S->ri[rr]] -= col[rr] * y0;
            } else {
                const double yr = yc[0], yi = yc[1];
                if (yr == 0.0 && yi == 0.0) continue;
                for (int64_t rr = cc + 1; rr < S->nr; rr++) {
                    double *tt = t + S->ri[rr] * 2;
                    tt[0] -= col[2 * rr] * yr - col[2 * rr + 1] * yi;
                    tt[1] -= col[2 * rr] * yi + col[2 * rr + 1] * yr;
                }
            }
        }
    }
    /* U x = y in pivot positions (U's row indices are pivot positions, the diagonal stored last) */
    for (int64_t k = 0; k < n; k++) memcpy(x + k * vs, t + F->p[k] * vs, vs * sizeof(double));
    for (int64_t j = n - 1; j >= 0; j--) {
        const int64_t q1 = F->Up[j + 1] - 1;
        double *xj = x + j * vs;
        if (vs == 1) {
            xj[0] /= F->Ux[q1];
            for (int64_t q = F->Up[j]; q < q1; q++) x[F->Ui[q]] -= F->Ux[q] * xj[0];
        } else {
            const double dr = F->Ux[2 * q1], di = F->Ux[2 * q1 + 1], den = dr * dr + di * di;
            const double a = xj[0], bb = xj[1];
            xj[0] = (a * dr + bb * di) / den;
            xj[1] = (bb * dr - a * di) / den;
            for (int64_t q = F->Up[j]; q < q1; q++) {
                double *xk = x + F->Ui[q] * 2;
                xk[0] -= F->Ux[2 * q] * xj[0] - F->Ux[2 * q + 1] * xj[1];
                xk[1] -= F->Ux[2 * q] * xj[1] + F->Ux[2 * q + 1] * xj[0];
            }
        }
    }
}

static void plu_free(PLU *F)
{
    for (int64_t s = 0; s < F->nsn; s++) { free(F->sn[s].ri); free(F->sn[s].v); free(F->sn[s].srch); }
    free(F->sn); free(F->col_sn); free(F->pinv); free(F->p); free(F->Up); free(F->Ui); free(F->Ux);
}

int cpb_factor_solve(int64_t n, const int64_t *Ap, const int64_t *Ai, const double *Ax, int cplx, const int64_t *row_perm,
                     const int64_t *col_perm, int64_t relax_rows, double pivot_rel, int64_t panel, int nthreads, double *b,
                     int64_t nrhs, double *out, char *err, size_t errlen)
{
    if (n < 1 || panel < 1 || nrhs < 0 || nthreads < 1) return ORC_BADARG;
    omp_set_num_threads(nthreads);
    const int vs = cplx ? 2 : 1;
    const size_t nn = (size_t)n, nnz = (size_t)Ap[n];
    int64_t *Bp = malloc((nn + 1) * sizeof(int64_t)), *Bi = malloc((nnz > 0 ? nnz : 1) * sizeof(int64_t));
    double *Bx = malloc((nnz > 0 ? nnz : 1) * vs * sizeof(double)), *Dr = malloc(nn * sizeof(double)), *Dc = malloc(nn * sizeof(double));
    PLU F;
    memset(&F, 0, sizeof(F));
    F.n = n; F.vs = vs;
    F.col_sn = malloc(nn * sizeof(int64_t)); F.pinv = malloc(nn * sizeof(int64_t)); F.p = malloc(nn * sizeof(int64_t));
    F.capU = 4 * (int64_t)nnz + n + 16;
    F.Up = malloc((nn + 1) * sizeof(int64_t)); F.Ui = malloc((size_t)F.capU * sizeof(int64_t)); F.Ux = malloc((size_t)F.capU * vs * sizeof(double));
    int rc = (Bp && Bi && Bx && Dr && Dc && F.col_sn && F.pinv && F.p && F.Up && F.Ui && F.Ux) ? ORC_OK : ORC_NOMEM;
    double t0 = omp_get_wtime(), t1 = t0, t2 = t0, t3 = t0;
    int iters = 0;
    if (rc == ORC_OK) rc = orc_permute(n, Ap, Ai, Ax, cplx, row_perm, col_perm, Bp, Bi, Bx);
    if (rc == ORC_OK) rc = orc_ruiz(n, Bp, Bi, Bx, cplx, 1, 20, 1e-2, Dr, Dc, &iters, err, errlen);
    t1 = omp_get_wtime();
    if (rc == ORC_OK) rc = factor(&F, Bp, Bi, Bx, relax_rows, pivot_rel, panel, err, errlen);
    t2 = omp_get_wtime();
    if (rc == ORC_OK && nrhs > 0) {
        /* x = Dc q (LU)^-1 p^T Dr b per right-hand side, the right-hand sides over the threads */
        int fail = 0;
#pragma omp parallel reduction(| : fail)
        {
            double *w = malloc(nn * vs * sizeof(double)), *t = malloc(nn * vs * sizeof(double));
            if (!w || !t) fail = 1;
#pragma omp for schedule(dynamic, 1)
            for (int64_t r = 0; r < nrhs; r++) {
                if (!w || !t) continue;
                double *br = b + (size_t)r * nn * vs;
                /* t = Dr b in B's row order (row k of B is row row_perm[k] of A) */
                for (int64_t k = 0; k < n; k++)
                    for (int c = 0; c < vs; c++) t[k * vs + c] = Dr[k] * br[row_perm[k] * vs + c];
                solve_one(&F, t, w);
                /* x[col_perm[k]] = Dc[k] w[k] */
                for (int64_t k = 0; k < n; k++)
                    for (int c = 0; c < vs; c++) br[col_perm[k] * vs + c] = Dc[k] * w[k * vs + c];
            }
            free(w); free(t);
        }
        if (fail) rc = ORC_NOMEM;
    }
    t3 = omp_get_wtime();
    if (out) {
        out[0] = t1 - t0; out[1] = t2 - t1; out[2] = t3 - t2;
        out[3] = rc == ORC_OK ? count_flops(&F) : 0.0;
        out[4] = (double)F.nsn;
        out[5] = (double)omp_get_max_threads();
    }
    free(Bp); free(Bi); free(Bx); free(Dr); free(Dc);
    plu_free(&F);
    return rc;
}
