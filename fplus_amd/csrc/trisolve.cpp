// trisolve.cpp -- fplus_tri_solve and the solve plans on the device, for CSC factors held by the caller
// (reference src/solve.c:62-74, src/solve_plan.c:64-207 and :661-719).
//
// A plan is built from the patterns alone.  For each of the four sweeps it lists, for every index, the entries that
// update it in the order of the plain column sweeps (reference src/solve.c:15-58):
//   L y = b     row i: L(i, j), j < i, columns ascending (entries on or above the diagonal of L do not feed the sweep;
//               entries above it update x after the sweep, as the column sweep does: the "upper" lists);
//   U x = y     row i: U(i, j), j > i, columns descending; the divisor is the first U(j, j) of column j;
//   U^T y = b   column j: U(i, j), i < j, in storage order; the divisor is the last U(j, j) of column j;
//   L^T x = y   column j: L(i, j), i > j, in storage order;
// plus the groups and levels of each sweep (trisolve_dev.hip).  The lists, with the position of every entry's value
// in Lx or Ux, live on the device; a solve uploads Lx, Ux and x, gathers the values into list order and runs the
// sweeps.  The results equal the plain sweeps' bit for bit, for every trans and for the plan and plan-free calls alike.
#include <algorithm>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "device_api.h"
#include "host_common.h"
#include "prof.h"
#include "trisolve_dev.h"

using namespace fpa;

namespace {

struct Lists {
    std::vector<int64_t> tp;     // index k: terms tp[k] .. tp[k+1]-1
    std::vector<int32_t> tidx;   // the index each term reads
    std::vector<int64_t> map;    // the term's value: position in Lx or Ux
};

struct SweepDev {
    int dir = 1;
    int64_t nterms = 0;
    std::vector<int32_t> lev;       // level offsets into the group list
    std::vector<double> lev_terms, lev_idx;
    DevBuf<int2> d_grp;
    DevBuf<int32_t> d_glist, d_nfirst, d_tidx;
    DevBuf<int64_t> d_tp, d_map;
};

// groups (runs of consecutive indices in sweep order, an index joining when it reads the one before it) and levels
void build_sweep(int64_t n, const Lists& Ls, int dir, SweepDev& S, hipStream_t s)
{
    S.dir = dir;
    S.nterms = (int64_t)Ls.tidx.size();
    std::vector<int32_t> gid(n > 0 ? n : 1);
    std::vector<int2> grp;
    for (int64_t t = 0; t < n; t++) {
        const int64_t k = dir > 0 ? t : n - 1 - t, prev = dir > 0 ? k - 1 : k + 1;
        bool join = false;
        if (t > 0 && grp.back().y < TS_GROUP)
            for (int64_t q = Ls.tp[k]; q < Ls.tp[k + 1] && !join; q++) join = Ls.tidx[q] == prev;
        if (join) {
            grp.back().y++;
            if (dir < 0) grp.back().x = (int32_t)k;
        } else {
            grp.push_back(make_int2((int32_t)k, 1));
        }
        gid[k] = (int32_t)grp.size() - 1;
    }
    const int32_t ng = (int32_t)grp.size();
    std::vector<int32_t> glev(ng > 0 ? ng : 1, 0), nfirst(n > 0 ? n : 1, 0);
    int32_t maxlev = 0;
    for (int32_t g = 0; g < ng; g++) {
        int32_t L = 0;
        for (int32_t k = grp[g].x; k < grp[g].x + grp[g].y; k++) {
            int32_t nf = 0;
            bool prefix = true;
            for (int64_t q = Ls.tp[k]; q < Ls.tp[k + 1]; q++) {
                const int32_t gj = gid[Ls.tidx[q]];
                if (gj != g) {
                    L = std::max(L, glev[gj] + 1);
                    if (prefix) nf++;
                } else {
                    prefix = false;
                }
            }
            nfirst[k] = nf;
        }
        glev[g] = L;
        maxlev = std::max(maxlev, L);
    }
    const int32_t nlev = ng > 0 ? maxlev + 1 : 0;
    S.lev.assign(nlev + 1, 0);
    S.lev_terms.assign(nlev, 0.0);
    S.lev_idx.assign(nlev, 0.0);
    for (int32_t g = 0; g < ng; g++) {
        S.lev[glev[g] + 1]++;
        S.lev_terms[glev[g]] += (double)(Ls.tp[grp[g].x + grp[g].y] - Ls.tp[grp[g].x]);
        S.lev_idx[glev[g]] += grp[g].y;
    }
    for (int32_t L = 0; L < nlev; L++) S.lev[L + 1] += S.lev[L];
    std::vector<int32_t> glist(ng > 0 ? ng : 1), fill(S.lev.begin(), S.lev.end());
    for (int32_t g = 0; g < ng; g++) glist[fill[glev[g]]++] = g;
    auto up = [s](auto& d, const auto& h, size_t cnt) { d.alloc(cnt > 0 ? cnt : 1); d.upload(h.data(), cnt, s); };
    up(S.d_grp, grp, grp.size());
    up(S.d_glist, glist, (size_t)ng);
    up(S.d_nfirst, nfirst, (size_t)n);
    up(S.d_tp, Ls.tp, (size_t)n + 1);
    up(S.d_tidx, Ls.tidx, Ls.tidx.size());
    up(S.d_map, Ls.map, Ls.map.size());
    dev_sync(s);
}

// the reference plan's supernode count (src/solve_plan.c:97-118, :133-160, :162, :189-199): -1 for the generic layout
int64_t reference_nsupernodes(int64_t n, const int64_t* Lp, const int64_t* Li, const int64_t* Up, const int64_t* Ui)
{
    for (int64_t j = 0; j < n; j++)
        if (Lp[j] >= Lp[j + 1] || Li[Lp[j]] != j) return -1;
    int64_t ns = 0;
    for (int64_t j = 0; j < n; j++) {
        bool join = false;
        if (j + 1 < n) {
            const int64_t a = Lp[j], la = Lp[j + 1] - a, b = Lp[j + 1], lb = Lp[j + 2] - b;
            join = la >= 2 && lb == la - 1 && Li[a + 1] == j + 1 && std::equal(Li + a + 2, Li + a + la, Li + b + 1);
        }
        if (!join) ns++;
    }
    // the blocked layout (not the plain sweeps) also needs the rows above the diagonal of every U column distinct
    int64_t nseg = 0, odd = 0;
    for (int64_t j = 0; j < n; j++) {
        int64_t prev = -2;
        bool diag = false;
        for (int64_t q = Up[j]; q < Up[j + 1]; q++) {
            const int64_t i = Ui[q];
            if (i > j || (i == j && diag)) odd++;
            if (i == j) diag = true;
            if (i >= j) { prev = -2; continue; }
            if (i != prev + 1) nseg++;
            prev = i;
        }
    }
    for (int64_t j = 0; j < n && !odd; j++)
        for (int64_t q = Lp[j] + 1; q < Lp[j + 1]; q++)
            if (Li[q] <= j) { odd++; break; }
    const bool plain = !odd && Lp[n] + Up[n] < 4000000 && (nseg == 0 || Up[n] - n < 35 * nseg) && n <= INT32_MAX;
    if (plain) return ns;
    std::vector<int64_t> stamp(n > 0 ? n : 1, -1);
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Up[j]; q < Up[j + 1]; q++) {
            const int64_t i = Ui[q];
            if (i >= j) continue;
            if (stamp[i] == j) return -1;
            stamp[i] = j;
        }
    return ns;
}

}  // namespace

struct fplus_solve_plan {
    int64_t n = 0, nnzL = 0, nnzU = 0, nsn = -1;
    SweepDev lf, ub, ut, lt;
    std::vector<int64_t> diag_first, diag_last;   // positions of U(j, j) in Ux (-1: none)
    DevBuf<int64_t> d_diag_first, d_diag_last;
    // L entries above the diagonal (only in unusual factors): rows, their lists
    int64_t nup_rows = 0, nup_terms = 0;
    DevBuf<int32_t> d_up_rows, d_up_tidx;
    DevBuf<int64_t> d_up_tp, d_up_map;
    // device workspace, reused across the solves of one plan (solves on one plan are serialized)
    mutable std::mutex mu;
    mutable DevBuf<double> d_Lx, d_Ux, d_v1, d_v2, d_dg, d_v3, d_b, d_w, d_tmp;
};

namespace {

int plan_create(int64_t n, const int64_t* Lp, const int64_t* Li, const int64_t* Up, const int64_t* Ui, fplus_solve_plan** out,
                char* err, size_t errlen)
{
    *out = nullptr;
    if (!Lp || !Up || (n > 0 && (!Li || !Ui))) { errf(err, errlen, "null factor arrays"); return FPLUS_ERR_BADARG; }
    int rc = check_compressed(n, Lp, Li, false, err, errlen);
    if (rc == FPLUS_OK) rc = check_compressed(n, Up, Ui, false, err, errlen);
    if (rc != FPLUS_OK) return rc;
    if (n > INT32_MAX - 1) { errf(err, errlen, "n too large for the device sweeps"); return FPLUS_ERR_BADARG; }
    hipStream_t s = devctx().stream;
    auto pl = std::make_unique<fplus_solve_plan>();
    pl->n = n; pl->nnzL = Lp[n]; pl->nnzU = Up[n];
    pl->nsn = reference_nsupernodes(n, Lp, Li, Up, Ui);
    const size_t nn = (size_t)(n > 0 ? n : 1);
    // rows of L and U: counting sorts
    auto row_lists = [n, nn](const int64_t* P, const int64_t* I, bool lower, bool ascending, Lists& Ls) {
        Ls.tp.assign(n + 1, 0);
        for (int64_t j = 0; j < n; j++)
            for (int64_t q = P[j]; q < P[j + 1]; q++)
                if (lower ? I[q] > j : I[q] < j) Ls.tp[I[q] + 1]++;
        for (int64_t i = 0; i < n; i++) Ls.tp[i + 1] += Ls.tp[i];
        Ls.tidx.resize(Ls.tp[n]);
        Ls.map.resize(Ls.tp[n]);
        std::vector<int64_t> pos(Ls.tp.begin(), Ls.tp.end() - 1);
        pos.resize(nn);
        for (int64_t t = 0; t < n; t++) {
            const int64_t j = ascending ? t : n - 1 - t;
            for (int64_t q = P[j]; q < P[j + 1]; q++)
                if (lower ? I[q] > j : I[q] < j) {
                    const int64_t o = pos[I[q]]++;
                    Ls.tidx[o] = (int32_t)j;
                    Ls.map[o] = q;
                }
        }
    };
    auto col_lists = [n](const int64_t* P, const int64_t* I, bool lower, Lists& Ls) {
        Ls.tp.assign(n + 1, 0);
        for (int64_t j = 0; j < n; j++) {
            int64_t c = 0;
            for (int64_t q = P[j]; q < P[j + 1]; q++) c += lower ? I[q] > j : I[q] < j;
            Ls.tp[j + 1] = Ls.tp[j] + c;
        }
        Ls.tidx.resize(Ls.tp[n]);
        Ls.map.resize(Ls.tp[n]);
        int64_t o = 0;
        for (int64_t j = 0; j < n; j++)
            for (int64_t q = P[j]; q < P[j + 1]; q++)
                if (lower ? I[q] > j : I[q] < j) { Ls.tidx[o] = (int32_t)I[q]; Ls.map[o] = q; o++; }
    };
    {
        Lists Ls;
        row_lists(Lp, Li, true, true, Ls);
        build_sweep(n, Ls, +1, pl->lf, s);
    }
    {
        Lists Ls;
        row_lists(Up, Ui, false, false, Ls);
        build_sweep(n, Ls, -1, pl->ub, s);
    }
    {
        Lists Ls;
        col_lists(Up, Ui, false, Ls);
        build_sweep(n, Ls, +1, pl->ut, s);
    }
    {
        Lists Ls;
        col_lists(Lp, Li, true, Ls);
        build_sweep(n, Ls, -1, pl->lt, s);
    }
    pl->diag_first.assign(nn, -1);
    pl->diag_last.assign(nn, -1);
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = Up[j]; q < Up[j + 1]; q++)
            if (Ui[q] == j) {
                if (pl->diag_first[j] < 0) pl->diag_first[j] = q;
                pl->diag_last[j] = q;
            }
    {   // L above its diagonal: row i takes L(i, j), j > i, columns ascending
        std::vector<int64_t> cnt(nn, 0);
        for (int64_t j = 0; j < n; j++)
            for (int64_t q = Lp[j]; q < Lp[j + 1]; q++)
                if (Li[q] < j) cnt[Li[q]]++;
        std::vector<int32_t> rows;
        std::vector<int64_t> slot(nn, -1), tp(1, 0);
        for (int64_t i = 0; i < n; i++)
            if (cnt[i] > 0) { slot[i] = (int64_t)rows.size(); rows.push_back((int32_t)i); tp.push_back(tp.back() + cnt[i]); }
        pl->nup_rows = (int64_t)rows.size();
        pl->nup_terms = tp.back();
        if (pl->nup_rows > 0) {
            std::vector<int32_t> tidx(pl->nup_terms);
            std::vector<int64_t> map(pl->nup_terms), pos(tp.begin(), tp.end() - 1);
            for (int64_t j = 0; j < n; j++)
                for (int64_t q = Lp[j]; q < Lp[j + 1]; q++)
                    if (Li[q] < j) {
                        const int64_t o = pos[slot[Li[q]]]++;
                        tidx[o] = (int32_t)j;
                        map[o] = q;
                    }
            pl->d_up_rows.alloc(rows.size()); pl->d_up_rows.upload(rows.data(), rows.size(), s);
            pl->d_up_tp.alloc(tp.size()); pl->d_up_tp.upload(tp.data(), tp.size(), s);
            pl->d_up_tidx.alloc(tidx.size()); pl->d_up_tidx.upload(tidx.data(), tidx.size(), s);
            pl->d_up_map.alloc(map.size()); pl->d_up_map.upload(map.data(), map.size(), s);
        }
    }
    pl->d_diag_first.alloc(nn); pl->d_diag_first.upload(pl->diag_first.data(), nn, s);
    pl->d_diag_last.alloc(nn); pl->d_diag_last.upload(pl->diag_last.data(), nn, s);
    dev_sync(s);
    *out = pl.release();
    return FPLUS_OK;
}

void run_sweep(const SweepDev& S, int vs, const double* tval, const double* diag, double* w, int32_t R, int skipzero, hipStream_t s)
{
    const double vb = 8.0 * vs;
    for (size_t L = 0; L + 1 < S.lev.size(); L++) {
        const double bytes = S.lev_terms[L] * (vb + 4.0 + R * vb) + S.lev_idx[L] * (2.0 * R * vb + (diag ? vb : 0.0) + 12.0);
        k::tsweep(vs, S.lev[L + 1] - S.lev[L], S.d_glist.p + S.lev[L], S.d_grp.p, S.d_tp.p, S.d_nfirst.p, S.d_tidx.p, tval, diag,
                  w, R, S.dir, skipzero, bytes, s);
    }
}

int plan_solve(const fplus_solve_plan* pl, int64_t n, const int64_t* Lp, const void* Lx, const int64_t* Up, const void* Ux,
               void* xv, int64_t nrhs, int64_t ldx, int trans, int is_complex)
{
    if (!pl || trans < 0 || trans > 2 || n < 0 || nrhs < 0 || ldx < n || n != pl->n || !Lp || !Up || Lp[n] != pl->nnzL ||
        Up[n] != pl->nnzU)
        return FPLUS_ERR_BADARG;
    const int vs = is_complex ? 2 : 1;
    const std::vector<int64_t>& dpos = trans == 0 ? pl->diag_first : pl->diag_last;
    const double* uxh = static_cast<const double*>(Ux);
    for (int64_t j = 0; j < n; j++) {   // a zero or missing pivot fails before x is touched
        const int64_t q = dpos[j];
        if (q < 0) return FPLUS_ERR_SINGULAR;
        if (uxh[q * vs] == 0.0 && (vs == 1 || uxh[q * vs + 1] == 0.0)) return FPLUS_ERR_SINGULAR;
    }
    if (n == 0 || nrhs == 0) return FPLUS_OK;
    std::lock_guard<std::mutex> lk(pl->mu);
    hipStream_t s = devctx().stream;
    const size_t sz = sizeof(double) * vs;
    pl->d_Lx.reserve((size_t)(pl->nnzL > 0 ? pl->nnzL : 1) * vs);
    pl->d_Ux.reserve((size_t)(pl->nnzU > 0 ? pl->nnzU : 1) * vs);
    pl->d_Lx.upload(static_cast<const double*>(Lx), (size_t)pl->nnzL * vs, s);
    pl->d_Ux.upload(uxh, (size_t)pl->nnzU * vs, s);
    const int cj = trans == 2 && vs == 2;
    const SweepDev& A = trans == 0 ? pl->lf : pl->ut;   // first sweep
    const SweepDev& B = trans == 0 ? pl->ub : pl->lt;   // second sweep
    const double* srcA = trans == 0 ? pl->d_Lx.p : pl->d_Ux.p;
    const double* srcB = trans == 0 ? pl->d_Ux.p : pl->d_Lx.p;
    pl->d_v1.reserve((size_t)(A.nterms > 0 ? A.nterms : 1) * vs);
    pl->d_v2.reserve((size_t)(B.nterms > 0 ? B.nterms : 1) * vs);
    pl->d_dg.reserve((size_t)n * vs);
    k::gather_conj(vs, A.nterms, A.d_map.p, srcA, pl->d_v1.p, cj, s);
    k::gather_conj(vs, B.nterms, B.d_map.p, srcB, pl->d_v2.p, cj, s);
    k::gather_conj(vs, n, trans == 0 ? pl->d_diag_first.p : pl->d_diag_last.p, pl->d_Ux.p, pl->d_dg.p, cj, s);
    if (trans == 0 && pl->nup_rows > 0) {
        pl->d_v3.reserve((size_t)pl->nup_terms * vs);
        k::gather_conj(vs, pl->nup_terms, pl->d_up_map.p, pl->d_Lx.p, pl->d_v3.p, 0, s);
    }
    // the right-hand sides: to the device (columns of n), passes of TS_RHS
    pl->d_b.reserve((size_t)(n * nrhs) * vs);
    for (int64_t r = 0; r < nrhs; r++)
        copy_h2d(reinterpret_cast<char*>(pl->d_b.p) + (size_t)r * n * sz, static_cast<const char*>(xv) + (size_t)r * ldx * sz, (size_t)n * sz, s);
    pl->d_w.reserve((size_t)(n * TS_RHS) * vs);
    if (trans == 0 && pl->nup_rows > 0) pl->d_tmp.reserve((size_t)(pl->nup_rows * TS_RHS) * vs);
    for (int64_t r0 = 0; r0 < nrhs; r0 += TS_RHS) {
        const int32_t R = (int32_t)std::min<int64_t>(TS_RHS, nrhs - r0);
        double* br = pl->d_b.p + (size_t)(r0 * n) * vs;
        k::rhs_in(vs, n, R, br, n, pl->d_w.p, s);
        if (trans == 0) {
            run_sweep(A, vs, pl->d_v1.p, nullptr, pl->d_w.p, R, 1, s);
            if (pl->nup_rows > 0) {
                k::lupper(vs, pl->nup_rows, pl->d_up_rows.p, pl->d_up_tp.p, pl->d_up_tidx.p, pl->d_v3.p, pl->d_w.p, pl->d_tmp.p, R, s);
                k::lupper_put(vs, pl->nup_rows, pl->d_up_rows.p, pl->d_tmp.p, pl->d_w.p, R, s);
            }
            run_sweep(B, vs, pl->d_v2.p, pl->d_dg.p, pl->d_w.p, R, 1, s);
        } else {
            run_sweep(A, vs, pl->d_v1.p, pl->d_dg.p, pl->d_w.p, R, 0, s);
            run_sweep(B, vs, pl->d_v2.p, nullptr, pl->d_w.p, R, 0, s);
        }
        k::rhs_out(vs, n, R, pl->d_w.p, br, n, s);
    }
    for (int64_t r = 0; r < nrhs; r++)
        copy_d2h(static_cast<char*>(xv) + (size_t)r * ldx * sz, reinterpret_cast<const char*>(pl->d_b.p) + (size_t)r * n * sz, (size_t)n * sz, s);
    dev_sync(s);
    return FPLUS_OK;
}

}  // namespace

extern "C" {

int fplus_solve_plan_create(int64_t n, const int64_t* Lp, const int64_t* Li, const int64_t* Up, const int64_t* Ui,
                            fplus_solve_plan** plan, char* err, size_t errlen)
{
    if (!plan) { errf(err, errlen, "null output pointer"); return FPLUS_ERR_BADARG; }
    *plan = nullptr;
    try {
        return plan_create(n, Lp, Li, Up, Ui, plan, err, errlen);
    } catch (const GpuError& e) {
        errf(err, errlen, "%s", e.what());
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_tri_solve_plan(const fplus_solve_plan* plan, int64_t n, const int64_t* Lp, const int64_t* Li, const void* Lx,
                         const int64_t* Up, const int64_t* Ui, const void* Ux, void* x, int64_t nrhs, int64_t ldx, int trans,
                         int is_complex)
{
    (void)Li; (void)Ui;   // the plan holds the patterns; the sizes are checked against it
    try {
        return plan_solve(plan, n, Lp, Lx, Up, Ux, x, nrhs, ldx, trans, is_complex);
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_tri_solve(int64_t n, const int64_t* Lp, const int64_t* Li, const void* Lx, const int64_t* Up, const int64_t* Ui,
                    const void* Ux, void* x, int trans, int is_complex)
{
    if (trans < 0 || trans > 2 || n < 0) return FPLUS_ERR_BADARG;
    fplus_solve_plan* pl = nullptr;
    char err[256];
    int rc = fplus_solve_plan_create(n, Lp, Li, Up, Ui, &pl, err, sizeof err);
    if (rc != FPLUS_OK) return rc;
    rc = fplus_tri_solve_plan(pl, n, Lp, Li, Lx, Up, Ui, Ux, x, 1, n, trans, is_complex);
    fplus_solve_plan_free(pl);
    return rc;
}

int64_t fplus_solve_plan_nsupernodes(const fplus_solve_plan* plan) { return plan ? plan->nsn : -1; }

void fplus_solve_plan_free(fplus_solve_plan* plan) { delete plan; }
}
