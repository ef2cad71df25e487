// order.cpp -- the host symbolic analysis: dense-line pinning, the fill-reducing column order, the column-etree
// postorder (reference src/order.c:21-62, :240-399).
//
// SuiteSparse COLAMD/AMD are not available to this build, so the fill-reducing order is this file's own minimum-degree
// ordering on the column intersection graph (the graph of A^T A), computed on the quotient graph whose initial elements
// are the rows of the interior pattern -- the formulation COLAMD uses -- with approximate external degrees, element
// absorption and mass elimination.  Its result differs from colamd_l's; the reference's tests pin no particular order,
// only that it is a permutation and that the factorization is accurate.
#include "host_common.h"
#include "symbolic.h"
#include "fplus_amd.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace {

// Liu's column-etree algorithm on the row-merged graph (each row represented by its first column), then a
// non-recursive postorder with children visited in increasing order (reference src/order.c:21-62).
int coletree_postorder(int64_t m, int64_t n, const int64_t* Ap, const int64_t* Ai, int64_t* post)
{
    std::vector<int64_t> firstcol(m > 0 ? m : 1, n), parent(n + 1, n), anc(n > 0 ? n : 1), root(n > 0 ? n : 1);
    for (int64_t c = 0; c < n; c++)
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            const int64_t r = Ai[q];
            if (r >= 0 && r < m && c < firstcol[r]) firstcol[r] = c;
        }
    auto find = [&](int64_t i) {
        int64_t r = i;
        while (anc[r] != r) r = anc[r];
        while (anc[i] != r) { const int64_t nx = anc[i]; anc[i] = r; i = nx; }
        return r;
    };
    for (int64_t c = 0; c < n; c++) {
        anc[c] = c; root[c] = c; parent[c] = n;
        for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) {
            const int64_t r = Ai[q];
            if (r < 0 || r >= m) continue;
            const int64_t fc = firstcol[r];
            if (fc >= c) continue;
            const int64_t set = find(fc), top = root[set];
            if (top != c) { parent[top] = c; anc[set] = c; root[c] = c; }
        }
    }
    std::vector<int64_t> head(n + 1, -1), next(n > 0 ? n : 1), stack(n + 1);
    for (int64_t v = n - 1; v >= 0; v--) { next[v] = head[parent[v]]; head[parent[v]] = v; }
    int64_t k = 0, sp = 0;
    stack[0] = n;
    while (sp >= 0) {
        const int64_t v = stack[sp], c = head[v];
        if (c >= 0) { head[v] = next[c]; stack[++sp] = c; }
        else { sp--; if (v < n) post[k++] = v; }
    }
    return k == n ? FPLUS_OK : FPLUS_ERR_INTERNAL;
}

// Minimum degree on the column intersection graph of the nv x (interior rows) pattern given row-wise (rows as the
// initial elements).  Columns with more than dense_col entries are ordered last; rows with more than dense_row
// entries take no part (COLAMD's dense-line rule).  order[k] = the column eliminated k-th.
void min_degree_ata(int64_t nv, const std::vector<std::vector<int32_t>>& rowcols, const std::vector<int64_t>& colcount,
                    int64_t* order)
{
    const int64_t ne0 = (int64_t)rowcols.size();
    const double dense_row = std::max(16.0, 10.0 * std::sqrt((double)std::max<int64_t>(nv, 1)));
    const double dense_col = std::max(16.0, 10.0 * std::sqrt((double)std::max<int64_t>(ne0, 1)));
    std::vector<std::vector<int32_t>> ev;      // element -> its variables
    std::vector<std::vector<int32_t>> ve(nv);  // variable -> its elements (absorbed ones dropped lazily)
    std::vector<char> ealive, vdone(nv, 0);
    ev.reserve(ne0 + nv);
    std::vector<char> vdense(nv, 0);
    for (int64_t v = 0; v < nv; v++) vdense[v] = (double)colcount[v] > dense_col;
    for (int64_t r = 0; r < ne0; r++) {
        if ((double)rowcols[r].size() > dense_row) continue;
        std::vector<int32_t> vars;
        for (int32_t v : rowcols[r]) if (!vdense[v]) vars.push_back(v);
        if (vars.empty()) continue;
        const int32_t e = (int32_t)ev.size();
        for (int32_t v : vars) ve[v].push_back(e);
        ev.push_back(std::move(vars));
        ealive.push_back(1);
    }
    int64_t nlive = 0;
    for (int64_t v = 0; v < nv; v++) nlive += !vdense[v];
    // degree buckets: doubly linked lists
    std::vector<int64_t> deg(nv, 0), head(nv + 1, -1), nxt(nv, -1), prv(nv, -1);
    auto unlink = [&](int64_t v) {
        if (prv[v] >= 0) nxt[prv[v]] = nxt[v]; else head[deg[v]] = nxt[v];
        if (nxt[v] >= 0) prv[nxt[v]] = prv[v];
    };
    auto link = [&](int64_t v, int64_t d) {
        deg[v] = d; prv[v] = -1; nxt[v] = head[d];
        if (head[d] >= 0) prv[head[d]] = v;
        head[d] = v;
    };
    for (int64_t v = nv - 1; v >= 0; v--) {
        if (vdense[v]) continue;
        double d = 0;
        for (int32_t e : ve[v]) d += (double)ev[e].size() - 1;
        link(v, (int64_t)std::min<double>(d, (double)std::max<int64_t>(nlive - 1, 0)));
    }
    std::vector<int64_t> mark(nv, 0), wstamp, w;
    int64_t stamp = 0, k = 0, mindeg = 0;
    std::vector<int32_t> Lp;
    while (k < nlive) {
        while (mindeg <= nv && head[mindeg] < 0) mindeg++;
        const int64_t p = head[mindeg];
        unlink(p);
        vdone[p] = 1;
        order[k++] = p;
        // the new element: the union of p's elements, p removed; those elements are absorbed
        stamp++;
        Lp.clear();
        mark[p] = stamp;
        for (int32_t e : ve[p]) {
            if (!ealive[e]) continue;
            ealive[e] = 0;
            for (int32_t v : ev[e]) if (!vdone[v] && mark[v] != stamp) { mark[v] = stamp; Lp.push_back(v); }
            std::vector<int32_t>().swap(ev[e]);
        }
        std::vector<int32_t>().swap(ve[p]);
        if (Lp.empty()) continue;
        const int32_t ep = (int32_t)ev.size();
        ealive.push_back(1);
        if ((int64_t)wstamp.size() < ep + 1) { wstamp.resize(2 * (ep + 1), 0); w.resize(2 * (ep + 1), 0); }
        // |e \ Lp| for every live element next to Lp
        stamp++;
        for (int32_t v : Lp)
            for (int32_t e : ve[v]) {
                if (!ealive[e]) continue;
                if (wstamp[e] != stamp) { wstamp[e] = stamp; w[e] = (int64_t)ev[e].size(); }
                w[e]--;
            }
        // degrees; elements inside Lp absorbed; variables left with the new element only are eliminated at once
        std::vector<int32_t> keep;
        keep.reserve(Lp.size());
        std::vector<int32_t> massel;
        for (int32_t v : Lp) {
            std::vector<int32_t>& E = ve[v];
            size_t o = 0;
            for (int32_t e : E) {
                if (!ealive[e]) continue;
                if (wstamp[e] == stamp && w[e] == 0) { ealive[e] = 0; std::vector<int32_t>().swap(ev[e]); continue; }
                E[o++] = e;
            }
            E.resize(o);
            if (E.empty()) massel.push_back(v);
            else keep.push_back(v);
        }
        for (int32_t v : massel) { unlink(v); vdone[v] = 1; order[k++] = v; std::vector<int32_t>().swap(ve[v]); }
        const int64_t lp = (int64_t)keep.size(), left = nlive - k;
        for (int32_t v : keep) {
            int64_t d = lp - 1;
            for (int32_t e : ve[v]) if (ealive[e] && wstamp[e] == stamp) d += w[e]; else if (ealive[e]) d += (int64_t)ev[e].size();
            d = std::min<int64_t>(d, std::max<int64_t>(left - 1, 0));
            ve[v].push_back(ep);
            unlink(v);
            link(v, d);
            if (d < mindeg) mindeg = d;
        }
        ev.push_back(std::move(keep));
    }
    for (int64_t v = 0; v < nv; v++) if (vdense[v]) order[k++] = v;
}

}  // namespace

extern "C" {

int fplus_coletree_postorder(int64_t m, int64_t n, const int64_t* Ap, const int64_t* Ai, int64_t* post)
{
    try {
        return coletree_postorder(m, n, Ap, Ai, post);
    } catch (...) {
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_analyze(int64_t n, const int64_t* Ap, const int64_t* Ai, int pin_dense, double tau_density, int64_t tau_min_nnz,
                  const unsigned char* tau_row_mask, const unsigned char* tau_col_mask, int postorder,
                  fplus_symbolic** sym, char* err, size_t errlen)
{
    return fplus_analyze_ord(n, Ap, Ai, pin_dense, tau_density, tau_min_nnz, tau_row_mask, tau_col_mask, postorder,
                             FPLUS_ORDER_COLAMD, sym, err, errlen);
}

// dense-line pinning, the interior order, the postorder (reference src/order.c:267-399)
int fplus_analyze_ord(int64_t n, const int64_t* Ap, const int64_t* Ai, int pin_dense, double tau_density, int64_t tau_min_nnz,
                      const unsigned char* tau_row_mask, const unsigned char* tau_col_mask, int postorder, int ordering,
                      fplus_symbolic** sym_out, char* err, size_t errlen)
{
    *sym_out = nullptr;
    int rc = fpa::check_compressed(n, Ap, Ai, true, err, errlen);
    if (rc != FPLUS_OK) return rc;
    if (ordering < FPLUS_ORDER_COLAMD || ordering > FPLUS_ORDER_AUTO) {
        fpa::errf(err, errlen, "ordering must be FPLUS_ORDER_COLAMD, _AMD_ATA or _AUTO (got %d)", ordering);
        return FPLUS_ERR_BADARG;
    }
    // The reference's FPLUS_ORDER_AMD_ATA runs SuiteSparse AMD on the pattern of A^T A (reference src/order.c:335-352);
    // no AMD is in this build, so it is refused rather than silently replaced.  FPLUS_ORDER_AUTO chooses between the two
    // only for small sparse matrices (n <= 4096, nnz <= 24 n) and keeps the COLAMD-formulation order otherwise; here it
    // always keeps it.
    if (ordering == FPLUS_ORDER_AMD_ATA) {
        fpa::errf(err, errlen, "FPLUS_ORDER_AMD_ATA (AMD on the pattern of A^T A) is not in this build: use FPLUS_ORDER_COLAMD or "
                               "FPLUS_ORDER_AUTO");
        return FPLUS_ERR_BADARG;
    }
    try {
        const int64_t nnz = Ap[n];
        auto* sym = new fplus_symbolic();
        sym->n = n; sym->nnz = nnz;
        std::vector<unsigned char> trow(n, 0), tcol(n, 0);
        if (pin_dense) {
            const double thr = std::max(tau_density * (double)n, (double)tau_min_nnz);
            if (tau_row_mask) std::copy(tau_row_mask, tau_row_mask + n, trow.begin());
            else {
                std::vector<int64_t> cnt(n, 0);
                for (int64_t q = 0; q < nnz; q++) cnt[Ai[q]]++;
                for (int64_t i = 0; i < n; i++) trow[i] = (double)cnt[i] > thr;
            }
            if (tau_col_mask) std::copy(tau_col_mask, tau_col_mask + n, tcol.begin());
            else for (int64_t j = 0; j < n; j++) tcol[j] = (double)(Ap[j + 1] - Ap[j]) > thr;
        }
        sym->n_tau_rows = sym->n_tau_cols = 0;
        for (int64_t i = 0; i < n; i++) { sym->n_tau_rows += trow[i] != 0; sym->n_tau_cols += tcol[i] != 0; }
        // rows: interior ascending, then dense ascending
        sym->row_perm.reserve(n);
        for (int64_t i = 0; i < n; i++) if (!trow[i]) sym->row_perm.push_back(i);
        for (int64_t i = 0; i < n; i++) if (trow[i]) sym->row_perm.push_back(i);
        // interior pattern: columns renumbered, rows renumbered ascending
        std::vector<int64_t> cmap(n, -1), rmap(n, -1), icol;
        int64_t ni = 0, mi = 0;
        for (int64_t j = 0; j < n; j++) if (!tcol[j]) { cmap[j] = ni++; icol.push_back(j); }
        for (int64_t i = 0; i < n; i++) if (!trow[i]) rmap[i] = mi++;
        std::vector<std::vector<int32_t>> rowcols(mi);
        std::vector<int64_t> colcount(ni, 0);
        for (int64_t j = 0; j < n; j++) {
            if (tcol[j]) continue;
            for (int64_t q = Ap[j]; q < Ap[j + 1]; q++)
                if (rmap[Ai[q]] >= 0) { rowcols[rmap[Ai[q]]].push_back((int32_t)cmap[j]); colcount[cmap[j]]++; }
        }
        std::vector<int64_t> ord(ni);
        min_degree_ata(ni, rowcols, colcount, ord.data());
        std::vector<std::vector<int32_t>>().swap(rowcols);
        sym->col_perm.reserve(n);
        if (postorder) {
            std::vector<int64_t> Bp(ni + 1, 0), Bi, post(ni > 0 ? ni : 1);
            for (int64_t jj = 0; jj < ni; jj++) {
                const int64_t c = icol[ord[jj]];
                for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) if (rmap[Ai[q]] >= 0) Bi.push_back(rmap[Ai[q]]);
                Bp[jj + 1] = (int64_t)Bi.size();
            }
            rc = coletree_postorder(mi, ni, Bp.data(), Bi.data(), post.data());
            if (rc != FPLUS_OK) { delete sym; fpa::errf(err, errlen, "column etree postorder failed"); return rc; }
            for (int64_t jj = 0; jj < ni; jj++) sym->col_perm.push_back(icol[ord[post[jj]]]);
        } else {
            for (int64_t jj = 0; jj < ni; jj++) sym->col_perm.push_back(icol[ord[jj]]);
        }
        for (int64_t j = 0; j < n; j++) if (tcol[j]) sym->col_perm.push_back(j);
        *sym_out = sym;
        return FPLUS_OK;
    } catch (const std::bad_alloc&) {
        fpa::errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

void fplus_symbolic_free(fplus_symbolic* sym) { delete sym; }
int64_t fplus_symbolic_n(const fplus_symbolic* sym) { return sym->n; }
int64_t fplus_symbolic_nnz(const fplus_symbolic* sym) { return sym->nnz; }
const int64_t* fplus_symbolic_row_perm(const fplus_symbolic* sym) { return sym->row_perm.data(); }
const int64_t* fplus_symbolic_col_perm(const fplus_symbolic* sym) { return sym->col_perm.data(); }
int64_t fplus_symbolic_n_dense_rows(const fplus_symbolic* sym) { return sym->n_tau_rows; }
int64_t fplus_symbolic_n_dense_cols(const fplus_symbolic* sym) { return sym->n_tau_cols; }

}  // extern "C"

extern "C" int fplus_amd_symbolic_from_orders(int64_t n, const int64_t* Ap, const int64_t* Ai, const int64_t* row_perm,
                                              const int64_t* col_perm, fplus_symbolic** sym_out, char* err, size_t errlen)
{
    *sym_out = nullptr;
    int rc = fpa::check_compressed(n, Ap, Ai, true, err, errlen);
    if (rc == FPLUS_OK) rc = fpa::check_perm(n, row_perm, "row_perm", err, errlen);
    if (rc == FPLUS_OK) rc = fpa::check_perm(n, col_perm, "col_perm", err, errlen);
    if (rc != FPLUS_OK) return rc;
    try {
        auto* sym = new fplus_symbolic();
        sym->n = n;
        sym->nnz = Ap[n];
        sym->row_perm.assign(row_perm, row_perm + n);
        sym->col_perm.assign(col_perm, col_perm + n);
        *sym_out = sym;
        return FPLUS_OK;
    } catch (const std::bad_alloc&) {
        fpa::errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}
