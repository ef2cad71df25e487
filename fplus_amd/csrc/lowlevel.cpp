// lowlevel.cpp -- the low-level entry points of the FPLUS interface on the device path: Ruiz on host arrays, the
// supernodal factorization of a given CSC matrix, the permute + Ruiz + factor pipelines (reference src/equil.c:186,
// src/factor_sn.c:609, src/pipeline.c:84-127); the row kernel's are in rowlu.cpp.
#include "host_common.h"
#include "snfactor.h"
#include "symbolic.h"

#include <algorithm>
#include <memory>
#include <new>

using namespace fpa;

namespace fpa {
// defined in numeric.cpp
int ruiz_device_entry(int64_t n, int64_t nnz, const int32_t* d_erow, const int32_t* d_ecol, double* d_x, int vs, int max_iter,
                      double tol, double* d_Dr, double* d_Dc, int verbose, int* iters, char* err, size_t errlen);
void export_factors_entry(const DevFactor& F, int is_complex, fplus_factors* f);
}  // namespace fpa

namespace {

int gpu_error(const std::exception& e, char* err, size_t errlen)
{
    errf(err, errlen, "GPU: %s", e.what());
    return FPLUS_ERR_GPU;
}

// canonical CSC: rows sorted, duplicates summed (the reference kernel sums duplicates as it scatters)
void canonicalize(int64_t n, const int64_t* Ap, const int64_t* Ai, const double* Ax, int vs, std::vector<int64_t>& Bp,
                  std::vector<int64_t>& Bi, std::vector<double>& Bx)
{
    Bp.assign(n + 1, 0);
    Bi.clear();
    Bx.clear();
    std::vector<std::pair<int64_t, int64_t>> col;
    for (int64_t j = 0; j < n; j++) {
        col.clear();
        for (int64_t q = Ap[j]; q < Ap[j + 1]; q++) col.push_back({Ai[q], q});
        std::stable_sort(col.begin(), col.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
        for (size_t t = 0; t < col.size(); t++) {
            if (t > 0 && col[t].first == col[t - 1].first) {
                for (int v = 0; v < vs; v++) Bx[Bx.size() - vs + v] += Ax[col[t].second * vs + v];
                continue;
            }
            Bi.push_back(col[t].first);
            for (int v = 0; v < vs; v++) Bx.push_back(Ax[col[t].second * vs + v]);
        }
        Bp[j + 1] = (int64_t)Bi.size();
    }
}

// factor the canonical CSC (Bp, Bi, Bx host) with optional Ruiz, no permutation; the factors exported
int factor_plain(int64_t n, const std::vector<int64_t>& Bp, const std::vector<int64_t>& Bi, const double* Bx, int vs,
                 int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose, int64_t relax_rows, double pivot_rel,
                 fplus_factors* out, double* Dr, double* Dc, char* err, size_t errlen)
{
    hipStream_t s = devctx().stream;
    const int64_t nnz = Bp[n];
    std::vector<int32_t> bi32(nnz), ecol(nnz);
    for (int64_t q = 0; q < nnz; q++) bi32[q] = (int32_t)Bi[q];
    for (int64_t j = 0; j < n; j++) for (int64_t q = Bp[j]; q < Bp[j + 1]; q++) ecol[q] = (int32_t)j;
    DevBuf<int64_t> d_Bp(n + 1);
    DevBuf<int32_t> d_Bi(nnz), d_ecol(nnz);
    DevBuf<double> d_Bx((size_t)(nnz > 0 ? nnz : 1) * vs);
    d_Bp.upload(Bp.data(), n + 1, s);
    d_Bi.upload(bi32.data(), nnz, s);
    d_ecol.upload(ecol.data(), nnz, s);
    d_Bx.upload(Bx, (size_t)nnz * vs, s);
    int rc = FPLUS_OK;
    if (equilibrate) {
        DevBuf<double> d_Dr(n > 0 ? n : 1), d_Dc(n > 0 ? n : 1);
        rc = ruiz_device_entry(n, nnz, d_Bi.p, d_ecol.p, d_Bx.p, vs, ruiz_max_iter, ruiz_tol, d_Dr.p, d_Dc.p, verbose, nullptr, err, errlen);
        if (rc != FPLUS_OK) return rc;
        d_Dr.download(Dr, n, s);
        d_Dc.download(Dc, n, s);
        dev_sync(s);
    } else if (Dr && Dc) {
        for (int64_t i = 0; i < n; i++) { Dr[i] = 1.0; Dc[i] = 1.0; }
    }
    DevFactor F;
    SnOptions so;
    so.relax_rows = relax_rows;
    so.pivot_rel = pivot_rel;
    so.panel = 256;
    so.verbose = verbose;
    rc = sn_factor_device(n, Bp.data(), Bi.data(), d_Bp.p, d_Bi.p, d_Bx.p, vs, so, F, err, errlen);
    if (rc != FPLUS_OK) return rc;
    export_factors_entry(F, vs == 2, out);
    return FPLUS_OK;
}

}  // namespace

extern "C" {

int fplus_ruiz(int64_t n, const int64_t* indptr, const int64_t* indices, void* data, int is_complex, int is_csc, int max_iter,
               double tol, int verbose, double* Dr, double* Dc, char* err, size_t errlen)
{
    int rc = check_compressed(n, indptr, indices, false, err, errlen);
    if (rc == FPLUS_OK) rc = check_finite(indptr[n], data, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const int vs = is_complex ? 2 : 1;
    const int64_t nnz = indptr[n];
    try {
        hipStream_t s = devctx().stream;
        std::vector<int32_t> erow(nnz), ecol(nnz);
        for (int64_t l = 0; l < n; l++)
            for (int64_t q = indptr[l]; q < indptr[l + 1]; q++) {
                erow[q] = (int32_t)(is_csc ? indices[q] : l);
                ecol[q] = (int32_t)(is_csc ? l : indices[q]);
            }
        DevBuf<int32_t> d_r(nnz), d_c(nnz);
        DevBuf<double> d_x((size_t)(nnz > 0 ? nnz : 1) * vs), d_Dr(n > 0 ? n : 1), d_Dc(n > 0 ? n : 1);
        d_r.upload(erow.data(), nnz, s);
        d_c.upload(ecol.data(), nnz, s);
        d_x.upload((const double*)data, (size_t)nnz * vs, s);
        rc = ruiz_device_entry(n, nnz, d_r.p, d_c.p, d_x.p, vs, max_iter, tol, d_Dr.p, d_Dc.p, verbose, nullptr, err, errlen);
        if (rc != FPLUS_OK) return rc;
        d_x.download((double*)data, (size_t)nnz * vs, s);
        d_Dr.download(Dr, n, s);
        d_Dc.download(Dc, n, s);
        dev_sync(s);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
    return FPLUS_OK;
}

int fplus_factor_sn(int64_t n, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex, int verbose, int64_t relax_rows,
                    int64_t panel, int use_gpu, double gpu_min_flops, double pivot_rel, fplus_factors* out, char* err, size_t errlen)
{
    (void)panel; (void)use_gpu; (void)gpu_min_flops;
    std::memset(out, 0, sizeof(*out));
    if (!(pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    int rc = check_compressed(n, Ap, Ai, false, err, errlen);
    if (rc == FPLUS_OK) rc = check_finite(Ap[n], Ax, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const int vs = is_complex ? 2 : 1;
    try {
        std::vector<int64_t> Bp, Bi;
        std::vector<double> Bx;
        canonicalize(n, Ap, Ai, (const double*)Ax, vs, Bp, Bi, Bx);
        return factor_plain(n, Bp, Bi, Bx.data(), vs, 0, 0, 0.0, verbose, relax_rows, pivot_rel, out, nullptr, nullptr, err, errlen);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_factor_sn_perm(int64_t n, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex, const int64_t* row_perm,
                         const int64_t* col_perm, int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                         int64_t relax_rows, int64_t panel, int use_gpu, double gpu_min_flops, double pivot_rel,
                         fplus_factors* out, double* Dr, double* Dc, char* err, size_t errlen)
{
    (void)panel; (void)use_gpu; (void)gpu_min_flops;
    std::memset(out, 0, sizeof(*out));
    int rc = check_compressed(n, Ap, Ai, true, err, errlen);
    if (rc == FPLUS_OK) rc = check_perm(n, row_perm, "row_perm", err, errlen);
    if (rc == FPLUS_OK) rc = check_perm(n, col_perm, "col_perm", err, errlen);
    if (rc == FPLUS_OK) rc = check_finite(Ap[n], Ax, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    if (!(pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    const int vs = is_complex ? 2 : 1;
    try {
        // B = A[row_perm, :][:, col_perm], canonical (src/pipeline.c:14-82)
        const int64_t nnz = Ap[n];
        std::vector<int64_t> inv(n), Bp(n + 1, 0), Bi(nnz);
        std::vector<double> Bx((size_t)nnz * vs);
        for (int64_t k2 = 0; k2 < n; k2++) inv[row_perm[k2]] = k2;
        for (int64_t k2 = 0; k2 < n; k2++) Bp[k2 + 1] = Bp[k2] + Ap[col_perm[k2] + 1] - Ap[col_perm[k2]];
        std::vector<std::pair<int64_t, int64_t>> col;
        const double* ax = (const double*)Ax;
        for (int64_t k2 = 0; k2 < n; k2++) {
            const int64_t c = col_perm[k2];
            col.clear();
            for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) col.push_back({inv[Ai[q]], q});
            std::sort(col.begin(), col.end());
            for (size_t t = 0; t < col.size(); t++) {
                Bi[Bp[k2] + t] = col[t].first;
                for (int v = 0; v < vs; v++) Bx[(Bp[k2] + t) * vs + v] = ax[col[t].second * vs + v];
            }
        }
        return factor_plain(n, Bp, Bi, Bx.data(), vs, equilibrate, ruiz_max_iter, ruiz_tol, verbose, relax_rows, pivot_rel, out, Dr, Dc,
                            err, errlen);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_factor_sn_sym(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex,
                        int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose, int64_t relax_rows, int64_t panel,
                        int use_gpu, double gpu_min_flops, double pivot_rel, fplus_factors* out, double* Dr, double* Dc,
                        char* err, size_t errlen)
{
    const int64_t n = sym->n;
    if (Ap[n] != sym->nnz) {
        std::memset(out, 0, sizeof(*out));
        errf(err, errlen, "the matrix has %ld nonzeros, the symbolic analysis %ld", (long)Ap[n], (long)sym->nnz);
        return FPLUS_ERR_BADARG;
    }
    return fplus_factor_sn_perm(n, Ap, Ai, Ax, is_complex, sym->row_perm.data(), sym->col_perm.data(), equilibrate, ruiz_max_iter,
                                ruiz_tol, verbose, relax_rows, panel, use_gpu, gpu_min_flops, pivot_rel, out, Dr, Dc, err, errlen);
}

}  // extern "C"
