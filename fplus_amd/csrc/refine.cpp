// refine.cpp -- refined solves: FP64 iterative refinement with GMRES corrections preconditioned by the factorization
// (GMRES-IR), on the device, for factorizations whose updates ran in FP32, whose pivots were replaced, or that reuse the
// pivot order of an earlier factorization.
//
// Everything runs in the equilibrated, analysed system the factorization solved: Abar = Dr B Dc with B = A[row_perm, :]
// [:, col_perm] and LU = Abar[p_lu, :].  With bb = the right-hand side in that system (Dr b, rows in pivot order) and u
// the solution in it (x = Dc u in the original order), a step computes the residual r = bb - Abar u in FP64, and solves
// the correction equation M^-1 Abar d = M^-1 r (M = LU) by restarted GMRES: left-preconditioned Arnoldi with classical
// Gram-Schmidt applied twice, Givens rotations on the host per right-hand side.  Each GMRES iteration is one product
// with Abar and one pass of the triangular sweeps; the right-hand sides of a batch (up to 16) go through both together.
#include "numeric_obj.h"
#include "prof.h"
#include "refine_dev.h"
#include "solve_dev.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

using namespace fpa;

namespace {

using cplx = std::complex<double>;

// the rows of the analysed pattern, once per pattern: row pointers, each entry's column and its CSC position
void ensure_csr(PatternCache& c, hipStream_t s)
{
    std::lock_guard<std::mutex> g(c.csr_mu);
    if (c.have_csr) return;
    const int64_t n = c.n, nnz = c.nnz;
    std::vector<int64_t> Rp(n + 1, 0), Rpos(nnz > 0 ? nnz : 1);
    std::vector<int32_t> Rj(nnz > 0 ? nnz : 1);
    for (int64_t q = 0; q < nnz; q++) Rp[c.Bi[q] + 1]++;
    for (int64_t i = 0; i < n; i++) Rp[i + 1] += Rp[i];
    std::vector<int64_t> next(Rp.begin(), Rp.end() - 1);
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = c.Bp[j]; q < c.Bp[j + 1]; q++) {
            const int64_t o = next[c.Bi[q]]++;
            Rj[o] = (int32_t)j;
            Rpos[o] = q;
        }
    c.d_Rp.alloc(n + 1);
    c.d_Rp.upload(Rp.data(), n + 1, s);
    c.d_Rj.alloc(Rj.size());
    c.d_Rj.upload(Rj.data(), nnz, s);
    c.d_Rpos.alloc(Rpos.size());
    c.d_Rpos.upload(Rpos.data(), nnz, s);
    c.have_csr = true;
}

struct Refiner {
    const fplus_numeric* num;
    const PatternCache& c;
    int vs;
    int64_t n;
    int32_t R;
    int m;                       // GMRES restart
    hipStream_t s;
    double nA = 0;               // ||Abar||_1
    double spmv_bytes = 0;
    std::vector<double> hbuf;    // host copies of dot products

    Refiner(const fplus_numeric* nm, PatternCache& pc, int32_t nrhs, int restart, hipStream_t st)
        : num(nm), c(pc), vs(nm->F.vs), n(nm->n), R(nrhs), m(restart), s(st)
    {
        const size_t vec = (size_t)std::max<int64_t>(n, 1) * R * vs;
        num->ir_V.reserve(vec * (size_t)(m + 1));
        num->ir_w.reserve(vec);
        num->ir_bb.reserve(vec);
        num->ir_u.reserve(vec);
        num->ir_r.reserve(vec);
        num->ir_part.reserve((size_t)k::dots_blocks() * (m + 1) * R * vs);
        num->ir_d.reserve((size_t)(m + 1) * R * vs);
        num->ir_c.reserve((size_t)(m + 1) * R * vs);
        num->ir_sc.reserve((size_t)R);
        num->ir_norm.reserve(1);
        spmv_bytes = (double)c.nnz * (8.0 * vs + 12.0 + (double)R * 8.0 * vs) + (double)n * R * 16.0 * vs;
    }
    double* V(int j) const { return num->ir_V.p + (size_t)j * n * R * vs; }
    int64_t vstride() const { return n * R * vs; }
    // d[j, r] = <V_j, w>, j < J, onto the host as complex values
    std::vector<cplx> dots(const double* Vb, int J, const double* w)
    {
        k::dots(vs, n, R, J, Vb, vstride(), w, num->ir_part.p, num->ir_d.p, s);
        hbuf.resize((size_t)J * R * vs);
        num->ir_d.download(hbuf.data(), hbuf.size(), s);
        std::vector<cplx> out((size_t)J * R);
        for (size_t t = 0; t < out.size(); t++) out[t] = vs == 2 ? cplx(hbuf[2 * t], hbuf[2 * t + 1]) : cplx(hbuf[t], 0.0);
        return out;
    }
    std::vector<double> norms(const double* w)
    {
        std::vector<cplx> d = dots(w, 1, w);
        std::vector<double> out(R);
        for (int r = 0; r < R; r++) out[r] = std::sqrt(std::max(0.0, d[r].real()));
        return out;
    }
    void upload_coef(const std::vector<cplx>& cf, int J)
    {
        hbuf.resize((size_t)J * R * vs);
        for (size_t t = 0; t < (size_t)J * R; t++) {
            if (vs == 2) { hbuf[2 * t] = cf[t].real(); hbuf[2 * t + 1] = cf[t].imag(); }
            else hbuf[t] = cf[t].real();
        }
        num->ir_c.upload(hbuf.data(), hbuf.size(), s);
    }
    // out = M^-1 Abar v (x-space to x-space)
    void apply(const double* v, double* out)
    {
        k::spmv_eq(vs, n, R, num->d_plu.p, c.d_Rp.p, c.d_Rj.p, c.d_Rpos.p, num->d_Abar.p, v, nullptr, out, spmv_bytes, s);
        device_sweeps(num, out, R, 0, s);
    }

    // one batch: b (device, column-major, leading dimension ldb) overwritten with x
    void run(double* b, int64_t ldb, const fplus_amd_refine_options& o, fplus_amd_refine_info& info)
    {
        double* bb = num->ir_bb.p;
        double* u = num->ir_u.p;
        double* r = num->ir_r.p;
        double* w = num->ir_w.p;
        const size_t vbytes = (size_t)n * R * vs * sizeof(double);
        k::gather_scale(vs, n, R, num->d_p.p, num->d_Dr.p, b, ldb, bb, s);
        hip_check(hipMemcpyAsync(u, bb, vbytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
        device_sweeps(num, u, R, 0, s);
        const std::vector<double> nb = norms(bb);
        std::vector<bool> done(R, false);
        for (int outer = 0;; outer++) {
            k::spmv_eq(vs, n, R, num->d_plu.p, c.d_Rp.p, c.d_Rj.p, c.d_Rpos.p, num->d_Abar.p, u, bb, r, spmv_bytes, s);
            const std::vector<double> nr = norms(r), nu = norms(u);
            double worst = 0;
            for (int q = 0; q < R; q++) {
                const double den = nA * nu[q] + nb[q];
                const double be = den > 0 ? nr[q] / den : 0.0;
                done[q] = be <= o.tol;
                worst = std::max(worst, be);
            }
            if (outer == 0) info.bwd_first = std::max(info.bwd_first, worst);
            if (o.verbose >= 1) printf("[FPLUS] refinement step %d: largest equilibrated backward error %.3e\n", outer, worst);
            const bool all = std::all_of(done.begin(), done.end(), [](bool d) { return d; });
            if (all || outer >= o.max_outer) {
                info.bwd_last = std::max(info.bwd_last, worst);
                if (!all) info.converged = 0;
                break;
            }
            info.outer++;
            // GMRES on M^-1 Abar d = M^-1 r, from d = 0
            hip_check(hipMemcpyAsync(V(0), r, vbytes, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
            device_sweeps(num, V(0), R, 0, s);
            const std::vector<double> beta = norms(V(0));
            std::vector<double> sc(R);
            for (int q = 0; q < R; q++) sc[q] = beta[q] > 0 ? 1.0 / beta[q] : 0.0;
            num->ir_sc.upload(sc.data(), R, s);
            k::scale_cols(vs, n, R, V(0), num->ir_sc.p, V(0), s);
            // per right-hand side: the Hessenberg columns (rotated), the rotations, the rotated right-hand side g
            std::vector<std::vector<cplx>> H(R, std::vector<cplx>((size_t)(m + 1) * m, cplx(0)));
            std::vector<std::vector<double>> cs(R, std::vector<double>(m, 1.0));
            std::vector<std::vector<cplx>> sn(R, std::vector<cplx>(m, cplx(0)));
            std::vector<std::vector<cplx>> g(R, std::vector<cplx>(m + 1, cplx(0)));
            std::vector<bool> stop(R, false);
            std::vector<int> Jq(R, 0);   // the iterations each right-hand side took (its Hessenberg columns)
            for (int q = 0; q < R; q++) {
                g[q][0] = beta[q];
                stop[q] = done[q] || beta[q] == 0.0;
            }
            int J = 0;
            for (int j = 0; j < m; j++) {
                apply(V(j), w);
                info.inner++;
                // classical Gram-Schmidt, twice
                std::vector<cplx> h = dots(V(0), j + 1, w);
                k::combine(vs, n, R, j + 1, V(0), vstride(), num->ir_d.p, -1.0, w, s);
                const std::vector<cplx> h2 = dots(V(0), j + 1, w);
                k::combine(vs, n, R, j + 1, V(0), vstride(), num->ir_d.p, -1.0, w, s);
                for (size_t t = 0; t < h.size(); t++) h[t] += h2[t];
                const std::vector<double> hn = norms(w);
                for (int q = 0; q < R; q++) sc[q] = hn[q] > 0 ? 1.0 / hn[q] : 0.0;
                num->ir_sc.upload(sc.data(), R, s);
                k::scale_cols(vs, n, R, w, num->ir_sc.p, V(j + 1), s);
                J = j + 1;
                bool all_stop = true;
                for (int q = 0; q < R; q++) {
                    if (stop[q]) continue;
                    std::vector<cplx>& Hq = H[q];
                    auto at = [&](int i, int jj) -> cplx& { return Hq[(size_t)jj * (m + 1) + i]; };
                    Jq[q] = j + 1;
                    for (int i = 0; i <= j; i++) at(i, j) = h[(size_t)i * R + q];
                    at(j + 1, j) = hn[q];
                    for (int i = 0; i < j; i++) {   // the earlier rotations
                        const cplx a = at(i, j), bq = at(i + 1, j);
                        at(i, j) = cs[q][i] * a + sn[q][i] * bq;
                        at(i + 1, j) = -std::conj(sn[q][i]) * a + cs[q][i] * bq;
                    }
                    const cplx a = at(j, j);
                    const double bq = hn[q], t = std::sqrt(std::norm(a) + bq * bq);
                    if (t == 0.0) { cs[q][j] = 1.0; sn[q][j] = 0.0; }
                    else if (std::abs(a) == 0.0) { cs[q][j] = 0.0; sn[q][j] = 1.0; }
                    else { cs[q][j] = std::abs(a) / t; sn[q][j] = a / std::abs(a) * bq / t; }
                    at(j, j) = cs[q][j] * a + sn[q][j] * bq;
                    at(j + 1, j) = 0.0;
                    g[q][j + 1] = -std::conj(sn[q][j]) * g[q][j];
                    g[q][j] = cs[q][j] * g[q][j];
                    if (std::abs(g[q][j + 1]) <= o.inner_tol * beta[q] || hn[q] == 0.0) stop[q] = true;
                    else all_stop = false;
                }
                if (all_stop) break;
            }
            // the corrections y (upper triangular solve per right-hand side), then u += V y
            std::vector<cplx> y((size_t)J * R, cplx(0));
            for (int q = 0; q < R; q++) {
                if (done[q] || beta[q] == 0.0) continue;
                const std::vector<cplx>& Hq = H[q];
                auto at = [&](int i, int jj) { return Hq[(size_t)jj * (m + 1) + i]; };
                for (int i = Jq[q] - 1; i >= 0; i--) {
                    cplx acc = g[q][i];
                    for (int jj = i + 1; jj < Jq[q]; jj++) acc -= at(i, jj) * y[(size_t)jj * R + q];
                    const cplx d = at(i, i);
                    y[(size_t)i * R + q] = d == cplx(0) ? cplx(0) : acc / d;
                }
            }
            upload_coef(y, J);
            k::combine(vs, n, R, J, V(0), vstride(), num->ir_c.p, 1.0, u, s);
        }
        k::scatter_scale(vs, n, R, num->d_q.p, num->d_Dc.p, u, b, ldb, s);
    }
};

}  // namespace

extern "C" void fplus_amd_default_refine_options(fplus_amd_refine_options* o)
{
    std::memset(o, 0, sizeof(*o));
    o->max_outer = 10;
    o->restart = 8;
    o->tol = 1e-14;
    o->inner_tol = 1e-4;
}

extern "C" int fplus_amd_solve_refined(const fplus_numeric* num, void* b, int b_on_device, int64_t nrhs, int64_t ldb,
                                       const fplus_amd_refine_options* ro, fplus_amd_refine_info* info, char* err, size_t errlen)
{
    if (num && num->row) return row_handle_error(err, errlen);
    fplus_amd_refine_options d;
    if (!ro) { fplus_amd_default_refine_options(&d); ro = &d; }
    fplus_amd_refine_info inf;
    std::memset(&inf, 0, sizeof(inf));
    inf.converged = 1;
    if (info) *info = inf;
    if (!num) { errf(err, errlen, "no factorization"); return FPLUS_ERR_BADARG; }
    const int64_t n = num->n;
    if (nrhs < 0 || ldb < n || ro->restart < 1 || ro->restart > 32 || ro->max_outer < 0 || !(ro->tol >= 0) || !(ro->inner_tol > 0)) {
        errf(err, errlen, "fplus_amd_solve_refined: bad nrhs, ldb or options");
        return FPLUS_ERR_BADARG;
    }
    if (!num->valid) { errf(err, errlen, "the last refactorization of this factorization failed: factor again"); return FPLUS_ERR_BADARG; }
    if (!num->cache || !num->d_Abar.p) { errf(err, errlen, "this factorization keeps no matrix for the residuals"); return FPLUS_ERR_BADARG; }
    if (nrhs == 0 || n == 0) return FPLUS_OK;
    const int vs = num->F.vs;
    try {
        hipStream_t s = devctx().stream;
        std::lock_guard<std::mutex> g(num->solve_mu);
        ensure_csr(*num->cache, s);
        double* db = static_cast<double*>(b);
        DevBuf<double> tmp;
        int64_t ld = ldb;
        if (!b_on_device) {   // the host's right-hand sides to the device (columns of n), x back at the end
            tmp.alloc((size_t)(n * nrhs) * vs);
            for (int64_t r = 0; r < nrhs; r++)
                tmp.upload(static_cast<const double*>(b) + (size_t)(r * ldb) * vs, (size_t)n * vs, s, (size_t)(r * n) * vs);
            db = tmp.p;
            ld = n;
        }
        const int32_t RP = k::solve_rhs_per_pass(vs);
        unsigned long long nA_bits = 0;
        for (int64_t r0 = 0; r0 < nrhs; r0 += RP) {
            const int32_t R = (int32_t)std::min<int64_t>(RP, nrhs - r0);
            Refiner rf(num, *num->cache, R, ro->restart, s);
            if (r0 == 0) {
                k::fill_u64(num->ir_norm.p, 1, 0ULL, s);
                k::norm1(vs, n, num->cache->d_Bp.p, num->d_Abar.p, num->ir_norm.p, s);
                num->ir_norm.download(&nA_bits, 1, s);
            }
            std::memcpy(&rf.nA, &nA_bits, 8);
            rf.run(db + (size_t)(r0 * ld) * vs, ld, *ro, inf);
        }
        if (!b_on_device)
            for (int64_t r = 0; r < nrhs; r++)
                tmp.download(static_cast<double*>(b) + (size_t)(r * ldb) * vs, (size_t)n * vs, s, (size_t)(r * n) * vs);
        dev_sync(s);
    } catch (const GpuError& e) {
        errf(err, errlen, "GPU: %s", e.what());
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    } catch (const std::exception& e) {
        errf(err, errlen, "%s", e.what());
        return FPLUS_ERR_INTERNAL;
    }
    if (info) *info = inf;
    return FPLUS_OK;
}

extern "C" int fplus_amd_refactor_solve(fplus_numeric* num, const void* Ax, int ax_on_device, void* b, int b_on_device, int64_t nrhs,
                                        int64_t ldb, int precision, const fplus_amd_refine_options* ro, fplus_amd_refine_info* info,
                                        int* fallback, char* err, size_t errlen)
{
    if (fallback) *fallback = 0;
    if (!num) { errf(err, errlen, "no factorization"); return FPLUS_ERR_BADARG; }
    if (num->row) return row_handle_error(err, errlen);
    const int64_t n = num->n;
    if (nrhs < 0 || ldb < n) { errf(err, errlen, "fplus_amd_refactor_solve: bad nrhs or ldb"); return FPLUS_ERR_BADARG; }
    fplus_amd_refine_options d;
    if (!ro) { fplus_amd_default_refine_options(&d); ro = &d; }
    const int vs = num->F.vs;
    // b's columns kept on the device for a second attempt
    DevBuf<double> keep;
    try {
        hipStream_t s = devctx().stream;
        keep.alloc((size_t)std::max<int64_t>(n * nrhs, 1) * vs);
        for (int64_t r = 0; r < nrhs; r++) {
            const double* col = static_cast<const double*>(b) + (size_t)(r * ldb) * vs;
            if (b_on_device)
                hip_check(hipMemcpyAsync(keep.p + (size_t)(r * n) * vs, col, (size_t)n * vs * sizeof(double), hipMemcpyDeviceToDevice, s),
                          "hipMemcpyAsync");
            else keep.upload(col, (size_t)n * vs, s, (size_t)(r * n) * vs);
        }
    } catch (const GpuError& e) {
        errf(err, errlen, "GPU: %s", e.what());
        return FPLUS_ERR_GPU;
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
    int rc = fplus_amd_refactor(num, Ax, ax_on_device, precision, 1, err, errlen);
    fplus_amd_refine_info inf;
    std::memset(&inf, 0, sizeof(inf));
    if (rc == FPLUS_OK) rc = fplus_amd_solve_refined(num, keep.p, 1, nrhs, n, ro, &inf, err, errlen);
    if (rc == FPLUS_OK && !inf.converged && precision != 0) {
        // the refinement stalled: the factorization again with FP64 updates, b solved again from the start
        if (fallback) *fallback = 1;
        try {
            hipStream_t s = devctx().stream;
            for (int64_t r = 0; r < nrhs; r++) {
                const double* col = static_cast<const double*>(b) + (size_t)(r * ldb) * vs;
                if (b_on_device)
                    hip_check(hipMemcpyAsync(keep.p + (size_t)(r * n) * vs, col, (size_t)n * vs * sizeof(double), hipMemcpyDeviceToDevice, s),
                              "hipMemcpyAsync");
                else keep.upload(col, (size_t)n * vs, s, (size_t)(r * n) * vs);
            }
        } catch (const GpuError& e) {
            errf(err, errlen, "GPU: %s", e.what());
            return FPLUS_ERR_GPU;
        }
        fplus_amd_refine_info inf2;
        std::memset(&inf2, 0, sizeof(inf2));
        rc = fplus_amd_refactor(num, Ax, ax_on_device, 0, 1, err, errlen);
        if (rc == FPLUS_OK) rc = fplus_amd_solve_refined(num, keep.p, 1, nrhs, n, ro, &inf2, err, errlen);
        inf2.outer += inf.outer;
        inf2.inner += inf.inner;
        inf2.bwd_first = inf.bwd_first;
        inf = inf2;
    }
    if (rc != FPLUS_OK) return rc;
    try {
        hipStream_t s = devctx().stream;
        for (int64_t r = 0; r < nrhs; r++) {
            double* col = static_cast<double*>(b) + (size_t)(r * ldb) * vs;
            if (b_on_device)
                hip_check(hipMemcpyAsync(col, keep.p + (size_t)(r * n) * vs, (size_t)n * vs * sizeof(double), hipMemcpyDeviceToDevice, s),
                          "hipMemcpyAsync");
            else keep.download(col, (size_t)n * vs, s, (size_t)(r * n) * vs);
        }
        dev_sync(s);
    } catch (const GpuError& e) {
        errf(err, errlen, "GPU: %s", e.what());
        return FPLUS_ERR_GPU;
    }
    if (info) *info = inf;
    return FPLUS_OK;
}

extern "C" int fplus_amd_refactor_sweep(fplus_numeric* num, const void* Ax, int64_t nshift, void* b, int64_t nrhs, int64_t ldb,
                                        int on_device, int precision, const fplus_amd_refine_options* ro,
                                        fplus_amd_refine_info* info, int* fallback, char* err, size_t errlen)
{
    if (!num) { errf(err, errlen, "no factorization"); return FPLUS_ERR_BADARG; }
    if (num->row) return row_handle_error(err, errlen);
    if (nshift < 0 || nrhs < 0 || ldb < num->n || !Ax || (nrhs > 0 && !b)) {
        errf(err, errlen, "fplus_amd_refactor_sweep: bad nshift, nrhs, ldb, Ax or b");
        return FPLUS_ERR_BADARG;
    }
    const int vs = num->F.vs;
    const int64_t nnz = num->cache->nnz;
    for (int64_t k = 0; k < nshift; k++) {
        const double* ax = static_cast<const double*>(Ax) + (size_t)(k * nnz) * vs;
        double* bk = static_cast<double*>(b) + (size_t)(k * ldb * nrhs) * vs;
        fplus_amd_refine_info ik;
        int fk = 0;
        const int rc = fplus_amd_refactor_solve(num, ax, on_device, bk, on_device, nrhs, ldb, precision, ro, &ik, &fk, err, errlen);
        if (info) info[k] = ik;
        if (fallback) fallback[k] = fk;
        if (rc != FPLUS_OK) {
            const size_t used = std::strlen(err ? err : "");
            if (err && used + 32 < errlen) std::snprintf(err + used, errlen - used, " (shift %ld)", (long)k);
            return rc;
        }
    }
    return FPLUS_OK;
}
