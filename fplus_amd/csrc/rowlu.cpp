// rowlu.cpp -- the row kernel's entry points (reference include/fplus.h:60-62, :137-147): fplus_factor_row (CSR in,
// CSR factors out) and the pipelines fplus_factor_row_perm / _sym (orders, Ruiz, the kernel, CSC factors; reference
// src/pipeline.c:129-225).  The elimination runs on the device (rowlu_dev.hip); the host lays out the rows, grows the
// device pools when a run reports them full, and assembles the factors the kernel leaves in its pools.
#include "rowlu.h"

#include "device_api.h"
#include "fplus_amd.h"
#include "host_common.h"
#include "rowlu_dev.h"
#include "symbolic.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <memory>
#include <new>
#include <thread>
#include <type_traits>

using namespace fpa;

namespace fpa {
// defined in numeric.cpp
int ruiz_device_entry(int64_t n, int64_t nnz, const int32_t* d_erow, const int32_t* d_ecol, double* d_x, int vs, int max_iter,
                      double tol, double* d_Dr, double* d_Dc, int verbose, int* iters, char* err, size_t errlen);

void compressed_transpose(int64_t n, const int64_t* P, const int64_t* I, std::vector<int64_t>& TP, std::vector<int64_t>& TI,
                          std::vector<int64_t>& map)
{
    const int64_t nnz = P[n];
    TP.assign(n + 1, 0);
    TI.resize(nnz);
    map.resize(nnz);
    for (int64_t q = 0; q < nnz; q++) TP[I[q] + 1]++;
    for (int64_t i = 0; i < n; i++) TP[i + 1] += TP[i];
    std::vector<int64_t> next(TP.begin(), TP.end() - 1);
    for (int64_t j = 0; j < n; j++)   // lines in order: the transposed indices come out ascending
        for (int64_t q = P[j]; q < P[j + 1]; q++) {
            const int64_t d = next[I[q]]++;
            TI[d] = j;
            map[d] = q;
        }
}

namespace {

thread_local fplus_amd_row_info t_row_info;
thread_local bool t_row_info_set = false;

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T> T* cmalloc(size_t cnt)
{
    T* p = static_cast<T*>(std::malloc((cnt > 0 ? cnt : 1) * sizeof(T)));
    if (!p) throw std::bad_alloc();
    return p;
}

void transpose_one(int64_t n, int vs, int64_t*& P, int64_t*& I, void*& X)
{
    std::vector<int64_t> TP, TI, map;
    compressed_transpose(n, P, I, TP, TI, map);
    const int64_t nnz = P[n];
    int64_t* nP = cmalloc<int64_t>(n + 1);
    int64_t* nI = cmalloc<int64_t>(nnz);
    double* nX = cmalloc<double>((size_t)nnz * vs);
    const double* x = static_cast<const double*>(X);
    std::copy(TP.begin(), TP.end(), nP);
    std::copy(TI.begin(), TI.end(), nI);
    for (int64_t t = 0; t < nnz; t++)
        for (int v = 0; v < vs; v++) nX[t * vs + v] = x[map[t] * vs + v];
    std::free(P); std::free(I); std::free(X);
    P = nP; I = nI; X = nX;
}

inline double absval(const double* x, int vs) { return vs == 2 ? std::hypot(x[0], x[1]) : std::fabs(x[0]); }

// the device state of one run of the kernel
struct RowRun {
    DevBuf<int64_t> off, base, cap_end, p, q;
    DevBuf<int32_t> nnz, lead, slot, cols, log_row, log_step, cand, scratch_col;
    DevBuf<double> vals, log_val, scratch_val, rmax;
    DevBuf<int32_t> rmaxc;
    DevBuf<unsigned long long> tops;   // pool top, log top
};

}  // namespace

void factors_to_csc(fplus_factors* f)
{
    if (f->csc) return;
    const int vs = f->is_complex ? 2 : 1;
    transpose_one(f->n, vs, f->Lp, f->Li, f->Lx);
    transpose_one(f->n, vs, f->Up, f->Ui, f->Ux);
    f->csc = 1;
}

// one system's state across the device runs of its elimination: the host layout of its rows, its device buffers,
// what the last run reported
struct RowSystem {
    int64_t n = 0, nnz = 0;
    int vs = 1, strategy = 1;
    double drop_tol = 0.0, pivot_rel = 0.0;
    const int64_t* Rp = nullptr;
    const int64_t* Ri = nullptr;
    const double* d_Rx = nullptr;
    RowLuStatus* d_status = nullptr;   // the run's report (its own buffer, or a slot of a batch's array)
    // the device arrays the runs use: the system's own buffers (R), or a batch's shared arenas until a pool grows
    struct Ptrs {
        int64_t *off = nullptr, *base = nullptr, *cap_end = nullptr, *p = nullptr, *q = nullptr;
        int32_t *nnz = nullptr, *lead = nullptr, *slot = nullptr, *cand = nullptr, *scratch_col = nullptr, *cols = nullptr;
        int32_t *log_row = nullptr, *log_step = nullptr, *rmaxc = nullptr;
        double *vals = nullptr, *log_val = nullptr, *scratch_val = nullptr, *rmax = nullptr;
        unsigned long long* tops = nullptr;
    } P;
    bool in_arena = false;             // P points into a batch's arenas, laid out and uploaded by the batch
    std::vector<int64_t> h_off, h_end, h_id;
    std::vector<int32_t> h_nnz, h_lead, h_slot, h_cols;
    int64_t pool_cap = 0, log_cap = 0;
    RowRun R;
    RowLuStatus st{};
    int runs = 0, restarts = 0;
    int32_t k0 = 0, resume = 0;
    bool fresh = true;
    int64_t merges = 0;
    double merge_bytes = 0.0;

    void layout()
    {
        // the rows as given: row i at [Rp[i], Rp[i + 1]) of the pool, its block exactly as long
        nnz = Rp[n];
        h_off.assign(Rp, Rp + n);
        h_end.assign(Rp + 1, Rp + n + 1);
        h_id.resize(n);
        h_nnz.resize(n); h_lead.resize(n); h_slot.resize(n); h_cols.resize(nnz);
        for (int64_t i = 0; i < n; i++) {
            h_nnz[i] = (int32_t)(Rp[i + 1] - Rp[i]);
            h_lead[i] = h_nnz[i] ? (int32_t)Ri[Rp[i]] : RL_NONE;
            h_slot[i] = (int32_t)i;
            h_id[i] = i;
        }
        for (int64_t q = 0; q < nnz; q++) h_cols[q] = (int32_t)Ri[q];
        pool_cap = 16 * nnz + 64 * n + 1024;
        log_cap = 4 * nnz + 2 * n + 1024;
    }

    // a pool grown in place: the entries (or multipliers) so far copied into a larger buffer of the system's own
    template <class B, class T> static void grow(B& buf, T*& cur, int64_t used, size_t per, size_t cap, hipStream_t s)
    {
        B nb(cap * per);
        if (used > 0) hip_check(hipMemcpyAsync(nb.p, cur, (size_t)used * per * sizeof(T), hipMemcpyDeviceToDevice, s), "pool growth");
        buf = std::move(nb);
        cur = buf.p;
    }

    // the buffers for the next run: laid out afresh, or the pool / log grown (GpuError when memory runs out)
    void prepare(hipStream_t s)
    {
        if (fresh && in_arena) {   // the batch laid out and uploaded the first run's arrays
            fresh = false;
            in_arena = false;
            k0 = 0;
            resume = 0;
        } else if (fresh) {
            R.off.alloc(n); R.base.alloc(n); R.cap_end.alloc(n); R.p.alloc(n); R.q.alloc(n);
            R.nnz.alloc(n); R.lead.alloc(n); R.slot.alloc(n); R.cand.alloc(n);
            R.scratch_col.alloc((size_t)16 * n);
            R.scratch_val.alloc((size_t)16 * n * vs);
            R.cols.alloc(pool_cap);
            R.vals.alloc((size_t)pool_cap * vs);
            R.log_row.alloc(log_cap); R.log_step.alloc(log_cap);
            R.log_val.alloc((size_t)log_cap * vs);
            R.tops.alloc(2);
            R.rmax.alloc(n);
            R.rmaxc.alloc(n);
            R.off.upload(h_off.data(), n, s); R.base.upload(h_off.data(), n, s); R.cap_end.upload(h_end.data(), n, s);
            R.p.upload(h_id.data(), n, s); R.q.upload(h_id.data(), n, s);
            R.nnz.upload(h_nnz.data(), n, s); R.lead.upload(h_lead.data(), n, s); R.slot.upload(h_slot.data(), n, s);
            R.cols.upload(h_cols.data(), nnz, s);
            if (nnz > 0)
                hip_check(hipMemcpyAsync(R.vals.p, d_Rx, (size_t)nnz * vs * sizeof(double), hipMemcpyDeviceToDevice, s), "row values");
            const unsigned long long tops[2] = {(unsigned long long)nnz, 0ull};
            R.tops.upload(tops, 2, s);
            P.off = R.off.p; P.base = R.base.p; P.cap_end = R.cap_end.p; P.p = R.p.p; P.q = R.q.p;
            P.nnz = R.nnz.p; P.lead = R.lead.p; P.slot = R.slot.p; P.cand = R.cand.p; P.scratch_col = R.scratch_col.p;
            P.cols = R.cols.p; P.log_row = R.log_row.p; P.log_step = R.log_step.p; P.rmaxc = R.rmaxc.p;
            P.vals = R.vals.p; P.log_val = R.log_val.p; P.scratch_val = R.scratch_val.p; P.rmax = R.rmax.p; P.tops = R.tops.p;
            fresh = false;
            k0 = 0;
            resume = 0;
        } else if (st.code == RL_POOL) {
            const int64_t used = std::min(st.pool_used, pool_cap);
            pool_cap *= 2;
            grow(R.cols, P.cols, used, 1, pool_cap, s);
            grow(R.vals, P.vals, used, vs, pool_cap, s);
        } else if (st.code == RL_LOG) {
            const int64_t used = std::min(st.log_used, log_cap);
            log_cap *= 2;
            grow(R.log_row, P.log_row, used, 1, log_cap, s);
            grow(R.log_step, P.log_step, used, 1, log_cap, s);
            grow(R.log_val, P.log_val, used, vs, log_cap, s);
        }
    }

    RowLuArgs args() const
    {
        RowLuArgs a{};
        a.n = (int32_t)n; a.strategy = strategy; a.drop_tol = drop_tol; a.pivot_rel = pivot_rel;
        a.off = P.off; a.base = P.base; a.cap_end = P.cap_end;
        a.nnz = P.nnz; a.lead = P.lead; a.slot = P.slot; a.p = P.p; a.q = P.q;
        a.cols = P.cols; a.vals = P.vals; a.pool_cap = pool_cap; a.pool_top = P.tops;
        a.log_row = P.log_row; a.log_step = P.log_step; a.log_val = P.log_val; a.log_cap = log_cap; a.log_top = P.tops + 1;
        a.cand = P.cand; a.scratch_col = P.scratch_col; a.scratch_val = P.scratch_val; a.status = d_status;
        a.rmax = P.rmax; a.rmaxc = P.rmaxc;
        a.k0 = k0; a.resume = resume;
        return a;
    }

    // after a run (st holds its report): whether another run is needed, and from where
    bool next()
    {
        runs++;
        merges += st.merges;
        merge_bytes += st.merge_bytes;
        if ((st.code == RL_POOL || st.code == RL_LOG) && runs <= 64) {   // grown in place, resumed at the step it stopped
            k0 = (int32_t)st.step;
            resume = st.phase;
            return true;
        }
        if (st.code == RL_RESTART && restarts < 8) {       // a pool ran out within a step: from the start, larger
            restarts++;
            pool_cap *= 2;
            log_cap *= 2;
            fresh = true;
            return true;
        }
        return false;
    }

    // what the factors are assembled from, copied back from the device
    std::vector<int64_t> h_doff, h_p, h_q;
    std::vector<int32_t> h_dnnz, h_dslot, h_pc, h_lr, h_ls;
    std::vector<double> h_pv, h_lv;

    // the pools and logs back to the host (the per-row arrays too, unless the caller copied them: skip_rows)
    void fetch(hipStream_t s, bool skip_rows = false)
    {
        if (st.code != RL_OK || n == 0) return;
        if (!skip_rows) {
            h_doff.resize(n); h_p.resize(n); h_q.resize(n); h_dnnz.resize(n); h_dslot.resize(n);
            copy_d2h(h_doff.data(), P.off, n * sizeof(int64_t), s);
            copy_d2h(h_dnnz.data(), P.nnz, n * sizeof(int32_t), s);
            copy_d2h(h_dslot.data(), P.slot, n * sizeof(int32_t), s);
            copy_d2h(h_p.data(), P.p, n * sizeof(int64_t), s);
            copy_d2h(h_q.data(), P.q, n * sizeof(int64_t), s);
        }
        h_pc.resize(std::max<int64_t>(st.pool_used, 0));
        h_pv.resize((size_t)std::max<int64_t>(st.pool_used, 0) * vs);
        h_lr.resize(st.log_used); h_ls.resize(st.log_used); h_lv.resize((size_t)st.log_used * vs);
        copy_d2h(h_pc.data(), P.cols, h_pc.size() * sizeof(int32_t), s);
        copy_d2h(h_pv.data(), P.vals, h_pv.size() * sizeof(double), s);
        copy_d2h(h_lr.data(), P.log_row, h_lr.size() * sizeof(int32_t), s);
        copy_d2h(h_ls.data(), P.log_step, h_ls.size() * sizeof(int32_t), s);
        copy_d2h(h_lv.data(), P.log_val, h_lv.size() * sizeof(double), s);
    }

    // the outcome: an error with the reference's messages, or the CSR factors assembled from what fetch copied back
    int result(fplus_factors* out, char* err, size_t errlen)
    {
        std::memset(out, 0, sizeof(*out));
        if (st.code == RL_SINGULAR) {
            errf(err, errlen, "Near-zero or zero pivot %.6e at step %ld (threshold %.6e). Matrix is singular.", st.best, (long)st.step, st.thr);
            return FPLUS_ERR_SINGULAR;
        }
        if (st.code == RL_POOL || st.code == RL_LOG || st.code == RL_RESTART) {
            errf(err, errlen, "out of memory (the row kernel's pools)");
            return FPLUS_ERR_NOMEM;
        }
        if (st.code != RL_OK) {
            errf(err, errlen, "Pivot entry missing at step %ld after swap.", (long)st.step);
            return FPLUS_ERR_INTERNAL;
        }
        // U row i is the row at position i; L row i its logged multipliers in step order
        const std::vector<int64_t>& d_off = h_doff;
        const std::vector<int64_t>& p = h_p;
        const std::vector<int64_t>& q = h_q;
        const std::vector<int32_t>& d_nnz = h_dnnz;
        const std::vector<int32_t>& slot = h_dslot;
        const std::vector<int32_t>& pc = h_pc;
        const std::vector<int32_t>& lr = h_lr;
        const std::vector<int32_t>& ls = h_ls;
        const std::vector<double>& pv = h_pv;
        const std::vector<double>& lv = h_lv;
    std::vector<int64_t> pos(n);
        for (int64_t i = 0; i < n; i++) pos[slot[i]] = i;
        auto dropped = [&](const double* v, int64_t j, int64_t i) { return drop_tol > 0.0 && j != i && absval(v, vs) <= drop_tol; };
        try {
            out->n = n;
            out->is_complex = vs == 2;
            out->csc = 0;
            // U (src/factor_row.c:1081-1133 with upper = 1).  With column exchanges the rows hold input column ids:
            // mapped to the final positions and sorted
            std::vector<int64_t> qinv;
            if (strategy != FPLUS_PIVOT_PARTIAL) {
                qinv.resize(n);
                for (int64_t j = 0; j < n; j++) qinv[q[j]] = j;
            }
            auto ucol = [&](int64_t t) -> int64_t { return qinv.empty() ? pc[t] : qinv[pc[t]]; };
            out->Up = cmalloc<int64_t>(n + 1);
            out->Up[0] = 0;
            for (int64_t i = 0; i < n; i++) {
                const int32_t r = slot[i];
                int64_t c = 0;
                for (int64_t t = d_off[r]; t < d_off[r] + d_nnz[r]; t++)
                    if (!dropped(&pv[t * vs], ucol(t), i) && ucol(t) >= i) c++;
                out->Up[i + 1] = out->Up[i] + c;
            }
            out->Ui = cmalloc<int64_t>(out->Up[n]);
            double* Ux = cmalloc<double>((size_t)out->Up[n] * vs);
            out->Ux = Ux;
            std::vector<std::pair<int64_t, int64_t>> row;
            for (int64_t i = 0, o = 0; i < n; i++) {
                const int32_t r = slot[i];
                row.clear();
                for (int64_t t = d_off[r]; t < d_off[r] + d_nnz[r]; t++)
                    if (!dropped(&pv[t * vs], ucol(t), i) && ucol(t) >= i) row.push_back({ucol(t), t});
                if (!qinv.empty()) std::sort(row.begin(), row.end());
                for (const auto& e : row) {
                    out->Ui[o] = e.first;
                    for (int v = 0; v < vs; v++) Ux[o * vs + v] = pv[e.second * vs + v];
                    o++;
                }
            }
            // L (upper = 0): the log grouped by row, stably, so every row's multipliers stay in step order
            std::vector<int64_t> lcnt(n + 1, 0);
            for (int64_t e = 0; e < st.log_used; e++)
                if (!dropped(&lv[e * vs], ls[e], pos[lr[e]])) lcnt[pos[lr[e]] + 1]++;
            out->Lp = cmalloc<int64_t>(n + 1);
            out->Lp[0] = 0;
            for (int64_t i = 0; i < n; i++) out->Lp[i + 1] = out->Lp[i] + 1 + lcnt[i + 1];
            out->Li = cmalloc<int64_t>(out->Lp[n]);
            double* Lx = cmalloc<double>((size_t)out->Lp[n] * vs);
            out->Lx = Lx;
            std::vector<int64_t> nxt(n);
            for (int64_t i = 0; i < n; i++) {
                const int64_t o = out->Lp[i];
                out->Li[o] = i;
                Lx[o * vs] = 1.0;
                if (vs == 2) Lx[o * vs + 1] = 0.0;
                nxt[i] = o + 1;
            }
            for (int64_t e = 0; e < st.log_used; e++) {
                const int64_t i = pos[lr[e]];
                if (dropped(&lv[e * vs], ls[e], i)) continue;
                const int64_t o = nxt[i]++;
                out->Li[o] = ls[e];
                for (int v = 0; v < vs; v++) Lx[o * vs + v] = lv[e * vs + v];
            }
            out->p = cmalloc<int64_t>(n);
            out->q = cmalloc<int64_t>(n);
            std::copy(p.begin(), p.end(), out->p);
            std::copy(q.begin(), q.end(), out->q);
            out->nsupernodes = 0;
            out->n_pivots_replaced = st.nrep;
        } catch (const std::bad_alloc&) {
            fplus_factors_free(out);
            errf(err, errlen, "out of memory");
            return FPLUS_ERR_NOMEM;
        }
        return FPLUS_OK;
    }
};

void record_info(const RowSystem* const* sys, size_t cnt, double t0, double t1)
{
    fplus_amd_row_info inf{};
    for (size_t i = 0; i < cnt; i++) {
        inf.runs += sys[i]->runs;
        inf.restarts += sys[i]->restarts;
        inf.merges += sys[i]->merges;
        inf.merge_bytes += sys[i]->merge_bytes;
        inf.pool_entries += sys[i]->st.pool_used;
        inf.multipliers += sys[i]->st.log_used;
    }
    inf.t_device = t1 - t0;
    inf.t_total = now_s() - t0;
    t_row_info = inf;
    t_row_info_set = true;
}

int row_factor_device(int64_t n, const int64_t* Rp, const int64_t* Ri, const double* d_Rx, int vs, double drop_tol, int strategy,
                      double pivot_rel, int verbose, fplus_factors* out, char* err, size_t errlen)
{
    std::memset(out, 0, sizeof(*out));
    if (n >= (int64_t)RL_NONE) { errf(err, errlen, "the row kernel takes n < 2^31 - 1 (got %ld)", (long)n); return FPLUS_ERR_BADARG; }
    const char* sname = strategy == FPLUS_PIVOT_FULL ? "full" : strategy == FPLUS_PIVOT_PARTIAL ? "partial" : "rook";
    if (verbose >= 1)
        printf("[FPLUS] LU factorisation start: n=%ld, nnz(A)=%ld, %s, drop_tol=%.1e, pivot=%s (row kernel, device)\n", (long)n,
               (long)Rp[n], vs == 2 ? "complex128" : "float64", drop_tol, sname);
    hipStream_t s = devctx().stream;
    RowSystem S;
    S.n = n; S.vs = vs; S.strategy = strategy; S.drop_tol = drop_tol; S.pivot_rel = pivot_rel;
    S.Rp = Rp; S.Ri = Ri; S.d_Rx = d_Rx;
    S.layout();
    DevBuf<RowLuStatus> d_status(1);
    S.d_status = d_status.p;
    const double t0 = now_s();
    for (;;) {
        try {
            S.prepare(s);
        } catch (const GpuError&) {
            errf(err, errlen, "out of memory (device: %ld pool entries, %ld multipliers)", (long)S.pool_cap, (long)S.log_cap);
            return FPLUS_ERR_NOMEM;
        }
        if (n == 0) break;
        row_lu(vs, S.args(), s);
        d_status.download(&S.st, 1, s);
        dev_sync(s);
        if (!S.next()) break;
    }
    const double t1 = now_s();
    S.fetch(s);
    const int rc = S.result(out, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const RowSystem* one = &S;
    record_info(&one, 1, t0, t1);
    if (verbose >= 1) {
        printf("[FPLUS] LU factorisation complete: nnz(L)=%ld, nnz(U)=%ld, total nnz=%ld, time=%.3fs (device elimination %.3fs in %d "
               "runs, %d from the start; pool %ld of %ld entries, %ld multipliers)\n", (long)out->Lp[n], (long)out->Up[n],
               (long)(out->Lp[n] + out->Up[n]), now_s() - t0, t1 - t0, S.runs, S.restarts, (long)S.st.pool_used, (long)S.pool_cap,
               (long)S.st.log_used);
        if (S.st.nrep > 0)
            printf("[FPLUS] ILU: %ld zero / negligible pivots replaced by drop_tol * max|a_ij| = %.3e\n", (long)S.st.nrep,
                   drop_tol * S.st.max0);
    }
    return FPLUS_OK;
}

// independent systems, one launch per round: every system that still needs a run gets one workgroup.  The first
// run's arrays of all systems live in shared arenas (one allocation and one upload per array), their inputs packed;
// a system whose pool or log grows moves that buffer to one of its own.  The factors are assembled on host threads.
int row_factor_batch(int64_t nsys, const int64_t* n, const int64_t* const* Rp, const int64_t* const* Ri, const double* const* Rx,
                     int vs, double drop_tol, int strategy, double pivot_rel, fplus_factors* out, int* rc, char* err, size_t errlen)
{
    hipStream_t s = devctx().stream;
    std::vector<std::unique_ptr<RowSystem>> sys(nsys);
    DevBuf<RowLuStatus> d_status(nsys > 0 ? nsys : 1);
    // the arenas' offsets: per-row arrays, scratch, pools, logs, counters
    std::vector<int64_t> o_row(nsys + 1, 0), o_pool(nsys + 1, 0), o_log(nsys + 1, 0), o_nnz(nsys + 1, 0);
    for (int64_t i = 0; i < nsys; i++) {
        sys[i].reset(new RowSystem());
        RowSystem& S = *sys[i];
        S.n = n[i]; S.vs = vs; S.strategy = strategy; S.drop_tol = drop_tol; S.pivot_rel = pivot_rel;
        S.Rp = Rp[i]; S.Ri = Ri[i];
        S.d_status = d_status.p + i;
        S.layout();
        o_row[i + 1] = o_row[i] + S.n;
        o_pool[i + 1] = o_pool[i] + S.pool_cap;
        o_log[i + 1] = o_log[i] + S.log_cap;
        o_nnz[i + 1] = o_nnz[i] + S.nnz;
    }
    const int64_t NR = o_row[nsys], NP = o_pool[nsys], NL = o_log[nsys], NZ = o_nnz[nsys];
    DevBuf<int64_t> a_off, a_base, a_end, a_p, a_q;
    DevBuf<int32_t> a_nnz, a_lead, a_slot, a_cand, a_scol, a_cols, a_lrow, a_lstep, a_rmaxc, a_incols;
    DevBuf<double> a_vals, a_lval, a_sval, a_rmax, a_in;
    DevBuf<unsigned long long> a_tops;
    try {
        a_off.alloc(NR); a_base.alloc(NR); a_end.alloc(NR); a_p.alloc(NR); a_q.alloc(NR);
        a_nnz.alloc(NR); a_lead.alloc(NR); a_slot.alloc(NR); a_cand.alloc(NR); a_rmaxc.alloc(NR); a_rmax.alloc(NR);
        a_scol.alloc((size_t)16 * NR); a_sval.alloc((size_t)16 * NR * vs);
        a_cols.alloc(NP); a_vals.alloc((size_t)NP * vs);
        a_lrow.alloc(NL); a_lstep.alloc(NL); a_lval.alloc((size_t)NL * vs);
        a_tops.alloc(2 * (nsys > 0 ? nsys : 1));
        a_in.alloc((size_t)(NZ > 0 ? NZ : 1) * vs);
        a_incols.alloc(NZ > 0 ? NZ : 1);
    } catch (const GpuError&) {
        errf(err, errlen, "out of memory (device: %ld rows, %ld pool entries, %ld multipliers in the batch)", (long)NR, (long)NP, (long)NL);
        return FPLUS_ERR_NOMEM;
    }
    {   // the first run's per-row arrays and counters, packed; the inputs' columns and values through a staging area
        std::vector<int64_t> off(NR), end(NR), id(NR);
        std::vector<int32_t> nz(NR), lead(NR), slot(NR), cols(NZ > 0 ? NZ : 1);
        std::vector<double> vals((size_t)(NZ > 0 ? NZ : 1) * vs);
        std::vector<unsigned long long> tops(2 * (nsys > 0 ? nsys : 1), 0ull);
        for (int64_t i = 0; i < nsys; i++) {
            const RowSystem& S = *sys[i];
            std::copy(S.h_off.begin(), S.h_off.end(), off.begin() + o_row[i]);
            std::copy(S.h_end.begin(), S.h_end.end(), end.begin() + o_row[i]);
            std::copy(S.h_id.begin(), S.h_id.end(), id.begin() + o_row[i]);
            std::copy(S.h_nnz.begin(), S.h_nnz.end(), nz.begin() + o_row[i]);
            std::copy(S.h_lead.begin(), S.h_lead.end(), lead.begin() + o_row[i]);
            std::copy(S.h_slot.begin(), S.h_slot.end(), slot.begin() + o_row[i]);
            std::copy(S.h_cols.begin(), S.h_cols.end(), cols.begin() + o_nnz[i]);
            std::copy(Rx[i], Rx[i] + S.nnz * vs, vals.begin() + o_nnz[i] * vs);
            tops[2 * i] = (unsigned long long)S.nnz;
        }
        a_off.upload(off.data(), NR, s); a_base.upload(off.data(), NR, s); a_end.upload(end.data(), NR, s);
        a_p.upload(id.data(), NR, s); a_q.upload(id.data(), NR, s);
        a_nnz.upload(nz.data(), NR, s); a_lead.upload(lead.data(), NR, s); a_slot.upload(slot.data(), NR, s);
        a_tops.upload(tops.data(), 2 * nsys, s);
        a_incols.upload(cols.data(), NZ, s);
        a_in.upload(vals.data(), (size_t)NZ * vs, s);
        for (int64_t i = 0; i < nsys; i++) {
            const int64_t z = sys[i]->nnz;
            if (z == 0) continue;
            hip_check(hipMemcpyAsync(a_cols.p + o_pool[i], a_incols.p + o_nnz[i], z * sizeof(int32_t), hipMemcpyDeviceToDevice, s),
                      "batch columns");
            hip_check(hipMemcpyAsync(a_vals.p + o_pool[i] * vs, a_in.p + o_nnz[i] * vs, z * vs * sizeof(double), hipMemcpyDeviceToDevice, s),
                      "batch values");
        }
    }
    std::vector<int64_t> active;
    for (int64_t i = 0; i < nsys; i++) {
        RowSystem& S = *sys[i];
        auto& P = S.P;
        P.off = a_off.p + o_row[i]; P.base = a_base.p + o_row[i]; P.cap_end = a_end.p + o_row[i];
        P.p = a_p.p + o_row[i]; P.q = a_q.p + o_row[i];
        P.nnz = a_nnz.p + o_row[i]; P.lead = a_lead.p + o_row[i]; P.slot = a_slot.p + o_row[i]; P.cand = a_cand.p + o_row[i];
        P.rmaxc = a_rmaxc.p + o_row[i]; P.rmax = a_rmax.p + o_row[i];
        P.scratch_col = a_scol.p + 16 * o_row[i]; P.scratch_val = a_sval.p + 16 * o_row[i] * vs;
        P.cols = a_cols.p + o_pool[i]; P.vals = a_vals.p + o_pool[i] * vs;
        P.log_row = a_lrow.p + o_log[i]; P.log_step = a_lstep.p + o_log[i]; P.log_val = a_lval.p + o_log[i] * vs;
        P.tops = a_tops.p + 2 * i;
        S.d_Rx = a_in.p + o_nnz[i] * vs;   // a system that starts over lays out its rows again from here
        S.in_arena = true;
        if (S.n > 0) active.push_back(i);
    }
    const double t0 = now_s();
    std::vector<RowLuArgs> args;
    DevBuf<RowLuArgs> d_args(nsys > 0 ? nsys : 1);
    std::vector<RowLuStatus> st(nsys > 0 ? nsys : 1);
    while (!active.empty()) {
        args.clear();
        for (int64_t i : active) {
            try {
                sys[i]->prepare(s);
            } catch (const GpuError&) {
                errf(err, errlen, "out of memory (device: system %ld, %ld pool entries, %ld multipliers)", (long)i, (long)sys[i]->pool_cap,
                     (long)sys[i]->log_cap);
                return FPLUS_ERR_NOMEM;
            }
            args.push_back(sys[i]->args());
        }
        d_args.upload(args.data(), args.size(), s);
        row_lu_batch(vs, strategy, d_args.p, (int)args.size(), s);
        d_status.download(st.data(), nsys, s);
        dev_sync(s);
        std::vector<int64_t> again;
        for (int64_t i : active) {
            sys[i]->st = st[i];
            if (sys[i]->next()) again.push_back(i);
        }
        active.swap(again);
    }
    const double t1 = now_s();
    // the per-row arrays of every system back in one copy each (a system that restarted keeps its own), then each
    // system's pool and log, then the factors assembled on host threads
    {
        std::vector<int64_t> off(NR), p(NR), q(NR);
        std::vector<int32_t> nz(NR), slot(NR);
        a_off.download(off.data(), NR, s); a_p.download(p.data(), NR, s); a_q.download(q.data(), NR, s);
        a_nnz.download(nz.data(), NR, s); a_slot.download(slot.data(), NR, s);
        for (int64_t i = 0; i < nsys; i++) {
            RowSystem& S = *sys[i];
            const bool own = S.P.off != a_off.p + o_row[i];
            if (!own && S.st.code == RL_OK && S.n > 0) {
                S.h_doff.assign(off.begin() + o_row[i], off.begin() + o_row[i + 1]);
                S.h_p.assign(p.begin() + o_row[i], p.begin() + o_row[i + 1]);
                S.h_q.assign(q.begin() + o_row[i], q.begin() + o_row[i + 1]);
                S.h_dnnz.assign(nz.begin() + o_row[i], nz.begin() + o_row[i + 1]);
                S.h_dslot.assign(slot.begin() + o_row[i], slot.begin() + o_row[i + 1]);
            }
            S.fetch(s, !own);
        }
    }
    std::vector<int> r(nsys, FPLUS_OK);
    {
        std::atomic<int64_t> next{0};
        auto work = [&]() {
            for (int64_t i; (i = next.fetch_add(1)) < nsys;) {
                char e2[256];
                r[i] = sys[i]->result(&out[i], e2, sizeof e2);
            }
        };
        const int nt = (int)std::min<int64_t>(std::max(1, host_threads()), std::max<int64_t>(nsys, 1));
        std::vector<std::thread> th;
        for (int t = 1; t < nt; t++) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
    }
    for (int64_t i = 0; i < nsys; i++) rc[i] = r[i];
    std::vector<const RowSystem*> ptrs;
    for (auto& S : sys) ptrs.push_back(S.get());
    record_info(ptrs.data(), ptrs.size(), t0, t1);
    return FPLUS_OK;
}

}  // namespace fpa

namespace {

int gpu_error(const std::exception& e, char* err, size_t errlen)
{
    errf(err, errlen, "GPU: %s", e.what());
    return FPLUS_ERR_GPU;
}

}  // namespace

extern "C" {

int fplus_factor_row(int64_t n, const int64_t* indptr, const int64_t* indices, const void* data, double drop_tol, int is_complex,
                     int verbose, int pivot_strategy, double pivot_rel, fplus_factors* out, char* err, size_t errlen)
{
    std::memset(out, 0, sizeof(*out));
    if (pivot_strategy != FPLUS_PIVOT_FULL && pivot_strategy != FPLUS_PIVOT_PARTIAL && pivot_strategy != FPLUS_PIVOT_ROOK) {
        errf(err, errlen, "pivot_strategy must be 0 (full), 1 (partial), or 2 (rook); got %d", pivot_strategy);
        return FPLUS_ERR_BADARG;
    }
    if (!(pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    if (!(drop_tol >= 0.0)) { errf(err, errlen, "drop_tol must be >= 0 (got %g)", drop_tol); return FPLUS_ERR_BADARG; }
    int rc = check_compressed(n, indptr, indices, true, err, errlen);   // the row kernel needs sorted, unique rows
    if (rc == FPLUS_OK) rc = check_finite(indptr[n], data, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const int vs = is_complex ? 2 : 1;
    try {
        hipStream_t s = devctx().stream;
        const int64_t nnz = indptr[n];
        DevBuf<double> d_x((size_t)(nnz > 0 ? nnz : 1) * vs);
        d_x.upload(static_cast<const double*>(data), (size_t)nnz * vs, s);
        return row_factor_device(n, indptr, indices, d_x.p, vs, drop_tol, pivot_strategy, pivot_rel, verbose, out, err, errlen);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_factor_row_perm(int64_t n, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex, const int64_t* row_perm,
                          const int64_t* col_perm, int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose,
                          double drop_tol, int pivot_strategy, double pivot_rel, fplus_factors* out, double* Dr, double* Dc,
                          char* err, size_t errlen)
{
    std::memset(out, 0, sizeof(*out));
    int rc = check_compressed(n, Ap, Ai, true, err, errlen);
    if (rc == FPLUS_OK) rc = check_perm(n, row_perm, "row_perm", err, errlen);
    if (rc == FPLUS_OK) rc = check_perm(n, col_perm, "col_perm", err, errlen);
    if (rc == FPLUS_OK) rc = check_finite(Ap[n], Ax, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    if (pivot_strategy != FPLUS_PIVOT_FULL && pivot_strategy != FPLUS_PIVOT_PARTIAL && pivot_strategy != FPLUS_PIVOT_ROOK) {
        errf(err, errlen, "pivot_strategy must be 0 (full), 1 (partial), or 2 (rook); got %d", pivot_strategy);
        return FPLUS_ERR_BADARG;
    }
    if (!(pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    if (!(drop_tol >= 0.0)) { errf(err, errlen, "drop_tol must be >= 0 (got %g)", drop_tol); return FPLUS_ERR_BADARG; }
    const int vs = is_complex ? 2 : 1;
    try {
        hipStream_t s = devctx().stream;
        const int64_t nnz = Ap[n];
        // B = A[row_perm, :][:, col_perm], canonical CSC (src/pipeline.c:14-82)
        std::vector<int64_t> inv(n), Bp(n + 1, 0), Bi(nnz), src(nnz);
        for (int64_t k = 0; k < n; k++) inv[row_perm[k]] = k;
        for (int64_t k = 0; k < n; k++) Bp[k + 1] = Bp[k] + Ap[col_perm[k] + 1] - Ap[col_perm[k]];
        std::vector<std::pair<int64_t, int64_t>> col;
        for (int64_t k = 0; k < n; k++) {
            const int64_t c = col_perm[k];
            col.clear();
            for (int64_t q = Ap[c]; q < Ap[c + 1]; q++) col.push_back({inv[Ai[q]], q});
            std::sort(col.begin(), col.end());
            for (size_t t = 0; t < col.size(); t++) { Bi[Bp[k] + t] = col[t].first; src[Bp[k] + t] = col[t].second; }
        }
        // the values gathered on the device, equilibrated there (CSC route), then laid out by rows
        std::vector<int32_t> erow(nnz), ecol(nnz);
        for (int64_t k = 0; k < n; k++)
            for (int64_t q = Bp[k]; q < Bp[k + 1]; q++) { erow[q] = (int32_t)Bi[q]; ecol[q] = (int32_t)k; }
        DevBuf<double> d_A((size_t)(nnz > 0 ? nnz : 1) * vs), d_B((size_t)(nnz > 0 ? nnz : 1) * vs), d_R((size_t)(nnz > 0 ? nnz : 1) * vs);
        DevBuf<int64_t> d_map(nnz > 0 ? nnz : 1);
        DevBuf<int32_t> d_er(nnz > 0 ? nnz : 1), d_ec(nnz > 0 ? nnz : 1);
        d_A.upload(static_cast<const double*>(Ax), (size_t)nnz * vs, s);
        d_map.upload(src.data(), nnz, s);
        k::gather_vals(vs, nnz, d_map.p, d_A.p, d_B.p, s);
        d_er.upload(erow.data(), nnz, s);
        d_ec.upload(ecol.data(), nnz, s);
        if (equilibrate) {
            DevBuf<double> d_Dr(n > 0 ? n : 1), d_Dc(n > 0 ? n : 1);
            rc = ruiz_device_entry(n, nnz, d_er.p, d_ec.p, d_B.p, vs, ruiz_max_iter, ruiz_tol, d_Dr.p, d_Dc.p, verbose, nullptr, err, errlen);
            if (rc != FPLUS_OK) return rc;
            d_Dr.download(Dr, n, s);
            d_Dc.download(Dc, n, s);
        } else {
            for (int64_t i = 0; i < n; i++) { Dr[i] = 1.0; Dc[i] = 1.0; }
        }
        std::vector<int64_t> Rp, Ri, rmap;
        compressed_transpose(n, Bp.data(), Bi.data(), Rp, Ri, rmap);
        d_map.upload(rmap.data(), nnz, s);
        k::gather_vals(vs, nnz, d_map.p, d_B.p, d_R.p, s);
        rc = row_factor_device(n, Rp.data(), Ri.data(), d_R.p, vs, drop_tol, pivot_strategy, pivot_rel, verbose, out, err, errlen);
        if (rc != FPLUS_OK) return rc;
        factors_to_csc(out);
        return FPLUS_OK;
    } catch (const GpuError& e) {
        fplus_factors_free(out);
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        fplus_factors_free(out);
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_factor_row_sym(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex,
                         int equilibrate, int ruiz_max_iter, double ruiz_tol, int verbose, double drop_tol, int pivot_strategy,
                         double pivot_rel, fplus_factors* out, double* Dr, double* Dc, char* err, size_t errlen)
{
    const int64_t n = sym->n;
    if (Ap[n] != sym->nnz) {
        std::memset(out, 0, sizeof(*out));
        errf(err, errlen, "the matrix has %ld nonzeros, the symbolic analysis %ld", (long)Ap[n], (long)sym->nnz);
        return FPLUS_ERR_BADARG;
    }
    return fplus_factor_row_perm(n, Ap, Ai, Ax, is_complex, sym->row_perm.data(), sym->col_perm.data(), equilibrate, ruiz_max_iter,
                                 ruiz_tol, verbose, drop_tol, pivot_strategy, pivot_rel, out, Dr, Dc, err, errlen);
}

int fplus_amd_factor_row_batch(int64_t nsys, const int64_t* n, const int64_t* const* indptr, const int64_t* const* indices,
                               const void* const* data, int is_complex, double drop_tol, int pivot_strategy, double pivot_rel,
                               fplus_factors* out, int* rc, char* err, size_t errlen)
{
    if (nsys < 0 || (nsys > 0 && (!n || !indptr || !indices || !data || !out || !rc))) {
        errf(err, errlen, "fplus_amd_factor_row_batch: bad nsys or null arrays");
        return FPLUS_ERR_BADARG;
    }
    for (int64_t i = 0; i < nsys; i++) { std::memset(&out[i], 0, sizeof(out[i])); rc[i] = FPLUS_OK; }
    if (pivot_strategy != FPLUS_PIVOT_FULL && pivot_strategy != FPLUS_PIVOT_PARTIAL && pivot_strategy != FPLUS_PIVOT_ROOK) {
        errf(err, errlen, "pivot_strategy must be 0 (full), 1 (partial), or 2 (rook); got %d", pivot_strategy);
        return FPLUS_ERR_BADARG;
    }
    if (!(pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    if (!(drop_tol >= 0.0)) { errf(err, errlen, "drop_tol must be >= 0 (got %g)", drop_tol); return FPLUS_ERR_BADARG; }
    const int vs = is_complex ? 2 : 1;
    std::vector<int64_t> idx, bn;
    std::vector<const int64_t*> bp, bi;
    for (int64_t i = 0; i < nsys; i++) {   // each system's checks; a system that fails them keeps its code and is left out
        char e2[256];
        int r = n[i] >= (int64_t)RL_NONE ? FPLUS_ERR_BADARG : check_compressed(n[i], indptr[i], indices[i], true, e2, sizeof e2);
        if (r == FPLUS_OK) r = check_finite(indptr[i][n[i]], data[i], is_complex, e2, sizeof e2);
        rc[i] = r;
        if (r != FPLUS_OK) continue;
        idx.push_back(i); bn.push_back(n[i]); bp.push_back(indptr[i]); bi.push_back(indices[i]);
    }
    try {
        std::vector<const double*> hx(idx.size());
        for (size_t t = 0; t < idx.size(); t++) hx[t] = static_cast<const double*>(data[idx[t]]);
        std::vector<fplus_factors> o(idx.size());
        std::vector<int> r(idx.size());
        const int ret = row_factor_batch((int64_t)idx.size(), bn.data(), bp.data(), bi.data(), hx.data(), vs, drop_tol, pivot_strategy,
                                         pivot_rel, o.data(), r.data(), err, errlen);
        if (ret != FPLUS_OK) {
            for (auto& f : o) fplus_factors_free(&f);
            return ret;
        }
        for (size_t t = 0; t < idx.size(); t++) { out[idx[t]] = o[t]; rc[idx[t]] = r[t]; }
        return FPLUS_OK;
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
}

int fplus_amd_row_stats(fplus_amd_row_info* info)
{
    if (!info || !t_row_info_set) return FPLUS_ERR_BADARG;
    *info = t_row_info;
    return FPLUS_OK;
}

}  // extern "C"
