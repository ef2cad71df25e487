// numeric.cpp -- the FPLUS C interface of the MI355X build: options, the pipelines (permute -> Ruiz -> supernodal LU),
// the numeric object with its device factors, the solves, the CSC export, and the low-level entry points.
//
// Follows the reference's high-level layer (src/numeric.c:15-131, src/pipeline.c:84-127): the same composition of the
// orders and scalings, the same error convention.  Every arithmetic step runs on the device; a missing device is
// FPLUS_ERR_GPU.
#include "host_common.h"
#include "host_pool.h"
#include "snfactor.h"
#include "solve_dev.h"
#include "symbolic.h"
#include "prof.h"
#include "fplus_amd.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <memory>
#include <new>
#include <thread>

using namespace fpa;

namespace {

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// the host side of a permuted pattern: B = A[row_perm, :][:, col_perm], canonical, and src[o] = the entry of A that
// lands at o (src/pipeline.c:14-82 computes the same arrays)
void permuted_pattern(int64_t n, const int64_t* Ap, const int64_t* Ai, const int64_t* rp, const int64_t* cp,
                      std::vector<int64_t>& Bp, std::vector<int64_t>& Bi, std::vector<int64_t>& src)
{
    const int64_t nnz = Ap[n];
    std::vector<int64_t> inv(n), rowp(n + 1, 0), tcol(nnz), tsrc(nnz), next(n + 1);
    for (int64_t k = 0; k < n; k++) inv[rp[k]] = k;
    Bp.assign(n + 1, 0);
    for (int64_t k = 0; k < n; k++) Bp[k + 1] = Bp[k] + Ap[cp[k] + 1] - Ap[cp[k]];
    for (int64_t q = 0; q < nnz; q++) rowp[inv[Ai[q]] + 1]++;
    for (int64_t r = 0; r < n; r++) rowp[r + 1] += rowp[r];
    std::copy(rowp.begin(), rowp.end() - 1, next.begin());
    for (int64_t k = 0; k < n; k++)
        for (int64_t q = Ap[cp[k]]; q < Ap[cp[k] + 1]; q++) {
            const int64_t t = next[inv[Ai[q]]]++;
            tcol[t] = k;
            tsrc[t] = q;
        }
    Bi.resize(nnz);
    src.resize(nnz);
    std::copy(Bp.begin(), Bp.end() - 1, next.begin());
    for (int64_t r = 0; r < n; r++)
        for (int64_t t = rowp[r]; t < rowp[r + 1]; t++) {
            const int64_t o = next[tcol[t]]++;
            Bi[o] = r;
            src[o] = tsrc[t];
        }
}

std::shared_ptr<PatternCache> make_cache(int64_t n, const int64_t* Ap, const int64_t* Ai, const int64_t* rp, const int64_t* cp,
                                         hipStream_t s)
{
    auto c = std::make_shared<PatternCache>();
    c->n = n;
    c->nnz = Ap[n];
    std::vector<int64_t> src;
    permuted_pattern(n, Ap, Ai, rp, cp, c->Bp, c->Bi, src);
    std::vector<int32_t> bi32(c->nnz), ecol(c->nnz);
    for (int64_t q = 0; q < c->nnz; q++) bi32[q] = (int32_t)c->Bi[q];
    for (int64_t j = 0; j < n; j++)
        for (int64_t q = c->Bp[j]; q < c->Bp[j + 1]; q++) ecol[q] = (int32_t)j;
    c->d_src.alloc(c->nnz);
    c->d_src.upload(src.data(), c->nnz, s);
    c->d_Bp.alloc(n + 1);
    c->d_Bp.upload(c->Bp.data(), n + 1, s);
    c->d_Bi.alloc(c->nnz);
    c->d_Bi.upload(bi32.data(), c->nnz, s);
    c->d_ecol.alloc(c->nnz);
    c->d_ecol.upload(ecol.data(), c->nnz, s);
    dev_sync(s);
    return c;
}

// Ruiz on the device (src/equil.c:186-441): the range guard, up to max_iter iterations, the fold of the prescale.
int ruiz_device(int64_t n, int64_t nnz, const int32_t* d_erow, const int32_t* d_ecol, double* d_x, int vs, int max_iter,
                double tol, double* d_Dr, double* d_Dc, int verbose, int* iters, char* err, size_t errlen)
{
    hipStream_t s = devctx().stream;
    const double t0 = now_s();
    DevBuf<unsigned long long> d_m(2 * (size_t)(n > 0 ? n : 1) + 2);
    DevBuf<double> d_sc(2 * (size_t)(n > 0 ? n : 1));
    DevBuf<k::RuizStatus> d_st(1);
    unsigned long long* amax = d_m.p + 2 * n;
    unsigned long long* amin = amax + 1;
    // range guard
    int k4 = 0;
    {
        const unsigned long long init[2] = {0ULL, (unsigned long long)0x7FF0000000000000ULL};   // 0, +inf
        copy_h2d(amax, init, sizeof(init), s);
        k::absminmax(nnz * vs, d_x, amax, amin, s);
        unsigned long long hv[2];
        copy_d2h(hv, amax, sizeof(hv), s);
        dev_sync(s);
        double vmax, vmin;
        std::memcpy(&vmax, &hv[0], 8);
        std::memcpy(&vmin, &hv[1], 8);
        if (vmax > 0.0 && (vmax > 0x1p500 || vmin < 0x1p-500)) {
            int emax = ilogb(vmax), emin = ilogb(vmin);
            if (emax - emin > 990) {
                k::fill_u64(d_m.p, 2 * n, 0ULL, s);
                k::line_absmax(vs, nnz, d_erow, d_ecol, d_x, d_m.p, n, s);
                unsigned long long lo = ULLONG_MAX;
                copy_h2d(amin, &lo, 8, s);
                k::min_positive(2 * n, d_m.p, amin, s);
                copy_d2h(&lo, amin, 8, s);
                dev_sync(s);
                double lmin;
                std::memcpy(&lmin, &lo, 8);
                const int elmin = ilogb(lmin);
                if (emax - elmin > 990) {
                    errf(err, errlen, "line maxima span 2^%d .. 2^%d (entries 2^%d ..): wider than Ruiz equilibration can handle "
                         "in double precision; factor without equilibration (equilibrate=False)", elmin, emax, emin);
                    return FPLUS_ERR_BADARG;
                }
                emin = elmin;
            }
            k4 = -((emax + emin) / 2 / 4) * 4;
            k::ruiz_scale_all(nnz * vs, d_x, ldexp(1.0, k4), s);
        }
    }
    k::fill_f64(d_Dr, n, 1.0, s);
    k::fill_f64(d_Dc, n, 1.0, s);
    int used = 0;
    double final_res = 0;
    for (int it = 0; it < max_iter; it++) {
        k::RuizStatus st0;
        st0.res_bits = 0;
        st0.bad_row = st0.bad_col = INT64_MAX;
        d_st.upload(&st0, 1, s);
        k::fill_u64(d_m.p, 2 * n, 0ULL, s);
        k::ruiz_rowmax(vs, nnz, d_erow, d_x, d_m.p, s);
        k::ruiz_line_finish(n, d_m.p, d_Dr, d_sc.p, d_st.p, 0, s);
        k::ruiz_scale_colmax(vs, nnz, d_erow, d_ecol, d_x, d_sc.p, d_m.p + n, s);
        k::ruiz_line_finish(n, d_m.p + n, d_Dc, d_sc.p + n, d_st.p, 1, s);
        k::ruiz_scale_cols(vs, nnz, d_ecol, d_x, d_sc.p + n, s);
        k::RuizStatus st;
        d_st.download(&st, 1, s);
        dev_sync(s);
        if (st.bad_row != INT64_MAX) {
            errf(err, errlen, "Row %ld of the matrix is entirely zero (structural singularity detected during equilibration).", (long)st.bad_row);
            return FPLUS_ERR_SINGULAR;
        }
        if (st.bad_col != INT64_MAX) {
            errf(err, errlen, "Column %ld of the matrix is entirely zero (structural singularity detected during equilibration).", (long)st.bad_col);
            return FPLUS_ERR_SINGULAR;
        }
        double res;
        std::memcpy(&res, &st.res_bits, 8);
        used = it + 1;
        final_res = res;
        if (verbose >= 2) printf("[FPLUS]   iter %3d  residual=%.3e\n", it + 1, res);
        if (res < tol) break;
    }
    if (k4 != 0) k::ruiz_scale_vec2(n, d_Dr, d_Dc, ldexp(1.0, k4 / 2), s);
    dev_sync(s);
    if (verbose >= 1)
        printf("[FPLUS] Ruiz %s: %d iters, final res=%.3e, time=%.3fs (device)\n", final_res < tol ? "converged" : "max_iter reached",
               used, final_res, now_s() - t0);
    if (iters) *iters = used;
    return FPLUS_OK;
}

int gpu_error(const std::exception& e, char* err, size_t errlen)
{
    errf(err, errlen, "GPU: %s", e.what());
    return FPLUS_ERR_GPU;
}

}  // namespace

#include "numeric_obj.h"

namespace {

// dependency levels of the blocks for the four sweeps, and the device copies the solves read
void prepare_solves(fplus_numeric* num, hipStream_t s)
{
    DevFactor& F = num->F;
    const int64_t n = F.n;
    const int32_t nb = (int32_t)F.blocks.size();
    double tp = now_s();
    auto lap = [&](int i) { const double t = now_s(); num->t_prep_part[i] = t - tp; tp = t; };
    HostPool pool(host_threads() + 1);
    std::vector<int32_t> colblk(n > 0 ? n : 1);
    // every block's U11^-1, one job per column, on the device while the host prepares the rest (kept until the
    // final synchronization below: the upload reads the host list)
    std::vector<int2> ujobs;
    ujobs.reserve(n > 0 ? n : 1);
    double uflops = 0;
    for (int32_t b = 0; b < nb; b++) {
        const double w = F.blocks[b].w;
        uflops += (F.vs == 1 ? 2.0 : 8.0) * w * w * w / 6;
        for (int32_t c = 0; c < F.blocks[b].w; c++) {
            colblk[F.blocks[b].c0 + c] = b;
            ujobs.push_back(make_int2(b, c));
        }
    }
    num->n_ujobs = (int64_t)ujobs.size();
    num->uinv_flops = uflops;
    if (!ujobs.empty()) {
        num->d_ujobs.reserve(ujobs.size());
        num->d_ujobs.upload(ujobs.data(), ujobs.size(), s);
        k::blk_uinv(F.vs, (int64_t)ujobs.size(), num->d_ujobs.p, F.d_blocks.p, uflops, s);
    }
    auto order_levels = [nb](const std::vector<int32_t>& lvl, std::vector<int32_t>& blk, std::vector<int32_t>& off) {
        int32_t mx = 0;
        for (int32_t b = 0; b < nb; b++) mx = std::max(mx, lvl[b]);
        off.assign(mx + 2, 0);
        for (int32_t b = 0; b < nb; b++) off[lvl[b] + 1]++;
        for (int32_t L = 0; L <= mx; L++) off[L + 1] += off[L];
        blk.resize(nb);
        std::vector<int32_t> fill(off.begin(), off.end() - 1);
        for (int32_t b = 0; b < nb; b++) blk[fill[lvl[b]]++] = b;
    };
    std::vector<int32_t> lvl(nb > 0 ? nb : 1, 0), rowlvl(n > 0 ? n : 1, 0);
    // L y = b: a block waits for every block that writes its pivot rows
    for (int32_t b = 0; b < nb; b++) {
        const BlockDesc& B = F.blocks[b];
        int32_t L = 0;
        for (int32_t c = 0; c < B.w; c++) L = std::max(L, rowlvl[B.c0 + c]);
        lvl[b] = L;
        for (int32_t r = B.w; r < B.nrows; r++) {
            int32_t& rl = rowlvl[F.pinv[F.block_rows[b][r]]];
            rl = std::max(rl, L + 1);
        }
    }
    order_levels(lvl, num->blk_lf, num->lev_lf);
    // L^T x = y: a block waits for the blocks holding its rows below
    for (int32_t b = nb - 1; b >= 0; b--) {
        const BlockDesc& B = F.blocks[b];
        int32_t L = 0;
        for (int32_t r = B.w; r < B.nrows; r++) L = std::max(L, lvl[colblk[F.pinv[F.block_rows[b][r]]]] + 1);
        lvl[b] = L;
    }
    // lvl is rewritten in place: blocks after b were assigned before b reads them
    order_levels(lvl, num->blk_ltb, num->lev_ltb);
    lap(0);
    // U x = y: the pieces of every block (counted, then placed in column order), and the
    // levels: a block waits for the blocks of its pieces' columns
    std::vector<int64_t> pb(nb + 1, 0);
    // (filled in full: no zero-initialisation)
    std::unique_ptr<int32_t[]> pj, pa;
    std::unique_ptr<int64_t[]> pu;
    {
        // column j's runs over the rows of blocks other than its own, in column order, over the columns [j0, j1)
        auto for_pieces = [&](int64_t j0, int64_t j1, auto&& fn) {
            for (int64_t j = j0; j < j1; j++) {
                const int32_t bj = colblk[j];
                int64_t uo = F.Up[j];
                for (int64_t r = F.Urp[j]; r < F.Urp[j + 1]; r++) {
                    const int64_t r0 = F.Ur_row[r], r1 = r0 + F.Ur_len[r];
                    for (int64_t i = r0; i < r1;) {
                        const int32_t b = colblk[i];
                        const BlockDesc& B = F.blocks[b];
                        const int64_t e = std::min<int64_t>(r1, (int64_t)B.c0 + B.w);
                        if (b != bj) {
                            if (e != (int64_t)B.c0 + B.w) throw std::logic_error("internal: a U run ends inside a block");
                            fn(b, j, i - B.c0, uo + (i - r0));
                        }
                        i = e;
                    }
                    uo += F.Ur_len[r];
                }
            }
        };
        // in parallel: chunks of columns count their pieces per block, the chunks' places within each block follow in
        // chunk order (so every block's pieces are in column order: a fixed summation order), the chunks fill them
        const int64_t nch = std::min<int64_t>(std::max<int64_t>(n, 1), 4 * (int64_t)pool.size()), per = (n + nch - 1) / nch;
        std::vector<std::vector<int64_t>> cnt(nch, std::vector<int64_t>(nb + 1, 0));
        pool.run(nch, [&](int64_t c, int) {
            std::vector<int64_t>& h = cnt[c];
            for_pieces(c * per, std::min(n, (c + 1) * per), [&](int32_t b, int64_t, int64_t, int64_t) { h[b]++; });
        });
        for (int32_t b = 0; b < nb; b++) {
            int64_t o = pb[b];
            for (int64_t c = 0; c < nch; c++) { const int64_t k = cnt[c][b]; cnt[c][b] = o; o += k; }
            pb[b + 1] = o;
        }
        const int64_t np = pb[nb];
        if (np >= ((int64_t)1 << 31)) throw std::length_error("too many U pieces for the solve's 32-bit indices");
        const size_t npa = (size_t)(np > 0 ? np : 1);
        pj.reset(new int32_t[npa]);
        pa.reset(new int32_t[npa]);
        pu.reset(new int64_t[npa]);
        pool.run(nch, [&](int64_t c, int) {
            std::vector<int64_t>& fill = cnt[c];
            for_pieces(c * per, std::min(n, (c + 1) * per), [&](int32_t b, int64_t j, int64_t a, int64_t u) {
                const int64_t o = fill[b]++;
                pj[o] = (int32_t)j; pa[o] = (int32_t)a; pu[o] = u - a;
            });
        });
        num->d_pu.alloc(npa);
        num->d_pu.upload(pu.get(), (size_t)np, s);
        num->d_pj.alloc(npa);
        num->d_pj.upload(pj.get(), (size_t)np, s);
        num->d_pa.alloc(npa);
        num->d_pa.upload(pa.get(), (size_t)np, s);
        dev_sync(s);
    }
    lap(1);
    // the levels of the two U sweeps from per-block lists of the distinct blocks they wait for, made in parallel
    std::vector<std::vector<int32_t>> dep(nb > 0 ? nb : 1);
    pool.run(nb, [&](int64_t b, int) {   // U x = y: the blocks of the pieces' columns (ascending: equal ones adjacent)
        std::vector<int32_t>& d = dep[b];
        d.clear();
        for (int64_t o = pb[b]; o < pb[b + 1]; o++) {
            const int32_t bj = colblk[pj[o]];
            if (d.empty() || d.back() != bj) d.push_back(bj);
        }
    });
    for (int32_t b = nb - 1; b >= 0; b--) {
        int32_t L = 0;
        for (const int32_t bj : dep[b]) L = std::max(L, lvl[bj] + 1);
        lvl[b] = L;
    }
    order_levels(lvl, num->blk_ub, num->lev_ub);
    {   // the persistent U sweep waits per block for these blocks (ascending)
        std::vector<int32_t> off(nb + 1, 0), all;
        for (int32_t b = 0; b < nb; b++) off[b + 1] = off[b] + (int32_t)dep[b].size();
        all.resize(off[nb] > 0 ? off[nb] : 1);
        for (int32_t b = 0; b < nb; b++) std::copy(dep[b].begin(), dep[b].end(), all.begin() + off[b]);
        num->d_udep_off.alloc(off.size());
        num->d_udep_off.upload(off.data(), off.size(), s);
        num->d_udep.alloc(all.size());
        num->d_udep.upload(all.data(), all.size(), s);
        dev_sync(s);
    }
    // U^T y = b: a block waits for the blocks holding the rows of its U columns above it
    pool.run(nb, [&](int64_t b, int) {
        const BlockDesc& B = F.blocks[b];
        std::vector<int32_t>& d = dep[b];
        d.clear();
        for (int64_t j = B.c0; j < (int64_t)B.c0 + B.w; j++)
            for (int64_t r = F.Urp[j]; r < F.Urp[j + 1]; r++) {
                const int64_t r0 = F.Ur_row[r], r1 = std::min<int64_t>(r0 + F.Ur_len[r], B.c0);
                for (int64_t i = r0; i < r1;) {
                    const int32_t bi = colblk[i];
                    d.push_back(bi);
                    i = (int64_t)F.blocks[bi].c0 + F.blocks[bi].w;
                }
            }
        std::sort(d.begin(), d.end());
        d.erase(std::unique(d.begin(), d.end()), d.end());
    });
    for (int32_t b = 0; b < nb; b++) {
        int32_t L = 0;
        for (const int32_t bi : dep[b]) L = std::max(L, lvl[bi] + 1);
        lvl[b] = L;
    }
    order_levels(lvl, num->blk_utf, num->lev_utf);
    lap(2);
    const double vb = 8.0 * F.vs;   // bytes per scalar
    const int R0 = k::solve_rhs_per_pass(F.vs);
    // contribution slots and the rows' lists of them (blocks in column order: a deterministic summation order)
    {
        std::vector<int64_t> slot_off(nb > 0 ? nb : 1), incp(n + 1, 0), incs;
        int64_t ns = 0;
        for (int32_t b = 0; b < nb; b++) {
            slot_off[b] = ns;
            ns += F.blocks[b].nrows - F.blocks[b].w;
        }
        // in parallel over chunks of blocks: each chunk counts its slots per row, every row's slots are placed chunk
        // by chunk (so still in block order), then the chunks fill their places
        const int32_t nch = (int32_t)std::min<int64_t>(std::max(nb, 1), pool.size()), per = (nb + nch - 1) / nch;
        std::vector<std::vector<int32_t>> cnt(nch);
        pool.run(nch, [&](int64_t c, int) {
            std::vector<int32_t>& h = cnt[c];
            h.assign(n > 0 ? n : 1, 0);
            for (int32_t b = (int32_t)c * per; b < std::min<int32_t>(nb, ((int32_t)c + 1) * per); b++) {
                const BlockDesc& B = F.blocks[b];
                const int32_t* br = F.block_rows[b].data();
                for (int32_t r = B.w; r < B.nrows; r++) h[F.pinv[br[r]]]++;
            }
        });
        for (int64_t i = 0; i < n; i++) {
            int64_t t = 0;
            for (int32_t c = 0; c < nch; c++) t += cnt[c][i];
            incp[i + 1] = incp[i] + t;
        }
        std::vector<std::vector<int64_t>> fill(nch);
        pool.run(nch, [&](int64_t c, int) {   // chunk c's first place in every row
            std::vector<int64_t>& f = fill[c];
            f.resize(n > 0 ? n : 1);
            for (int64_t i = 0; i < n; i++) {
                int64_t o = incp[i];
                for (int32_t c2 = 0; c2 < c; c2++) o += cnt[c2][i];
                f[i] = o;
            }
        });
        incs.resize(ns > 0 ? ns : 1);
        pool.run(nch, [&](int64_t c, int) {
            std::vector<int64_t>& f = fill[c];
            for (int32_t b = (int32_t)c * per; b < std::min<int32_t>(nb, ((int32_t)c + 1) * per); b++) {
                const BlockDesc& B = F.blocks[b];
                const int32_t* br = F.block_rows[b].data();
                for (int32_t r = B.w; r < B.nrows; r++) incs[f[F.pinv[br[r]]]++] = slot_off[b] + r - B.w;
            }
        });
        num->nslots = ns;
        auto up64 = [s](DevBuf<int64_t>& d, const std::vector<int64_t>& h) { d.alloc(h.size()); d.upload(h.data(), h.size(), s); };
        up64(num->d_slot_off, slot_off);
        up64(num->d_incp, incp);
        up64(num->d_incs, incs);
        lap(3);
        {   // the rows each L level gathers: the pivot rows of its blocks
            std::vector<int32_t> grows;
            const size_t nlev0 = num->lev_lf.size() - 1;
            num->lev_gr.assign(nlev0 + 1, 0);
            num->bytes_lf_gather.assign(nlev0, 0.0);
            for (size_t L = 0; L < nlev0; L++) {
                for (int32_t k2 = num->lev_lf[L]; k2 < num->lev_lf[L + 1]; k2++) {
                    const BlockDesc& B = F.blocks[num->blk_lf[k2]];
                    for (int32_t c = 0; c < B.w; c++) {
                        grows.push_back(B.c0 + c);
                        num->bytes_lf_gather[L] += (double)(incp[B.c0 + c + 1] - incp[B.c0 + c]) * (8.0 + R0 * 8.0 * F.vs);
                    }
                }
                num->lev_gr[L + 1] = (int32_t)grows.size();
            }
            num->d_grows.alloc(grows.size() > 0 ? grows.size() : 1);
            num->d_grows.upload(grows.data(), grows.size(), s);
            dev_sync(s);
        }
        // push tiles (64 rows below a block each) and triangle tiles (64 of its rows) per level, and the bytes of every
        // level of the four sweeps
        std::vector<int2> tiles, ptiles;
        const size_t nlev = num->lev_lf.size() - 1;
        num->lev_lft.assign(nlev + 1, 0);
        num->lev_lpt.assign(nlev + 1, 0);
        num->bytes_lf_pull.assign(nlev, 0.0);
        num->bytes_lf_push.assign(nlev, 0.0);
        for (size_t L = 0; L < nlev; L++) {
            for (int32_t k2 = num->lev_lf[L]; k2 < num->lev_lf[L + 1]; k2++) {
                const int32_t b = num->blk_lf[k2];
                const BlockDesc& B = F.blocks[b];
                for (int32_t r0 = 0; r0 < B.nrows - B.w; r0 += 64) tiles.push_back(make_int2(b, r0));
                for (int32_t r0 = 0; r0 < B.w; r0 += 64) ptiles.push_back(make_int2(b, r0));
                num->bytes_lf_pull[L] += (double)B.w * (B.w + 1) / 2 * vb + 2.0 * B.w * R0 * vb;
                num->bytes_lf_push[L] += (double)(B.nrows - B.w) * B.w * vb + (double)(B.nrows - B.w) * R0 * vb;
            }
            num->lev_lft[L + 1] = (int32_t)tiles.size();
            num->lev_lpt[L + 1] = (int32_t)ptiles.size();
        }
        num->d_lft.alloc(tiles.size() > 0 ? tiles.size() : 1);
        num->d_lft.upload(tiles.data(), tiles.size(), s);
        num->d_lpt.alloc(ptiles.size() > 0 ? ptiles.size() : 1);
        num->d_lpt.upload(ptiles.data(), ptiles.size(), s);
        dev_sync(s);
        {   // the persistent L sweep: per block the blocks writing its pivot rows (those whose rows below hold them), and
            // the items level by level: the blocks' gathers, their triangle tiles, their push tiles
            std::vector<std::vector<int32_t>> tgt(nb > 0 ? nb : 1);
            pool.run(nb, [&](int64_t b, int) {
                const BlockDesc& B = F.blocks[b];
                const int32_t* br = F.block_rows[b].data();
                std::vector<int32_t>& t = tgt[b];
                t.clear();
                for (int32_t r = B.w; r < B.nrows; r++) t.push_back(colblk[F.pinv[br[r]]]);
                std::sort(t.begin(), t.end());
                t.erase(std::unique(t.begin(), t.end()), t.end());
            });
            std::vector<int32_t> off(nb + 1, 0);
            for (int32_t b = 0; b < nb; b++)
                for (const int32_t t : tgt[b]) off[t + 1]++;
            for (int32_t b = 0; b < nb; b++) off[b + 1] += off[b];
            std::vector<int32_t> all(off[nb] > 0 ? off[nb] : 1), fl(off.begin(), off.end() - 1);
            for (int32_t b = 0; b < nb; b++)
                for (const int32_t t : tgt[b]) all[fl[t]++] = b;
            num->d_ldep_off.alloc(off.size());
            num->d_ldep_off.upload(off.data(), off.size(), s);
            num->d_ldep.alloc(all.size());
            num->d_ldep.upload(all.data(), all.size(), s);
            std::vector<int4> items;
            double bytes = 0;
            for (size_t L = 0; L < nlev; L++) {
                const int32_t k0 = num->lev_lf[L], k1 = num->lev_lf[L + 1];
                for (int32_t k2 = k0; k2 < k1; k2++) {
                    const int32_t b = num->blk_lf[k2];
                    for (int32_t r0 = 0; r0 < F.blocks[b].w; r0 += 64) items.push_back(make_int4(b, r0, 0, 0));
                }
                for (int32_t k2 = k0; k2 < k1; k2++) {
                    const int32_t b = num->blk_lf[k2];
                    for (int32_t r0 = 0; r0 < F.blocks[b].w; r0 += 64) items.push_back(make_int4((1 << 29) | b, r0, 0, 0));
                }
                for (int32_t k2 = k0; k2 < k1; k2++) {
                    const int32_t b = num->blk_lf[k2];
                    for (int32_t r0 = 0; r0 < F.blocks[b].nrows - F.blocks[b].w; r0 += 64) items.push_back(make_int4((2 << 29) | b, r0, 0, 0));
                }
                bytes += num->bytes_lf_gather[L] + num->bytes_lf_pull[L] + num->bytes_lf_push[L];
            }
            if (items.size() >= ((size_t)1 << 31)) throw std::length_error("too many items for the persistent L sweep");
            num->dfl_nitems = (int32_t)items.size();
            num->d_dfl_items.alloc(items.size() > 0 ? items.size() : 1);
            num->d_dfl_items.upload(items.data(), items.size(), s);
            num->df_bytes_l = bytes;
            dev_sync(s);
        }
    }
    {
        const size_t nub = num->lev_ub.size() - 1;
        num->bytes_ub.assign(nub, 0.0);
        num->bytes_ub_prod.assign(nub, 0.0);
        num->lev_ubt.assign(nub + 1, 0);
        num->lev_upt.assign(nub + 1, 0);
        std::vector<int4> ubt, ubr;
        std::vector<int2> upt;
        num->lev_ubr.assign(nub + 1, 0);
        num->ub_slots_max = 0;
        for (size_t L = 0; L < nub; L++) {
            int32_t slot = 0;
            for (int32_t k2 = num->lev_ub[L]; k2 < num->lev_ub[L + 1]; k2++) {
                const int32_t b = num->blk_ub[k2];
                const BlockDesc& B = F.blocks[b];
                double ent = 0;
                for (int64_t o = pb[b]; o < pb[b + 1]; o++) ent += (double)(B.w - pa[o]);
                num->bytes_ub[L] += ent * vb + (double)(pb[b + 1] - pb[b]) * (12.0 + R0 * vb) + 2.0 * B.w * R0 * vb;
                num->bytes_ub_prod[L] += (double)B.w * (B.w + 1) / 2 * vb + 2.0 * B.w * R0 * vb;
                for (int32_t r0 = 0; r0 < B.w; r0 += 64) upt.push_back(make_int2(b, r0));
                const int32_t first = slot;
                for (int64_t p = pb[b]; p < pb[b + 1]; p += k::UB_PC)
                    ubt.push_back(make_int4(b, (int32_t)p, (int32_t)std::min<int64_t>(k::UB_PC, pb[b + 1] - p), slot++));
                if (slot > first)
                    for (int32_t row0 = 0; row0 < B.w; row0 += 16) ubr.push_back(make_int4(b, row0, first, slot - first));
            }
            num->lev_ubt[L + 1] = (int32_t)ubt.size();
            num->lev_ubr[L + 1] = (int32_t)ubr.size();
            num->lev_upt[L + 1] = (int32_t)upt.size();
            num->ub_slots_max = std::max<int64_t>(num->ub_slots_max, slot);
        }
        num->d_ubt.alloc(ubt.size() > 0 ? ubt.size() : 1);
        num->d_ubt.upload(ubt.data(), ubt.size(), s);
        num->d_ubr.alloc(ubr.size() > 0 ? ubr.size() : 1);
        num->d_ubr.upload(ubr.data(), ubr.size(), s);
        num->d_upt.alloc(upt.size() > 0 ? upt.size() : 1);
        num->d_upt.upload(upt.data(), upt.size(), s);
        dev_sync(s);
        {   // the persistent U sweep's items, level by level: a block's dot-product groups (of at most 4 UB_PC pieces
            // each: a few batches per item), its reduction, its triangle tiles
            std::vector<int4> items;
            std::vector<int64_t> ubase(nb > 0 ? nb : 1, 0);
            int64_t pbase = 0;
            double bytes = 0;
            for (size_t L = 0; L < nub; L++) {
                const int32_t k0 = num->lev_ub[L], k1 = num->lev_ub[L + 1];
                std::vector<int32_t> ngrp(k1 - k0, 0);
                for (int32_t k2 = k0; k2 < k1; k2++) {
                    const int32_t b = num->blk_ub[k2];
                    const int64_t np = pb[b + 1] - pb[b];
                    if (np == 0) continue;
                    const int64_t ng = (np + 4 * k::UB_PC - 1) / (4 * k::UB_PC), per = (np + ng - 1) / ng;
                    int32_t g = 0;
                    for (int64_t p = pb[b]; p < pb[b + 1]; p += per, g++)
                        items.push_back(make_int4(b, (int32_t)p, (int32_t)std::min<int64_t>(per, pb[b + 1] - p), g));
                    ngrp[k2 - k0] = g;
                    ubase[b] = pbase;
                    pbase += (int64_t)g * F.blocks[b].w;
                }
                for (int32_t k2 = k0; k2 < k1; k2++)
                    if (ngrp[k2 - k0] > 0) items.push_back(make_int4((1 << 29) | num->blk_ub[k2], ngrp[k2 - k0], 0, 0));
                for (int32_t k2 = k0; k2 < k1; k2++) {
                    const int32_t b = num->blk_ub[k2];
                    for (int32_t r0 = 0; r0 < F.blocks[b].w; r0 += 64) items.push_back(make_int4((2 << 29) | b, r0, ngrp[k2 - k0] > 0, 0));
                }
                bytes += num->bytes_ub[L] + num->bytes_ub_prod[L];
            }
            if (items.size() >= ((size_t)1 << 31)) throw std::length_error("too many items for the persistent U sweep");
            num->dfu_nitems = (int32_t)items.size();
            num->d_dfu_items.alloc(items.size() > 0 ? items.size() : 1);
            num->d_dfu_items.upload(items.data(), items.size(), s);
            num->d_ubase.alloc(ubase.size());
            num->d_ubase.upload(ubase.data(), ubase.size(), s);
            num->df_part = pbase;
            num->df_bytes_u = bytes;
            dev_sync(s);
        }
        const size_t nut = num->lev_utf.size() - 1, nlt = num->lev_ltb.size() - 1;
        {   // the transposed sweeps' dot-product tiles (one column each) per level
            auto tiles_of = [&](const std::vector<int32_t>& lev, const std::vector<int32_t>& blk, std::vector<int32_t>& lt, DevBuf<int2>& d) {
                std::vector<int2> tt;
                lt.assign(lev.size(), 0);
                for (size_t L = 0; L + 1 < lev.size(); L++) {
                    for (int32_t k2 = lev[L]; k2 < lev[L + 1]; k2++)
                        for (int32_t t0 = 0; t0 < F.blocks[blk[k2]].w; t0++) tt.push_back(make_int2(blk[k2], t0));
                    lt[L + 1] = (int32_t)tt.size();
                }
                d.alloc(tt.size() > 0 ? tt.size() : 1);
                d.upload(tt.data(), tt.size(), s);
            };
            tiles_of(num->lev_utf, num->blk_utf, num->lev_utt, num->d_utt);
            tiles_of(num->lev_ltb, num->blk_ltb, num->lev_ltt, num->d_ltt);
            // and the triangle-product tiles (64 rows each) per level
            auto ptiles_of = [&](const std::vector<int32_t>& lev, const std::vector<int32_t>& blk, std::vector<int32_t>& lt, DevBuf<int2>& d) {
                std::vector<int2> tt;
                lt.assign(lev.size(), 0);
                for (size_t L = 0; L + 1 < lev.size(); L++) {
                    for (int32_t k2 = lev[L]; k2 < lev[L + 1]; k2++)
                        for (int32_t r0 = 0; r0 < F.blocks[blk[k2]].w; r0 += 64) tt.push_back(make_int2(blk[k2], r0));
                    lt[L + 1] = (int32_t)tt.size();
                }
                d.alloc(tt.size() > 0 ? tt.size() : 1);
                d.upload(tt.data(), tt.size(), s);
            };
            ptiles_of(num->lev_utf, num->blk_utf, num->lev_utp, num->d_utp);
            ptiles_of(num->lev_ltb, num->blk_ltb, num->lev_ltp, num->d_ltp);
            dev_sync(s);
        }
        num->bytes_utf.assign(nut, 0.0);
        num->bytes_ltb.assign(nlt, 0.0);
        for (size_t L = 0; L < nut; L++)
            for (int32_t k2 = num->lev_utf[L]; k2 < num->lev_utf[L + 1]; k2++) {
                const BlockDesc& B = F.blocks[num->blk_utf[k2]];
                num->bytes_utf[L] += (double)(F.Up[B.c0 + B.w] - F.Up[B.c0]) * (vb + 4.0) + 2.0 * B.w * R0 * vb;
            }
        for (size_t L = 0; L < nlt; L++)
            for (int32_t k2 = num->lev_ltb[L]; k2 < num->lev_ltb[L + 1]; k2++) {
                const BlockDesc& B = F.blocks[num->blk_ltb[k2]];
                num->bytes_ltb[L] += (double)B.nrows * B.w * vb + (double)B.nrows * (4.0 + R0 * vb);
            }
    }
    lap(4);
    auto up32 = [s](DevBuf<int32_t>& d, const std::vector<int32_t>& h) { d.alloc(h.size()); d.upload(h.data(), h.size(), s); };
    up32(num->d_blk_lf, num->blk_lf);
    up32(num->d_blk_ub, num->blk_ub);
    up32(num->d_blk_utf, num->blk_utf);
    up32(num->d_blk_ltb, num->blk_ltb);
    std::vector<int32_t> pinv32(n > 0 ? n : 1);
    for (int64_t i = 0; i < n; i++) pinv32[i] = (int32_t)F.pinv[i];
    up32(num->d_pinv, pinv32);
    num->d_Up.alloc(n + 1);
    num->d_Up.upload(F.Up.data(), n + 1, s);
    num->d_p.alloc(n > 0 ? n : 1);
    num->d_p.upload(num->p.data(), n, s);
    num->d_q.alloc(n > 0 ? n : 1);
    num->d_q.upload(num->q.data(), n, s);
    num->d_Dr.alloc(n > 0 ? n : 1);
    num->d_Dr.upload(num->Dr.data(), n, s);
    num->d_Dc.alloc(n > 0 ? n : 1);
    num->d_Dc.upload(num->Dc.data(), n, s);
    dev_sync(s);
    lap(5);
}

}  // namespace

namespace fpa {
SweepScratch sweep_scratch_size(const fplus_numeric* num, int32_t R)
{
    const int vs = num->F.vs;
    SweepScratch z;
    z.n_y = (size_t)std::max<int64_t>(num->F.n, 1) * R * vs;
    z.n_contrib = (size_t)(num->nslots > 0 ? num->nslots : 1) * R * vs;
    z.n_part = (size_t)std::max<int64_t>(num->ub_slots_max, 1) * 256 * k::solve_rhs_per_pass(vs) * vs;   // a slot: 256 rows
    if (solve_dataflow()) {
        z.n_part = std::max(z.n_part, (size_t)std::max<int64_t>(num->df_part, 1) * R * vs);
        z.n_cnt = (size_t)k::df_counters((int64_t)num->F.blocks.size());
    }
    return z;
}

static std::atomic<bool> g_solve_dataflow{false};
bool solve_dataflow() { return g_solve_dataflow.load(std::memory_order_relaxed); }
void set_solve_dataflow(bool on) { g_solve_dataflow.store(on, std::memory_order_relaxed); }

// the sweeps on w (n x R row-major, pivot order, R <= solve_rhs_per_pass) in place, with the factorization's own
// scratch buffers (grown here when too small)
void device_sweeps(const fplus_numeric* num, double* w, int32_t R, int trans, hipStream_t s)
{
    SweepScratch z = sweep_scratch_size(num, R);
    num->d_y.reserve(z.n_y);
    z.y = num->d_y.p;
    if (trans == 0) {
        num->d_contrib.reserve(z.n_contrib);
        num->d_ub_part.reserve(z.n_part);
        z.contrib = num->d_contrib.p;
        z.part = num->d_ub_part.p;
        if (z.n_cnt) {
            num->d_dfcnt.reserve(z.n_cnt);
            z.cnt = num->d_dfcnt.p;
        }
    }
    device_sweeps(num, w, R, trans, z, s);
}

// the same with caller-given scratch (z's pointers, of at least z's sizes for R; transposed sweeps use y only)
void device_sweeps(const fplus_numeric* num, double* w, int32_t R, int trans, const SweepScratch& z, hipStream_t s)
{
    const DevFactor& F = num->F;
    const int vs = F.vs;
    const BlockDesc* blocks = F.d_blocks.p;
    if (trans == 0 && z.cnt && solve_dataflow()) {
        // L y = b from w into y, then U x = y from y into w, each as one persistent launch
        const int32_t nb = (int32_t)F.blocks.size();
        k::fill_i32(z.cnt, k::df_counters(nb), 0, s);
        k::lsweep_df(vs, num->dfl_nitems, num->d_dfl_items.p, blocks, num->d_ldep_off.p, num->d_ldep.p, num->d_incp.p, num->d_incs.p,
                     num->d_slot_off.p, w, z.y, z.contrib, z.cnt, nb, R, num->df_bytes_l, s);
        k::usweep_df(vs, num->dfu_nitems, num->d_dfu_items.p, blocks, num->d_udep_off.p, num->d_udep.p, num->d_pa.p, num->d_pj.p,
                     num->d_pu.p, F.d_Ux.p, num->d_ubase.p, w, z.y, z.part, z.cnt, nb, R, num->df_bytes_u, s);
        // a wait that gave up (which the ordering of the items rules out) leaves the results incomplete: the call fails
        // and later solves take the level sweeps
        int32_t ab = 0;
        hip_check(hipMemcpyAsync(&ab, z.cnt + 2, sizeof(int32_t), hipMemcpyDeviceToHost, s), "persistent sweeps' status");
        hip_check(hipStreamSynchronize(s), "persistent sweeps");
        if (ab) {
            set_solve_dataflow(false);
            throw GpuError("internal: a persistent sweep gave up waiting for its inputs; later solves use the level sweeps");
        }
        return;
    }
    if (trans == 0) {
        // L y = b from w into y, then U x = y from y into w
        double* y = z.y;
        for (size_t L = 0; L + 1 < num->lev_lf.size(); L++) {
            k::lfwd_gather(vs, num->lev_gr[L + 1] - num->lev_gr[L], num->d_grows.p + num->lev_gr[L], num->d_incp.p, num->d_incs.p,
                           z.contrib, w, R, num->bytes_lf_gather[L], s);
            k::lfwd_pull(vs, num->lev_lpt[L + 1] - num->lev_lpt[L], num->d_lpt.p + num->lev_lpt[L], blocks, w, y, R,
                         num->bytes_lf_pull[L], s);
            k::lfwd_push(vs, num->lev_lft[L + 1] - num->lev_lft[L], num->d_lft.p + num->lev_lft[L], blocks, num->d_slot_off.p, y,
                         z.contrib, R, num->bytes_lf_push[L], s);
        }
        double* part = z.part;
        for (size_t L = 0; L + 1 < num->lev_ub.size(); L++) {
            k::ub_dots(vs, num->lev_ubt[L + 1] - num->lev_ubt[L], num->d_ubt.p + num->lev_ubt[L], blocks, num->d_pa.p, num->d_pj.p,
                       num->d_pu.p, F.d_Ux.p, w, part, R, num->bytes_ub[L], s);
            k::ub_red(vs, num->lev_ubr[L + 1] - num->lev_ubr[L], num->d_ubr.p + num->lev_ubr[L], blocks, part, y, R, s);
            k::ub_prod(vs, num->lev_upt[L + 1] - num->lev_upt[L], num->d_upt.p + num->lev_upt[L], blocks, y, w, R,
                       num->bytes_ub_prod[L], s);
        }
    } else {
        // U^T y = b from w into y, then L^T x = y from y into w
        const int cj = trans == 2 && vs == 2;
        double* y = z.y;
        for (size_t L = 0; L + 1 < num->lev_utf.size(); L++) {
            k::utdots(vs, num->lev_utt[L + 1] - num->lev_utt[L], num->d_utt.p + num->lev_utt[L], blocks, num->d_Up.p, num->F.d_Ui.p,
                      F.d_Ux.p, y, w, R, cj, num->bytes_utf[L], s);
            k::ut_prod(vs, num->lev_utp[L + 1] - num->lev_utp[L], num->d_utp.p + num->lev_utp[L], blocks, w, y, R, cj, 0.0, s);
        }
        for (size_t L = 0; L + 1 < num->lev_ltb.size(); L++) {
            k::ltdots(vs, num->lev_ltt[L + 1] - num->lev_ltt[L], num->d_ltt.p + num->lev_ltt[L], blocks, num->d_pinv.p, w, y, R, cj,
                      num->bytes_ltb[L], s);
            k::lt_prod(vs, num->lev_ltp[L + 1] - num->lev_ltp[L], num->d_ltp.p + num->lev_ltp[L], blocks, y, w, R, cj, 0.0, s);
        }
    }
}

// solve for nrhs right-hand sides in device memory b (column-major, ld), in passes of solve_rhs_per_pass columns
void solve_on_device(const fplus_numeric* num, double* b, int64_t nrhs, int64_t ldb, int tr, hipStream_t s)
{
    const int vs = num->F.vs;
    const int64_t n = num->n;
    const int32_t RP = k::solve_rhs_per_pass(vs);
    num->d_w.reserve((size_t)(n * RP * vs));
    for (int64_t r0 = 0; r0 < nrhs; r0 += RP) {
        const int32_t R = (int32_t)std::min<int64_t>(RP, nrhs - r0);
        double* br = b + (size_t)(r0 * ldb) * vs;
        if (tr == 0) k::gather_scale(vs, n, R, num->d_p.p, num->d_Dr.p, br, ldb, num->d_w.p, s);
        else k::gather_scale(vs, n, R, num->d_q.p, num->d_Dc.p, br, ldb, num->d_w.p, s);
        device_sweeps(num, num->d_w.p, R, tr, s);
        if (tr == 0) k::scatter_scale(vs, n, R, num->d_q.p, num->d_Dc.p, num->d_w.p, br, ldb, s);
        else k::scatter_scale(vs, n, R, num->d_p.p, num->d_Dr.p, num->d_w.p, br, ldb, s);
    }
}

}  // namespace fpa

namespace {

// the factors as CSC arrays on the host (src/factor_sn.c:741-808 layout): L with the unit diagonal first, then the
// supernode's later pivot rows, then its other rows ascending by pivot index; U as stored (segments, diagonal last)
void export_factors(const DevFactor& F, int is_complex, fplus_factors* f)
{
    const int64_t n = F.n;
    const int vs = F.vs;
    const size_t sz = vs * sizeof(double);
    hipStream_t s = devctx().stream;
    std::memset(f, 0, sizeof(*f));
    f->n = n; f->is_complex = is_complex; f->csc = 1; f->nsupernodes = F.nsn;
    int64_t nnzL = 0;
    for (int64_t sidx = 0; sidx < F.nsn; sidx++) {
        const int64_t w = F.sn_first[sidx + 1] - F.sn_first[sidx], nr = (int64_t)F.sn_rows[sidx].size();
        nnzL += w * nr - w * (w - 1) / 2;
    }
    f->Lp = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    f->Li = (int64_t*)malloc((size_t)(nnzL > 0 ? nnzL : 1) * sizeof(int64_t));
    f->Lx = malloc((size_t)(nnzL > 0 ? nnzL : 1) * sz);
    f->Up = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    f->Ui = (int64_t*)malloc((size_t)(F.nnzU > 0 ? F.nnzU : 1) * sizeof(int64_t));
    f->Ux = malloc((size_t)(F.nnzU > 0 ? F.nnzU : 1) * sz);
    f->p = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    f->q = (int64_t*)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    if (!f->Lp || !f->Li || !f->Lx || !f->Up || !f->Ui || !f->Ux || !f->p || !f->q) { fplus_factors_free(f); throw std::bad_alloc(); }
    std::memcpy(f->Up, F.Up.data(), (n + 1) * sizeof(int64_t));
    for (int64_t j = 0, o = 0; j < n; j++)
        for (int64_t r = F.Urp[j]; r < F.Urp[j + 1]; r++)
            for (int32_t t = 0; t < F.Ur_len[r]; t++) f->Ui[o++] = F.Ur_row[r] + t;
    F.d_Ux.download((double*)f->Ux, (size_t)F.nnzU * vs, s);
    for (int64_t i = 0; i < n; i++) { f->p[i] = F.p[i]; f->q[i] = i; }
    // block values to the host, block by block
    std::vector<std::vector<double>> bv(F.blocks.size());
    for (size_t b = 0; b < F.blocks.size(); b++) {
        const BlockDesc& B = F.blocks[b];
        bv[b].resize((size_t)B.nrows * B.w * vs);
        copy_d2h(bv[b].data(), B.vals, bv[b].size() * sizeof(double), s);
    }
    dev_sync(s);
    std::vector<int32_t> colblk(n > 0 ? n : 1);
    for (size_t b = 0; b < F.blocks.size(); b++)
        for (int32_t c = 0; c < F.blocks[b].w; c++) colblk[F.blocks[b].c0 + c] = (int32_t)b;
    std::vector<int64_t> posin(n > 0 ? n : 1, -1);   // block row of a row in the current column's block
    double* Lx = (double*)f->Lx;
    int64_t o = 0;
    std::vector<std::pair<int64_t, int64_t>> ord;
    for (int64_t sidx = 0; sidx < F.nsn; sidx++) {
        const std::vector<int64_t>& rows = F.sn_rows[sidx];
        const int64_t c0 = F.sn_first[sidx], w = F.sn_first[sidx + 1] - c0, nr = (int64_t)rows.size();
        ord.clear();
        for (int64_t t = w; t < nr; t++) ord.push_back({F.pinv[rows[t]], rows[t]});
        std::sort(ord.begin(), ord.end());
        int32_t curb = -1;
        for (int64_t cc = 0; cc < w; cc++) {
            const int64_t j = c0 + cc;
            const int32_t b = colblk[j];
            const BlockDesc& B = F.blocks[b];
            if (b != curb) {
                if (curb >= 0) for (int32_t r : F.block_rows[curb]) posin[r] = -1;
                for (int32_t r = 0; r < B.nrows; r++) posin[F.block_rows[b][r]] = r;
                curb = b;
            }
            const int64_t bc = j - B.c0;
            const double* col = bv[b].data() + (size_t)bc * B.nrows * vs;
            f->Lp[j] = o;
            f->Li[o] = j;
            Lx[o * vs] = 1.0;
            if (vs == 2) Lx[o * vs + 1] = 0.0;
            o++;
            auto put = [&](int64_t row, int64_t pidx) {
                f->Li[o] = pidx;
                const int64_t r = posin[row];
                for (int v = 0; v < vs; v++) Lx[o * vs + v] = r >= 0 ? col[r * vs + v] : 0.0;
                o++;
            };
            for (int64_t rr = cc + 1; rr < w; rr++) put(rows[rr], c0 + rr);
            for (const auto& pr : ord) put(pr.second, pr.first);
        }
        if (curb >= 0) for (int32_t r : F.block_rows[curb]) posin[r] = -1;
    }
    f->Lp[n] = o;
}

int check_options(const fplus_options* o, char* err, size_t errlen)
{
    if (o->pivot_strategy != FPLUS_PIVOT_PARTIAL && o->pivot_strategy != FPLUS_PIVOT_ROOK && o->pivot_strategy != FPLUS_PIVOT_FULL) {
        errf(err, errlen, "pivot_strategy must be FPLUS_PIVOT_PARTIAL, _ROOK or _FULL (got %d)", o->pivot_strategy);
        return FPLUS_ERR_BADARG;
    }
    if (!(o->drop_tol >= 0.0)) { errf(err, errlen, "drop_tol must be >= 0 (got %g)", o->drop_tol); return FPLUS_ERR_BADARG; }
    if (!(o->pivot_rel >= 0.0)) { errf(err, errlen, "pivot_rel must be >= 0"); return FPLUS_ERR_BADARG; }
    return FPLUS_OK;
}

int64_t panel_width(int64_t panel)
{
    // the structure does not depend on the width: 256 columns, the widest block the solve kernels take
    (void)panel;
    return 256;
}

// the common core: values on the device (d_Bx, permuted, canonical pattern cache c), Ruiz, the factorization
int factor_core(fplus_numeric* num, const PatternCache& c, DevBuf<double>& d_Bx, int vs, int equilibrate, int ruiz_max_iter,
                double ruiz_tol, int64_t relax_rows, int64_t panel, double pivot_rel, int verbose, DevBuf<double>& d_Drp,
                DevBuf<double>& d_Dcp, const fplus_amd_options& ao, char* err, size_t errlen)
{
    hipStream_t s = devctx().stream;
    const int64_t n = c.n;
    const double t0 = now_s();
    d_Drp.alloc(n > 0 ? n : 1);
    d_Dcp.alloc(n > 0 ? n : 1);
    int rc = FPLUS_OK;
    if (equilibrate) rc = ruiz_device(n, c.nnz, c.d_Bi.p, c.d_ecol.p, d_Bx.p, vs, ruiz_max_iter, ruiz_tol, d_Drp.p, d_Dcp.p, verbose,
                                      &num->ruiz_iters, err, errlen);
    else { k::fill_f64(d_Drp.p, n, 1.0, s); k::fill_f64(d_Dcp.p, n, 1.0, s); }
    if (rc != FPLUS_OK) return rc;
    const double t1 = now_s();
    SnOptions so;
    so.relax_rows = relax_rows;
    so.pivot_rel = pivot_rel;
    so.panel = panel_width(panel);
    so.verbose = verbose;
    so.precision = ao.precision;
    so.record = ao.refactor != 0;
    so.sel = selector_args(ao.precision >= 2 ? ao.selector : nullptr, d_Drp.p, d_Dcp.p);
    if (verbose >= 1)
        printf("[FPLUS] LU factorisation start: n=%ld, nnz(A)=%ld, %s, supernodal left-looking (device), pivot=partial, relax_rows=%ld, panel=%ld\n",
               (long)n, (long)c.nnz, vs == 2 ? "complex128" : "float64", (long)relax_rows, (long)so.panel);
    rc = sn_factor_device(n, c.Bp.data(), c.Bi.data(), c.d_Bp.p, c.d_Bi.p, d_Bx.p, vs, so, num->F, err, errlen);
    if (rc == FPLUS_ERR_INTERNAL && num->F.coop_aborted) {
        // the cooperative panel LU needs its whole grid resident at once; when another process holds compute units it
        // can give up at a barrier: the factorization runs again with the one-workgroup panel LU
        if (verbose >= 1) printf("[FPLUS] the cooperative panel LU gave up at a barrier: factoring again with the one-workgroup panel LU\n");
        dev_sync(s);
        num->F = DevFactor();
        so.coop_panel_lu = false;
        rc = sn_factor_device(n, c.Bp.data(), c.Bi.data(), c.d_Bp.p, c.d_Bi.p, d_Bx.p, vs, so, num->F, err, errlen);
        num->F.coop_retried = true;
    }
    num->t_ruiz = t1 - t0;
    num->t_factor = now_s() - t1;
    if (rc == FPLUS_OK && verbose >= 1)
        printf("[FPLUS] LU factorisation complete: nnz(U)=%ld, supernodes=%ld, panels=%ld, levels=%ld, time=%.3fs (host %.3fs, wait %.3fs, "
               "of which %.3fs within the panel LU); workspace rows per panel: %.0f pivotal, %.0f not (max %ld)\n",
               (long)num->F.nnzU, (long)num->F.nsn, (long)num->F.panels, (long)num->F.levels, num->t_factor, num->F.t_host, num->F.t_wait,
               num->F.t_wait_lu,
               (double)num->F.rows_upper / std::max<int64_t>(1, num->F.panels), (double)num->F.rows_lower / std::max<int64_t>(1, num->F.panels),
               (long)num->F.rows_lower_max);
    if (rc == FPLUS_OK && verbose >= 1)
        printf("[FPLUS]   host phases: search %.3fs, tasks/levels %.3fs (union order and workspace rows %.3fs, of which rows not "
               "pivotal %.3fs; tasks and levels %.3fs, of which the level pass %.3fs), symbolic with pivots %.3fs (search %.3fs "
               "[candidate pass %.3fs], U %.3fs, join %.3fs, prune %.3fs), out %.3fs\n", num->F.t_p1, num->F.t_tasks, num->F.t_ws,
               num->F.t_ws_low, num->F.t_lev, num->F.t_lev_pass, num->F.t_p3, num->F.t_p3_dfs, num->F.t_p3_split, num->F.t_p3_u,
               num->F.t_p3_join,
               num->F.t_p3_prune, num->F.t_out);
    if (rc == FPLUS_OK && verbose >= 1) {
        const double* h = num->F.upd_ncol_flops;
        const double pf = std::max(num->F.upd_prod_flops, 1.0);
        printf("[FPLUS]   block updates: product %.3e flops, issued on the matrix cores %.3e (%.2fx); by task width <=8 %.1f%%, "
               "<=16 %.1f%%, <=32 %.1f%%, <=64 %.1f%%, more %.1f%%; %ld of the %ld tiles of shared levels atomic\n",
               num->F.upd_prod_flops, num->F.upd_mfma_flops, num->F.upd_mfma_flops / pf, 100 * h[0] / pf, 100 * h[1] / pf,
               100 * h[2] / pf, 100 * h[3] / pf, 100 * h[4] / pf, (long)num->F.atomic_tiles, (long)num->F.shared_level_tiles);
        for (int b = 0; b < 13; b++) {
            const auto& H = num->F.lev_hist[b];
            if (H.levels)
                printf("[FPLUS]     levels with <= %5d tiles: %7ld levels, %9ld tiles, issued %.3e flops, mean longest segment %.0f\n",
                       1 << b, (long)H.levels, (long)H.tiles, H.issued, H.maxm / (double)H.levels);
        }
    }
    return rc;
}

}  // namespace

extern "C" {

void fplus_factors_free(fplus_factors* f)
{
    if (!f) return;
    free(f->Lp); free(f->Li); free(f->Lx); free(f->Up); free(f->Ui); free(f->Ux); free(f->p); free(f->q);
    f->Lp = f->Li = f->Up = f->Ui = f->p = f->q = nullptr;
    f->Lx = f->Ux = nullptr;
}

void fplus_default_options(fplus_options* o)
{
    o->pin_dense = 1; o->tau_density = 0.5; o->tau_min_nnz = 10; o->postorder = 1;
    o->equilibrate = 1; o->ruiz_max_iter = 20; o->ruiz_tol = 1e-2;
    o->pivot_rel = FPLUS_PIVOT_REL_DEFAULT; o->pivot_strategy = FPLUS_PIVOT_PARTIAL; o->drop_tol = 0.0;
    o->relax_rows = 32; o->panel = 512;
    o->use_gpu = 1; o->gpu_min_flops = 5e7; o->verbose = 0;
}

int fplus_analyze_opts(int64_t n, const int64_t* Ap, const int64_t* Ai, const fplus_options* opt, fplus_symbolic** sym,
                       char* err, size_t errlen)
{
    fplus_options d;
    if (!opt) { fplus_default_options(&d); opt = &d; }
    return fplus_analyze(n, Ap, Ai, opt->pin_dense, opt->tau_density, opt->tau_min_nnz, nullptr, nullptr, opt->postorder, sym, err, errlen);
}

void fplus_numeric_free(fplus_numeric* num)
{
    if (!num) return;
    fplus_solve_plan_free(num->rplan);
    fplus_factors_free(&num->f);
    delete num;
}

static int check_amd_options(const fplus_amd_options* ao, char* err, size_t errlen)
{
    if (ao->precision < 0 || ao->precision > 4) {
        errf(err, errlen, "precision must be 0 (FP64), 1 (FP32), 2 (per block, by the selector), 3 (FP64, FP32 differences logged) "
                          "or 4 (BF16 hi/lo split)");
        return FPLUS_ERR_BADARG;
    }
    if (ao->precision == 2 && !ao->selector) { errf(err, errlen, "precision 2 needs a selector (fplus_amd_options.selector)"); return FPLUS_ERR_BADARG; }
    if (ao->static_pivots < 0 || ao->static_pivots > 1) { errf(err, errlen, "static_pivots must be 0 or 1"); return FPLUS_ERR_BADARG; }
    return FPLUS_OK;
}

// rook / full pivoting or an incomplete factorization: the row kernel on the permuted, equilibrated matrix, the
// factors in CSC with their solve plan, the orders and scalings composed (reference src/numeric.c:61-80)
static int factor_row_entry(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex,
                            const fplus_options* opt, fplus_numeric** out, char* err, size_t errlen)
{
    const int64_t n = sym->n;
    std::unique_ptr<fplus_numeric> num(new fplus_numeric());
    num->n = n;
    num->is_complex = is_complex;
    num->row = true;
    num->opt = *opt;
    num->rp = sym->row_perm;
    num->cp = sym->col_perm;
    std::vector<double> Drp(n > 0 ? n : 1), Dcp(n > 0 ? n : 1);
    const double t0 = now_s();
    int rc = fplus_factor_row_sym(sym, Ap, Ai, Ax, is_complex, opt->equilibrate, opt->ruiz_max_iter, opt->ruiz_tol, opt->verbose,
                                  opt->drop_tol, opt->pivot_strategy, opt->pivot_rel, &num->f, Drp.data(), Dcp.data(), err, errlen);
    if (rc != FPLUS_OK) return rc;
    num->exported = true;
    const int64_t* rp = sym->row_perm.data();
    const int64_t* cp = sym->col_perm.data();
    num->p.resize(n); num->q.resize(n); num->Dr.resize(n); num->Dc.resize(n);
    for (int64_t k2 = 0; k2 < n; k2++) {
        num->p[k2] = rp[num->f.p[k2]];   // p = row_perm[p_lu]
        num->q[k2] = cp[num->f.q[k2]];   // q = col_perm[q_lu]
        num->Dr[rp[k2]] = Drp[k2];
        num->Dc[cp[k2]] = Dcp[k2];
    }
    rc = fplus_solve_plan_create(n, num->f.Lp, num->f.Li, num->f.Up, num->f.Ui, &num->rplan, err, errlen);
    if (rc != FPLUS_OK) return rc;
    num->t_factor = num->t_total = now_s() - t0;
    *out = num.release();
    return FPLUS_OK;
}

static int factor_entry(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, bool device_values,
                        int is_complex, const fplus_options* opt, const fplus_amd_options* ao, fplus_numeric** out, char* err,
                        size_t errlen)
{
    *out = nullptr;
    fplus_options d;
    if (!opt) { fplus_default_options(&d); opt = &d; }
    fplus_amd_options ad;
    if (!ao) { fplus_amd_default_options(&ad); ao = &ad; }
    int rc = check_options(opt, err, errlen);
    if (rc == FPLUS_OK) rc = check_amd_options(ao, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const int64_t n = sym->n;
    if (Ap[n] != sym->nnz) {
        errf(err, errlen, "the matrix has %ld nonzeros, the symbolic analysis %ld", (long)Ap[n], (long)sym->nnz);
        return FPLUS_ERR_BADARG;
    }
    if (opt->pivot_strategy != FPLUS_PIVOT_PARTIAL || opt->drop_tol > 0.0) {
        if (device_values || ao->refactor || ao->precision != 0) {
            errf(err, errlen, "rook / full pivoting and drop_tol > 0 run the row kernel on host values, without the device "
                              "factorization's extensions (refactor, precision, device-resident values)");
            return FPLUS_ERR_BADARG;
        }
        try {
            return factor_row_entry(sym, Ap, Ai, Ax, is_complex, opt, out, err, errlen);
        } catch (const GpuError& e) {
            return gpu_error(e, err, errlen);
        } catch (const std::bad_alloc&) {
            errf(err, errlen, "out of memory");
            return FPLUS_ERR_NOMEM;
        }
    }
    rc = check_compressed(n, Ap, Ai, true, err, errlen);
    if (rc == FPLUS_OK && !device_values) rc = check_finite(Ap[n], Ax, is_complex, err, errlen);
    if (rc != FPLUS_OK) return rc;
    const int vs = is_complex ? 2 : 1;
    std::unique_ptr<fplus_numeric> num;
    try {
        const double t0 = now_s();
        hipStream_t s = devctx().stream;
        std::shared_ptr<PatternCache> c;
        {
            std::lock_guard<std::mutex> g(sym->mu);
            if (!sym->cache) sym->cache = make_cache(n, Ap, Ai, sym->row_perm.data(), sym->col_perm.data(), s);
            c = sym->cache;
        }
        num.reset(new fplus_numeric());
        num->n = n;
        num->is_complex = is_complex;
        const int64_t nnz = c->nnz;
        // the new values' equilibrated, analysed form goes into the factorization's own buffer (d_Abar, which the
        // refined solves read), so that a captured replay finds its input at the same address every time
        DevBuf<double> d_Ax;
        if (num->d_Abar.n < (size_t)(nnz > 0 ? nnz : 1) * vs) num->d_Abar.alloc((size_t)(nnz > 0 ? nnz : 1) * vs);
        DevBuf<double>& d_Bx = num->d_Abar;
        const double* src = static_cast<const double*>(Ax);
        if (!device_values) {
            d_Ax.alloc((size_t)(nnz > 0 ? nnz : 1) * vs);
            d_Ax.upload((const double*)Ax, (size_t)nnz * vs, s);
            src = d_Ax.p;
        }
        k::gather_vals(vs, nnz, c->d_src.p, src, d_Bx.p, s);
        num->cache = c;
        num->opt = *opt;
        num->aopt = *ao;
        num->rp = sym->row_perm;
        num->cp = sym->col_perm;
        DevBuf<double>& d_Drp = num->d_Drp;
        DevBuf<double>& d_Dcp = num->d_Dcp;
        rc = factor_core(num.get(), *c, d_Bx, vs, opt->equilibrate, opt->ruiz_max_iter, opt->ruiz_tol, opt->relax_rows, opt->panel,
                         opt->pivot_rel, opt->verbose, d_Drp, d_Dcp, *ao, err, errlen);
        if (rc != FPLUS_OK) return rc;
        // (d_Abar now holds the values equilibrated in place by Ruiz)
        num->d_plu.alloc(n > 0 ? n : 1);
        num->d_plu.upload(num->F.p.data(), n, s);
        std::vector<double> Drp(n), Dcp(n);
        d_Drp.download(Drp.data(), n, s);
        d_Dcp.download(Dcp.data(), n, s);
        dev_sync(s);
        const int64_t* rp = sym->row_perm.data();
        const int64_t* cp = sym->col_perm.data();
        num->p.resize(n); num->q.resize(n); num->Dr.resize(n); num->Dc.resize(n);
        for (int64_t k2 = 0; k2 < n; k2++) {
            num->p[k2] = rp[num->F.p[k2]];    // p = row_perm[p_lu]
            num->q[k2] = cp[k2];              // q = col_perm[q_lu], q_lu the identity
            num->Dr[rp[k2]] = Drp[k2];
            num->Dc[cp[k2]] = Dcp[k2];
        }
        const double tps = now_s();
        prepare_solves(num.get(), s);
        num->t_prepare = now_s() - tps;
        if (opt->verbose >= 1)
            printf("[FPLUS] solve preparation %.3fs (L levels %.3f, U pieces %.3f, U levels %.3f, slots %.3f, tiles %.3f, "
                   "uploads %.3f); %zu lower and %zu upper sweep levels\n", num->t_prepare, num->t_prep_part[0],
                   num->t_prep_part[1], num->t_prep_part[2], num->t_prep_part[3], num->t_prep_part[4], num->t_prep_part[5],
                   num->lev_lf.size() - 1, num->lev_ub.size() - 1);
        num->t_total = now_s() - t0;
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    } catch (const std::logic_error& e) {
        errf(err, errlen, "%s", e.what());
        return FPLUS_ERR_INTERNAL;
    }
    *out = num.release();
    return FPLUS_OK;
}

int fplus_factor(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, int is_complex,
                 const fplus_options* opt, fplus_numeric** out, char* err, size_t errlen)
{
    return factor_entry(sym, Ap, Ai, Ax, false, is_complex, opt, nullptr, out, err, errlen);
}

// a row-kernel factorization's solve: the right-hand sides scaled and permuted in groups of 32, the CSC factors' sweeps
// on the device through the plan, the result scattered back (reference src/numeric.c:85-131)
static int solve_row(const fplus_numeric* num, void* b, int64_t nrhs, int64_t ldb, int trans, char* err, size_t errlen)
{
    const int64_t n = num->n;
    const int vs = num->is_complex ? 2 : 1;
    const fplus_factors* f = &num->f;
    const int64_t* p = num->p.data();
    const int64_t* q = num->q.data();
    const double* Dr = num->Dr.data();
    const double* Dc = num->Dc.data();
    const int64_t g = std::min<int64_t>(nrhs, 32);
    std::vector<double> w((size_t)n * g * vs);
    double* x = static_cast<double*>(b);
    int rc = FPLUS_OK;
    std::lock_guard<std::mutex> lk(num->solve_mu);
    for (int64_t r0 = 0; r0 < nrhs && rc == FPLUS_OK; r0 += g) {
        const int64_t m = std::min(g, nrhs - r0);
        for (int64_t r = 0; r < m; r++) {
            const double* xb = x + (size_t)((r0 + r) * ldb) * vs;
            double* ww = w.data() + (size_t)(r * n) * vs;
            for (int64_t k2 = 0; k2 < n; k2++) {
                const int64_t src = trans == 0 ? p[k2] : q[k2];
                const double d = trans == 0 ? Dr[src] : Dc[src];
                for (int v = 0; v < vs; v++) ww[k2 * vs + v] = d * xb[src * vs + v];
            }
        }
        rc = fplus_tri_solve_plan(num->rplan, n, f->Lp, f->Li, f->Lx, f->Up, f->Ui, f->Ux, w.data(), m, n, trans, num->is_complex);
        if (rc != FPLUS_OK) break;
        for (int64_t r = 0; r < m; r++) {
            double* xb = x + (size_t)((r0 + r) * ldb) * vs;
            const double* ww = w.data() + (size_t)(r * n) * vs;
            for (int64_t k2 = 0; k2 < n; k2++) {
                const int64_t dst = trans == 0 ? q[k2] : p[k2];
                const double d = trans == 0 ? Dc[dst] : Dr[dst];
                for (int v = 0; v < vs; v++) xb[dst * vs + v] = d * ww[k2 * vs + v];
            }
        }
    }
    if (rc == FPLUS_ERR_SINGULAR) errf(err, errlen, "fplus_solve: zero or missing diagonal in U");
    else if (rc != FPLUS_OK) errf(err, errlen, "fplus_solve: error %d", rc);
    return rc;
}

int fplus_solve(const fplus_numeric* num, void* b, int64_t nrhs, int64_t ldb, int trans, char* err, size_t errlen)
{
    const int64_t n = num->n;
    if (!num->valid) { errf(err, errlen, "the last refactorization of this factorization failed: factor again"); return FPLUS_ERR_BADARG; }
    if (trans < 0 || trans > 2 || nrhs < 0 || ldb < n) { errf(err, errlen, "fplus_solve: bad trans, nrhs or ldb"); return FPLUS_ERR_BADARG; }
    if (nrhs == 0 || n == 0) return FPLUS_OK;
    if (num->row) return solve_row(num, b, nrhs, ldb, trans, err, errlen);
    const int vs = num->F.vs;
    const int tr = vs == 1 ? (trans ? 1 : 0) : trans;
    try {
        hipStream_t s = devctx().stream;
        std::lock_guard<std::mutex> g(num->solve_mu);
        num->d_b.reserve((size_t)(n * nrhs * vs));
        for (int64_t r = 0; r < nrhs; r++)
            num->d_b.upload((const double*)b + (size_t)(r * ldb) * vs, (size_t)n * vs, s, (size_t)(r * n) * vs);
        solve_on_device(num, num->d_b.p, nrhs, n, tr, s);
        for (int64_t r = 0; r < nrhs; r++)
            num->d_b.download((double*)b + (size_t)(r * ldb) * vs, (size_t)n * vs, s, (size_t)(r * n) * vs);
        dev_sync(s);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    }
    return FPLUS_OK;
}

const fplus_factors* fplus_numeric_factors(const fplus_numeric* num)
{
    std::lock_guard<std::mutex> g(num->mu);
    if (!num->exported) {
        try {
            export_factors(num->F, num->is_complex, &num->f);
            num->exported = true;
        } catch (...) {
            return nullptr;
        }
    }
    return &num->f;
}
const int64_t* fplus_numeric_p(const fplus_numeric* num) { return num->p.data(); }
const int64_t* fplus_numeric_q(const fplus_numeric* num) { return num->q.data(); }
const double* fplus_numeric_Dr(const fplus_numeric* num) { return num->Dr.data(); }
const double* fplus_numeric_Dc(const fplus_numeric* num) { return num->Dc.data(); }

}  // extern "C"

namespace fpa {
int ruiz_device_entry(int64_t n, int64_t nnz, const int32_t* d_erow, const int32_t* d_ecol, double* d_x, int vs, int max_iter,
                      double tol, double* d_Dr, double* d_Dc, int verbose, int* iters, char* err, size_t errlen)
{
    return ruiz_device(n, nnz, d_erow, d_ecol, d_x, vs, max_iter, tol, d_Dr, d_Dc, verbose, iters, err, errlen);
}
void export_factors_entry(const DevFactor& F, int is_complex, fplus_factors* f) { export_factors(F, is_complex, f); }
}  // namespace fpa

#include "fplus_amd.h"

extern "C" int fplus_amd_numeric_stats(const fplus_numeric* num, fplus_amd_stats* st)
{
    if (!num || !st || num->row) return FPLUS_ERR_BADARG;
    const DevFactor& F = num->F;
    std::memset(st, 0, sizeof(*st));
    for (int64_t s = 0; s < F.nsn; s++) {
        const int64_t w = F.sn_first[s + 1] - F.sn_first[s], nr = (int64_t)F.sn_rows[s].size();
        st->nnzL += w * nr - w * (w - 1) / 2;
    }
    st->nnzU = F.nnzU;
    st->nsupernodes = F.nsn;
    st->nblocks = (int64_t)F.blocks.size();
    st->panels = F.panels;
    st->levels = F.levels;
    st->flops = F.flops;
    st->gemm_flops = F.gemm_flops;
    st->gemm_bytes = F.gemm_bytes;
    st->t_ruiz = num->t_ruiz;
    st->t_factor = num->t_factor;
    st->t_host = F.t_host;
    st->t_wait = F.t_wait;
    st->t_total = num->t_total;
    st->ruiz_iters = num->ruiz_iters;
    st->shared_levels = F.shared_levels;
    st->pivot_ties = F.pivot_ties;
    st->max_block_w = F.max_block_w;
    st->coop_retried = F.coop_retried ? 1 : 0;
    st->precision = F.precision;
    st->refactor_replaced = num->refactor_replaced;
    st->refactor_min_ratio = num->refactor_min_ratio;
    st->replay_bytes = F.replay ? (double)F.replay->device_bytes() : 0.0;
    if (F.precision == 1) {
        st->f32_blocks = st->nblocks;
        st->f32_flop_share = 1.0;
    } else if (F.precision == 2 && F.d_blk_f32.p && !F.blocks.empty()) {
        // the selector's marks, from the device
        std::vector<uint8_t> mk(F.blocks.size());
        try {
            F.d_blk_f32.download(mk.data(), mk.size(), devctx().stream);
            dev_sync(devctx().stream);
        } catch (const GpuError&) {
            return FPLUS_ERR_GPU;
        }
        double tot = 0, f32 = 0;
        for (size_t b = 0; b < mk.size(); b++) {
            const double fl = b < F.blk_upd_flops.size() ? F.blk_upd_flops[b] : 0.0;
            tot += fl;
            if (mk[b]) {
                st->f32_blocks++;
                f32 += fl;
            }
        }
        st->f32_flop_share = tot > 0 ? f32 / tot : 0.0;
    }
    return FPLUS_OK;
}

extern "C" void fplus_amd_default_options(fplus_amd_options* o)
{
    std::memset(o, 0, sizeof(*o));
}

extern "C" int fplus_amd_factor_ex(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* Ax, int ax_on_device,
                                   int is_complex, const fplus_options* opt, const fplus_amd_options* aopt, fplus_numeric** num,
                                   char* err, size_t errlen)
{
    return factor_entry(sym, Ap, Ai, Ax, ax_on_device != 0, is_complex, opt, aopt, num, err, errlen);
}

extern "C" int fplus_amd_refactor(fplus_numeric* num, const void* Ax, int ax_on_device, int precision, int static_pivots, char* err,
                                  size_t errlen)
{
    if (!num) { errf(err, errlen, "no factorization"); return FPLUS_ERR_BADARG; }
    if (num->row) return row_handle_error(err, errlen);
    if (!num->F.replay) {
        errf(err, errlen, "this factorization was made without fplus_amd_options.refactor = 1: nothing to replay");
        return FPLUS_ERR_BADARG;
    }
    if (precision < 0 || precision > 4 || static_pivots < 0 || static_pivots > 1) {
        errf(err, errlen, "precision must be 0, 1, 2, 3 or 4 and static_pivots 0 or 1");
        return FPLUS_ERR_BADARG;
    }
    if (precision == 2 && !num->aopt.selector) {
        errf(err, errlen, "precision 2 needs the selector given with the factorization (fplus_amd_options.selector)");
        return FPLUS_ERR_BADARG;
    }
    const PatternCache& c = *num->cache;
    const int vs = num->F.vs;
    const int64_t n = num->n, nnz = c.nnz;
    if (!ax_on_device) {
        const int rc = check_finite(nnz, Ax, num->is_complex, err, errlen);
        if (rc != FPLUS_OK) return rc;
    }
    try {
        std::lock_guard<std::mutex> g1(num->solve_mu);
        std::lock_guard<std::mutex> g2(num->mu);
        const double t0 = now_s();
        hipStream_t s = devctx().stream;
        // the new values' equilibrated, analysed form goes into the factorization's own buffer (d_Abar, which the
        // refined solves read), so that a captured replay finds its input at the same address every time
        DevBuf<double> d_Ax;
        if (num->d_Abar.n < (size_t)(nnz > 0 ? nnz : 1) * vs) num->d_Abar.alloc((size_t)(nnz > 0 ? nnz : 1) * vs);
        DevBuf<double>& d_Bx = num->d_Abar;
        const double* src = static_cast<const double*>(Ax);
        if (!ax_on_device) {
            d_Ax.alloc((size_t)(nnz > 0 ? nnz : 1) * vs);
            d_Ax.upload(static_cast<const double*>(Ax), (size_t)nnz * vs, s);
            src = d_Ax.p;
        }
        k::gather_vals(vs, nnz, c.d_src.p, src, d_Bx.p, s);
        num->valid = false;
        int rc = FPLUS_OK;
        if (num->opt.equilibrate)
            rc = ruiz_device(n, nnz, c.d_Bi.p, c.d_ecol.p, d_Bx.p, vs, num->opt.ruiz_max_iter, num->opt.ruiz_tol, num->d_Drp.p,
                             num->d_Dcp.p, num->opt.verbose, &num->ruiz_iters, err, errlen);
        if (rc != FPLUS_OK) return rc;
        const double t1 = now_s();
        int64_t replaced = 0;
        double minr = 0;
        const SelectorArgs sa = selector_args(precision == 2 ? num->aopt.selector : nullptr, num->d_Drp.p, num->d_Dcp.p);
        rc = sn_refactor_device(num->F, c.d_Bp.p, c.d_Bi.p, d_Bx.p, num->opt.pivot_rel, static_pivots, precision, sa, &replaced,
                                &minr, err, errlen);
        num->refactor_replaced = replaced;
        num->refactor_min_ratio = minr;
        if (rc != FPLUS_OK) return rc;
        std::vector<double> Drp(n), Dcp(n);
        num->d_Drp.download(Drp.data(), n, s);
        num->d_Dcp.download(Dcp.data(), n, s);
        for (int64_t k2 = 0; k2 < n; k2++) {
            num->Dr[num->rp[k2]] = Drp[k2];
            num->Dc[num->cp[k2]] = Dcp[k2];
        }
        num->d_Dr.upload(num->Dr.data(), n, s);
        num->d_Dc.upload(num->Dc.data(), n, s);
        if (num->n_ujobs > 0) k::blk_uinv(vs, num->n_ujobs, num->d_ujobs.p, num->F.d_blocks.p, num->uinv_flops, s);
        dev_sync(s);
        if (num->exported) {
            fplus_factors_free(&num->f);
            num->exported = false;
        }
        num->valid = true;
        num->t_ruiz = t1 - t0;
        num->t_factor = now_s() - t1;
        num->t_total = now_s() - t0;
        if (num->opt.verbose >= 1)
            printf("[FPLUS] refactorization in the recorded pivot order: %.3fs (Ruiz %.3fs), %s updates, %ld pivots replaced, "
                   "smallest |pivot| / colmax %.3e\n", num->t_total, num->t_ruiz,
                   precision == 0 ? "FP64" : (precision == 1 ? "FP32" : (precision == 2 ? "per-block (selector)" : (precision == 3 ? "FP64 (logged)" : "BF16 hi/lo"))),
                   (long)replaced, minr);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    } catch (const std::bad_alloc&) {
        errf(err, errlen, "out of memory");
        return FPLUS_ERR_NOMEM;
    } catch (const std::exception& e) {
        errf(err, errlen, "%s", e.what());
        return FPLUS_ERR_INTERNAL;
    }
    return FPLUS_OK;
}

extern "C" int fplus_amd_tune(const char* name, int64_t value)
{
    if (!name) return FPLUS_ERR_BADARG;
    if (std::strcmp(name, "panel_lu_max_workgroups") == 0) {
        if (value < 0 || value > 512) return FPLUS_ERR_BADARG;
        k::set_getf2_coop_max_wg((int)value);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "solve_noalloc_limit_bytes") == 0) {
        if (value < 0 || !set_noalloc_limit((size_t)value)) return FPLUS_ERR_BADARG;
        return FPLUS_OK;
    }
    if (std::strcmp(name, "update_tile_policy") == 0) {
        if (value < 0 || value > 4 || !set_update_tile_policy((int)value)) return FPLUS_ERR_BADARG;
        return FPLUS_OK;
    }
    if (std::strcmp(name, "refactor_graph") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_refactor_graph(value != 0);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "refactor_pipeline") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_refactor_pipeline(value != 0);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "host_threads") == 0) {
        if (!set_host_threads((int)std::max<int64_t>(std::min<int64_t>(value, 65), -1))) return FPLUS_ERR_BADARG;
        return FPLUS_OK;
    }
    if (std::strcmp(name, "deterministic_updates") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_deterministic_updates(value != 0);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "update_atomic_rows") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_update_atomic_rows(value != 0);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "solve_dataflow") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_solve_dataflow(value != 0);
        return FPLUS_OK;
    }
    if (std::strcmp(name, "xcd_tile_order") == 0) {
        if (value != 0 && value != 1) return FPLUS_ERR_BADARG;
        set_xcd_tile_order(value != 0);
        return FPLUS_OK;
    }
    return FPLUS_ERR_BADARG;
}

extern "C" {
int fplus_amd_factor_device(const fplus_symbolic* sym, const int64_t* Ap, const int64_t* Ai, const void* d_Ax, int is_complex,
                            const fplus_options* opt, fplus_numeric** num, char* err, size_t errlen)
{
    return factor_entry(sym, Ap, Ai, d_Ax, true, is_complex, opt, nullptr, num, err, errlen);
}

int fplus_amd_solve_device(const fplus_numeric* num, void* d_b, int64_t nrhs, int64_t ldb, int trans, char* err, size_t errlen)
{
    if (num->row) return row_handle_error(err, errlen);
    const int64_t n = num->n;
    if (!num->valid) { errf(err, errlen, "the last refactorization of this factorization failed: factor again"); return FPLUS_ERR_BADARG; }
    if (trans < 0 || trans > 2 || nrhs < 0 || ldb < n) { errf(err, errlen, "fplus_amd_solve_device: bad trans, nrhs or ldb"); return FPLUS_ERR_BADARG; }
    if (nrhs == 0 || n == 0) return FPLUS_OK;
    const int vs = num->F.vs;
    const int tr = vs == 1 ? (trans ? 1 : 0) : trans;
    try {
        hipStream_t s = devctx().stream;
        std::lock_guard<std::mutex> g(num->solve_mu);
        solve_on_device(num, static_cast<double*>(d_b), nrhs, ldb, tr, s);
        dev_sync(s);
    } catch (const GpuError& e) {
        return gpu_error(e, err, errlen);
    }
    return FPLUS_OK;
}

int fplus_amd_dev_alloc(size_t bytes, void** ptr)
{
    try {
        devctx();
        return hipMalloc(ptr, bytes > 0 ? bytes : 1) == hipSuccess ? FPLUS_OK : FPLUS_ERR_NOMEM;
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    }
}
int fplus_amd_dev_free(void* ptr) { return hipFree(ptr) == hipSuccess ? FPLUS_OK : FPLUS_ERR_GPU; }
int fplus_amd_dev_copy(void* dst, const void* src, size_t bytes, int kind)
{
    try {
        hipStream_t s = devctx().stream;
        if (kind == 0) copy_h2d(dst, src, bytes, s);
        else if (kind == 1) copy_d2h(dst, src, bytes, s);
        else if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) return FPLUS_ERR_GPU;
        return hipStreamSynchronize(s) == hipSuccess ? FPLUS_OK : FPLUS_ERR_GPU;
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    }
}
int fplus_amd_thread_stream(int on)
{
    try {
        set_thread_stream(on != 0);
        return FPLUS_OK;
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    }
}
int fplus_amd_dev_sync(void)
{
    try {
        return hipStreamSynchronize(devctx().stream) == hipSuccess ? FPLUS_OK : FPLUS_ERR_GPU;
    } catch (const GpuError&) {
        return FPLUS_ERR_GPU;
    }
}
int fplus_amd_profile(int on)
{
    KernelProf& p = kprof();
    std::lock_guard<std::mutex> g(p.mu);
    p.collect();
    if (on) p.sums.clear();
    p.on = on != 0;
    return FPLUS_OK;
}
int fplus_amd_profile_get(const char* kernel, double* ms, int64_t* launches, double* work)
{
    KernelProf& p = kprof();
    std::lock_guard<std::mutex> g(p.mu);
    p.collect();
    auto it = p.sums.find(kernel);
    if (it == p.sums.end()) { *ms = 0; *launches = 0; *work = 0; return FPLUS_OK; }
    *ms = it->second.ms; *launches = it->second.launches; *work = it->second.work;
    return FPLUS_OK;
}
}  // extern "C"
