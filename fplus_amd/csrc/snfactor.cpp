// snfactor.cpp -- the supernodal left-looking LU with partial pivoting: the host runs the symbolic half of the
// reference's panel kernel (reference src/factor_sn.c:97-210 search / relaxed joins, src/factor_sn_panel.inc:30-390
// the three phases, src/factor_sn.c:609-808 the driver), the device every arithmetic operation.
//
// Per panel of P columns:
//   1. host: each column's depth-first search over the supernodes complete before the panel (its reached segments and
//      non-pivotal rows), the union search giving a topological order, the update tasks and their dependency levels,
//      the rows the panel touches (the workspace rows);
//   2. device: the panel's columns of A scattered into the workspace, every earlier supernode applied block by block
//      (TRSM of the reached segment, then GEMM + relative-index scatter), one launch pair per level;
//   3. device: dense partial-pivot LU of the workspace rows not yet pivotal; the P pivot positions come back;
//   4. host: per column, the within-panel search, the U pattern, the relaxed-join decision and the new supernodes,
//      pruning -- the reference's logic, fed with the device's pivots;
//   5. device: the new supernode blocks and the U values copied out of the workspace.
// The partition, the patterns and the pivot order are those of the reference (up to pivot candidates within
// round-off of each other); the values agree to round-off (another summation order).
#include "snfactor.h"

#include "host_common.h"
#include "host_pool.h"
#include "prof.h"
#include "selector_dev.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <exception>
#include <functional>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <thread>

namespace fpa {

static std::atomic<bool> g_xcd_order{true};
bool xcd_tile_order() { return g_xcd_order.load(std::memory_order_relaxed); }
void set_xcd_tile_order(bool on) { g_xcd_order.store(on, std::memory_order_relaxed); }
static std::atomic<int> g_tile_policy{0};
int update_tile_policy() { return g_tile_policy.load(std::memory_order_relaxed); }
bool set_update_tile_policy(int p)
{
    if (p < 0 || p > 4) return false;
    g_tile_policy.store(p, std::memory_order_relaxed);
    return true;
}
static std::atomic<int> g_host_threads{0};
int host_threads()
{
    const int n = g_host_threads.load(std::memory_order_relaxed);
    return n > 0 ? n : 15;
}
bool set_host_threads(int n)
{
    if (n < 0 || n > 64) return false;
    g_host_threads.store(n, std::memory_order_relaxed);
    return true;
}
static std::atomic<bool> g_atomic_rows{true};
bool update_atomic_rows() { return g_atomic_rows.load(std::memory_order_relaxed); }
void set_update_atomic_rows(bool on) { g_atomic_rows.store(on, std::memory_order_relaxed); }
static std::atomic<bool> g_det_updates{false};
bool deterministic_updates() { return g_det_updates.load(std::memory_order_relaxed); }
void set_deterministic_updates(bool on) { g_det_updates.store(on, std::memory_order_relaxed); }
static std::atomic<bool> g_refactor_graph{false};
bool refactor_graph() { return g_refactor_graph.load(std::memory_order_relaxed); }
static std::atomic<bool> g_refactor_pipeline{true};
bool refactor_pipeline() { return g_refactor_pipeline.load(std::memory_order_relaxed); }
void set_refactor_pipeline(bool on) { g_refactor_pipeline.store(on, std::memory_order_relaxed); }
void set_refactor_graph(bool on) { g_refactor_graph.store(on, std::memory_order_relaxed); }

namespace {

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct HSn {
    int64_t c0 = 0, w = 0;
    std::vector<int64_t> ri;      // rows: the pivot rows of its columns first (column order), then the others
    std::vector<int64_t> srch;    // closed: its rows past the pivot rows, the search list
    int64_t srch_end = -1;        // end of the pivotal part of srch once pruned
    bool closed = false;
    double nzero = 0;             // explicit zeros from relaxed joins
    std::vector<int32_t> blocks;  // device blocks in column order
    int64_t nr() const { return (int64_t)ri.size(); }
};

// device arena of fixed-address chunks
template <class T> struct Arena {
    std::vector<DevBuf<T>>* chunks;
    size_t used = 0, cap = 0;
    size_t min_chunk;
    T* take(size_t cnt)
    {
        if (chunks->empty() || used + cnt > cap) {
            const size_t sz = std::max(min_chunk, cnt);
            chunks->emplace_back(sz);
            used = 0;
            cap = sz;
        }
        T* p = chunks->back().p + used;
        used += (cnt + 7) & ~(size_t)7;
        return p;
    }
};

// one helper thread running one job at a time: a panel's output (blocks and U out of the workspace) while the
// next panel's search runs.  wait() returns when the job is done and rethrows what it threw.
class SideWorker {
public:
    SideWorker() : th_([this] { loop(); }) {}
    ~SideWorker()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        th_.join();
    }
    SideWorker(const SideWorker&) = delete;
    SideWorker& operator=(const SideWorker&) = delete;
    void submit(std::function<void()> fn)
    {
        std::lock_guard<std::mutex> lk(mu_);
        job_ = std::move(fn);
        busy_ = true;
        cv_.notify_all();
    }
    void wait()
    {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return !busy_; });
        if (err_) {
            std::exception_ptr e = err_;
            err_ = nullptr;
            std::rethrow_exception(e);
        }
    }

private:
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            cv_.wait(lk, [this] { return stop_ || job_; });
            if (!job_) return;
            std::function<void()> fn = std::move(job_);
            job_ = nullptr;
            lk.unlock();
            std::exception_ptr e;
            try {
                fn();
            } catch (...) {
                e = std::current_exception();
            }
            lk.lock();
            err_ = e;
            busy_ = false;
            done_cv_.notify_all();
        }
    }
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void()> job_;
    bool busy_ = false, stop_ = false;
    std::exception_ptr err_;
    std::thread th_;
};

class Driver {
public:
    Driver(int64_t n, const int64_t* Bp, const int64_t* Bi, const int64_t* d_Bp, const int32_t* d_Bi, const double* d_Bx,
           int vs, const SnOptions& opt, DevFactor& F, char* err, size_t errlen)
        : n_(n), Bp_(Bp), Bi_(Bi), d_Bp_(d_Bp), d_Bi_(d_Bi), d_Bx_(d_Bx), vs_(vs), opt_(opt), F_(F), err_(err), errlen_(errlen)
    {
    }
    int run();

private:
    // the search (reference src/factor_sn.c:97-154)
    const int64_t* srch_rows(int64_t s, int64_t open, int64_t* end) const
    {
        const HSn& S = sn_[s];
        if (s == open) { *end = S.nr() - S.w; return S.ri.data() + S.w; }
        *end = S.srch_end >= 0 ? S.srch_end : S.nr() - S.w;
        return S.srch.data();
    }
    // the search's marks and stacks (one set per thread of the parallel per-column search), 32-bit: rows, columns,
    // supernodes and marks (below 3 n + P) all fit, and half the footprint keeps more of the marks in cache
    struct SWork {
        std::vector<int32_t> rmark, smark, seg0, topo, cand, stack, pos;
        void init(size_t nn)
        {
            rmark.assign(nn, 0); smark.assign(nn, 0); seg0.assign(nn, 0); topo.assign(nn, 0);
            cand.assign(nn, 0); stack.assign(nn, 0); pos.assign(nn, 0);
        }
    };
    // the search from supernode `root` entered at column `col`, its rows read from position pos0 of its search list;
    // a row not pivotal is a new candidate unless marked or in_list(row) (the column's candidates known before)
    template <class InList>
    int64_t dfs_in(SWork& w, int64_t root, int64_t col, int64_t top, int64_t stamp, int64_t open, int64_t* ncand,
                   int64_t pos0, const InList& in_list) const
    {
        int64_t sp = 0;
        w.stack[0] = (int32_t)root; w.smark[root] = (int32_t)stamp; w.seg0[root] = (int32_t)col; w.pos[root] = (int32_t)pos0;
        while (sp >= 0) {
            const int64_t s = w.stack[sp];
            int64_t end;
            const int64_t* rows = srch_rows(s, open, &end);
            int64_t q = w.pos[s], t = -1;
            for (; q < end; q++) {
                const int64_t i = rows[q], k = pinv_[i];
                if (k < 0) {
                    if (w.rmark[i] != stamp && !in_list(i)) { w.rmark[i] = (int32_t)stamp; w.cand[(*ncand)++] = (int32_t)i; }
                    continue;
                }
                const int64_t u = col_sn_[k];
                if (w.smark[u] == stamp) { if (k < w.seg0[u]) w.seg0[u] = (int32_t)k; continue; }
                t = u; w.smark[u] = (int32_t)stamp; w.seg0[u] = (int32_t)k;
                break;
            }
            if (t >= 0) { w.pos[s] = (int32_t)(q + 1); w.pos[t] = 0; w.stack[++sp] = (int32_t)t; }
            else { sp--; w.topo[--top] = (int32_t)s; }
        }
        return top;
    }
    int64_t dfs(SWork& w, int64_t root, int64_t col, int64_t top, int64_t stamp, int64_t open, int64_t* ncand) const
    {
        return dfs_in(w, root, col, top, stamp, open, ncand, 0, [](int64_t) { return false; });
    }
    bool relax_ok(const HSn& S, int64_t n_miss, int64_t extra) const
    {
        const int64_t mm = n_miss + extra;
        if (!relax_auto_) return mm <= relax_;
        if (mm > rz_cap_) return false;
        const int64_t w = S.w + 1, r = S.nr() + n_miss;
        if (w <= rz_w_[0]) return true;
        const double z = S.nzero + (double)n_miss * (double)S.w + (double)extra, tot = (double)w * (double)r;
        const double zmax = w <= rz_w_[1] ? rz_z_[0] : (w <= rz_w_[2] ? rz_z_[1] : rz_z_[2]);
        return z <= zmax * tot;
    }
    void prune(int64_t s, int64_t pivrow)
    {
        HSn& S = sn_[s];
        if (!S.closed || S.srch_end >= 0) return;
        const int64_t len = (int64_t)S.srch.size();
        int64_t* rows = S.srch.data();
        bool hit = false;
        for (int64_t r = 0; r < len; r++) if (rows[r] == pivrow) { hit = true; break; }
        if (!hit) return;
        int64_t a = 0, b = len;
        while (a < b) {
            if (pinv_[rows[a]] >= 0) { a++; continue; }
            b--;
            std::swap(rows[a], rows[b]);
        }
        S.srch_end = b;
    }
    int panel(int64_t j0, int64_t j1);
    // the panel's output, on the helper thread; its inputs are moved into out_ by panel()
    int panel_out();
    struct OutState {
        int64_t j0 = 0, j1 = 0, Rl = 0, ldw = 0, u_begin = 0;
        double* Wl = nullptr;
        std::vector<int32_t> univ, lrows;
        std::vector<std::pair<int64_t, int64_t>> touched;
        std::vector<URun> pruns;
    } out_;
    int out_rc_ = 0;
    bool out_pending_ = false;
    // wait for the previous panel's output; its status (or its exception)
    int join_out()
    {
        if (!out_pending_) return FPLUS_OK;
        out_pending_ = false;
        side_.wait();
        return out_rc_;
    }
    const BlockDesc* d_blocks_p() const { return F_.d_blocks.p; }
    void d_blocks_upload()
    {
        const size_t nb = F_.blocks.size();
        if (nb > blocks_up_) up_.upload(F_.d_blocks.p + blocks_up_, F_.blocks.data() + blocks_up_, nb - blocks_up_, s_);
        blocks_up_ = nb;
    }
    size_t blocks_up_ = 0;
    void add_block(int64_t s, int64_t b0, int64_t b1, int64_t j0, const std::vector<int32_t>& lpos,
                   std::vector<ExtractCol>& ex, std::vector<int32_t>& expos);

    int64_t n_;
    const int64_t *Bp_, *Bi_, *d_Bp_;
    const int32_t* d_Bi_;
    const double* d_Bx_;
    int vs_;
    SnOptions opt_;
    DevFactor& F_;
    char* err_;
    size_t errlen_;

    std::vector<HSn> sn_;
    std::vector<int64_t> col_sn_, pinv_, p_, miss_;
    std::vector<int32_t> bpos_;   // a row's position in the open supernode's rows, -1 outside it
    std::vector<int64_t> piv_c_;   // phase 3: a column's candidates pivotal in the panel
    std::vector<std::vector<int64_t>> c_s_, c_kf_;   // phase 1's results per panel column
    std::vector<std::vector<int32_t>> c_np_;         // and its candidates
    std::vector<std::vector<int32_t>> ch_rows_;   // per chunk of columns: the rows not yet pivotal, first appearance
    std::vector<int32_t> reader_;   // per workspace U row: the update task whose segment holds it
    // phase 3's view of the columns' candidates: their lower workspace positions (as at the panel's start), the same as
    // a bitset per column, and per lower position the panel column that pivoted it
    std::vector<std::vector<int32_t>> c_lp_;
    std::vector<uint64_t> cbits_;
    std::vector<int32_t> pivcol_;
    std::vector<std::vector<int32_t>> succ_;   // per update task: the later tasks whose segments its rows below reach
    std::vector<int32_t> det_next_, det_stamp_;   // deterministic levels: per workspace row, the next free sub-level
    std::vector<std::vector<int64_t>> tmark_;   // one mark array per thread of the pool, indexed by task
    // per thread of the pool, indexed by workspace row: the (panel, level) that last wrote the row and the row tile that
    // wrote it (the rows two tasks of one level write)
    std::vector<std::vector<int64_t>> wstamp_, wwho_;
    SWork ws0_;                     // the sequential phases' marks
    std::vector<SWork> wsT_;        // one per thread of the parallel search
    HostPool pool_{host_threads()};   // (15 by default: one core left for the helper thread)
    int64_t relax_ = 0, rz_w_[3] = {4, 16, 48}, rz_cap_ = 32;
    double rz_z_[3] = {0.8, 0.1, 0.05};
    bool relax_auto_ = false;
    int64_t open_ = -1;
    std::vector<int32_t> hrowpos_;    // workspace position of a row, -1 outside the panel
    std::vector<int32_t> lposmap_;    // final lower-section position of a row after the panel LU
    std::vector<int64_t> ustamp_, uidx_;
    std::vector<int64_t> Up_, Urp_;
    std::vector<int32_t> Ur_row_, Ur_len_;   // U as runs of consecutive rows
    // device workspace
    DevBuf<double> d_W_;
    DevBuf<int32_t> d_rowmap_, d_univ_, d_cols_, d_pos_;
    DevBuf<double> d_colmax_;
    DevBuf<UpdTask> d_tasks_;
    DevBuf<int2> d_ijobs_;           // the new blocks' inverse forms: (block, column) jobs and the product tiles
    DevBuf<GemmTile> d_itiles_;
    std::vector<int32_t> new_blocks_;   // the blocks made by the current panel
    DevBuf<UpdTile> d_tiles_;
    DevBuf<double> d_X_;
    DevBuf<PanelOut> d_pout_;        // the panel LU's status and pivots
    DevBuf<double> d_coop_;          // the cooperative panel LU's candidate slots
    DevBuf<int32_t> d_coop_cnt_;     // and its barrier counter
    int64_t coop_counted_ = 0;       // the arrivals the counter holds
    HostBuf<PanelOut> h_pout_;       // per sub-panel: the status and its pivots, written by the panel LU kernel
    PanelOut* h_pout_dev_ = nullptr; // the device's address of h_pout_
    EventList sp_ev_;                // the end of each sub-panel's copies
    DevBuf<URun> d_runs_;
    DevBuf<ExtractCol> d_ex_;
    DevBuf<int32_t> d_newblk_;       // the panel's new blocks (features, selector)
    DevBuf<int32_t> d_lrows_;        // the row at each lower workspace position after the panel LU
    UploadStage up_;                 // every upload of a panel goes through page-locked staging (one area per panel,
                                     // two in turn): the host goes on while the copies and the kernels run
    Arena<double> vals_{nullptr, 0, 0, 0};
    std::vector<DevBuf<double>> iv_chunks_;   // the blocks' inverse forms: needed only while factoring, freed with the driver
    Arena<double> ivals_{nullptr, 0, 0, 0};
    Arena<double> linv_{nullptr, 0, 0, 0};
    Arena<double> uinv_{nullptr, 0, 0, 0};
    Arena<int32_t> rows_{nullptr, 0, 0, 0};
    hipStream_t s_ = nullptr;
    Replay* rec_ = nullptr;          // recording what a refactorization replays (F_.replay), or null
    ReplayPanel rec_panel_{};        // the panel being recorded
    SideWorker side_;   // last: destroyed (joined) first, before the state its job uses
};

void Driver::add_block(int64_t s, int64_t b0, int64_t b1, int64_t j0, const std::vector<int32_t>& lpos,
                       std::vector<ExtractCol>& ex, std::vector<int32_t>& expos)
{
    HSn& S = sn_[s];
    const int64_t r0 = b0 - S.c0, nrows = S.nr() - r0, w = b1 - b0;
    BlockDesc B;
    B.vals = vals_.take((size_t)(nrows * w * vs_));
    B.ivals = ivals_.take((size_t)(nrows * w * vs_));
    B.linv = linv_.take((size_t)(w * w * vs_));
    B.uinv = uinv_.take((size_t)(w * w * vs_));
    B.rows = rows_.take((size_t)nrows);
    B.nrows = (int32_t)nrows;
    B.w = (int32_t)w;
    B.c0 = (int32_t)b0;
    B.sn = (int32_t)s;
    F_.max_block_w = std::max<int64_t>(F_.max_block_w, w);
    std::vector<int32_t> hr(nrows);
    const int64_t pos_off = (int64_t)expos.size();
    for (int64_t r = 0; r < nrows; r++) {
        const int64_t row = S.ri[r0 + r];
        hr[r] = (int32_t)row;
        if (lpos[row] < 0) throw std::logic_error("internal: a block row outside the panel's lower rows");
        expos.push_back(lpos[row]);
    }
    for (int64_t c = 0; c < w; c++) {   // the rows go to the device with the first column (the row at each position)
        ExtractCol e;
        e.dst = B.vals + (size_t)c * nrows * vs_;
        e.rdst = c == 0 ? B.rows : nullptr;
        e.pos_off = pos_off;
        e.nrows = (int32_t)nrows;
        e.ci = (int32_t)(b0 - j0 + c);
        ex.push_back(e);
    }
    const int32_t id = (int32_t)F_.blocks.size();
    new_blocks_.push_back(id);
    F_.blocks.push_back(B);
    F_.block_rows.push_back(std::move(hr));
    S.blocks.push_back(id);
}

int Driver::panel(int64_t j0, int64_t j1)
{
    const int64_t np = j1 - j0, n = n_;
    const int64_t open0 = open_, we_open = open0 >= 0 ? sn_[open0].w : 0, nr_open = open0 >= 0 ? sn_[open0].nr() : 0;
    const double th0 = now_s();
    // ---- phase 1: per-column search over the supernodes complete before j0 (src/factor_sn_panel.inc:34-58); the
    // columns' searches are independent (they read the structure as it stood at j0): threads take columns in turn,
    // each with its own marks, and the results are joined in column order
    std::vector<int64_t> ext_s, ext_kf, ext_off(np + 1);
    std::vector<std::vector<int32_t>>& c_np = c_np_;   // per column: its candidates (rows not pivotal), search order
    if ((int64_t)c_np.size() < np) c_np.resize(np);
    {
        std::vector<std::vector<int64_t>>& c_s = c_s_;
        std::vector<std::vector<int64_t>>& c_kf = c_kf_;
        if ((int64_t)c_s.size() < np) { c_s.resize(np); c_kf.resize(np); }
        auto search = [&](int64_t ci, SWork& w) {
            const int64_t c = j0 + ci, stamp = c + 1;
            int64_t top = n, ncand = 0;
            for (int64_t q = Bp_[c]; q < Bp_[c + 1]; q++) {
                const int64_t i = Bi_[q], k = pinv_[i];
                if (k < 0) {
                    if (w.rmark[i] != stamp) { w.rmark[i] = (int32_t)stamp; w.cand[ncand++] = (int32_t)i; }
                    continue;
                }
                const int64_t sn = col_sn_[k];
                if (w.smark[sn] == stamp) { if (k < w.seg0[sn]) w.seg0[sn] = (int32_t)k; continue; }
                top = dfs(w, sn, k, top, stamp, open_, &ncand);
            }
            c_s[ci].resize(n - top);
            c_kf[ci].resize(n - top);
            for (int64_t t = top; t < n; t++) { c_s[ci][t - top] = w.topo[t]; c_kf[ci][t - top] = w.seg0[w.topo[t]]; }
            c_np[ci].resize(ncand);
            for (int64_t q = 0; q < ncand; q++) c_np[ci][q] = (int32_t)w.cand[q];
        };
        pool_.run(np, [&](int64_t ci, int tid) { search(ci, wsT_[tid]); });
        for (int64_t ci = 0; ci < np; ci++) {
            ext_off[ci] = (int64_t)ext_s.size();
            ext_s.insert(ext_s.end(), c_s[ci].begin(), c_s[ci].end());
            ext_kf.insert(ext_kf.end(), c_kf[ci].begin(), c_kf[ci].end());
        }
    }
    ext_off[np] = (int64_t)ext_s.size();
    const double th_p1 = now_s();
    F_.t_p1 += th_p1 - th0;
    // ---- the union of the reached supernodes, in a topological order of the updates (the reference takes one from a
    // search over the union, src/factor_sn_panel.inc:61-70): increasing supernode index is one -- a block's rows below
    // its pivot rows were not pivotal when its supernode opened, so the supernodes whose pivot rows they become opened
    // later, and a supernode's blocks follow in column order.  The order only places the U rows in the workspace and
    // orders the tasks; the levels are the same for every topological order.
    std::vector<int64_t> un;
    {
        const int64_t us = 2 * n + 1 + j0;
        for (size_t q = 0; q < ext_s.size(); q++)
            if (ustamp_[ext_s[q]] != us) { ustamp_[ext_s[q]] = us; un.push_back(ext_s[q]); }
        std::sort(un.begin(), un.end());
        for (size_t u = 0; u < un.size(); u++) uidx_[un[u]] = (int64_t)u;
    }
    // ---- the workspace rows: the U rows (pivotal before j0) first, each union supernode's reached rows as one run (the
    // union of its columns' segments [kf, we) is the longest one), then the rows not yet pivotal
    std::vector<int32_t> univ;
    std::vector<int64_t> ukmin(un.size(), INT64_MAX), ubase(un.size(), 0);
    for (size_t q = 0; q < ext_s.size(); q++) {
        const int64_t u = uidx_[ext_s[q]];
        ukmin[u] = std::min(ukmin[u], ext_kf[q]);
    }
    for (size_t u = 0; u < un.size(); u++) {
        ubase[u] = (int64_t)univ.size();
        if (ukmin[u] == INT64_MAX) continue;
        const HSn& S = sn_[un[u]];
        const int64_t we = un[u] == open0 ? we_open : S.w;
        for (int64_t cc = ukmin[u] - S.c0; cc < we; cc++) {
            const int64_t row = S.ri[cc];
            if (hrowpos_[row] >= 0) { errf(err_, errlen_, "internal: a U row in two segments"); return FPLUS_ERR_INTERNAL; }
            hrowpos_[row] = (int32_t)univ.size();
            univ.push_back((int32_t)row);
        }
    }
    const int64_t Ru = (int64_t)univ.size();
    {
        // the rows not yet pivotal, in order of first appearance over the columns' candidate lists: chunks of columns
        // list theirs in parallel (each with its own marks), then the lists are joined in chunk order -- the same order
        // as one pass over all the lists
        // (one chunk per thread: the join reads every chunk's list, most rows are in all of them)
        const double tl0 = now_s();
        const int64_t nch = std::min<int64_t>(np, (int64_t)pool_.size()), per = (np + nch - 1) / nch;
        if ((int64_t)ch_rows_.size() < nch) ch_rows_.resize(nch);
        pool_.run(nch, [&](int64_t c, int tid) {
            SWork& w = wsT_[tid];
            const int64_t stamp = 2 * n + 2 + j0 + c;
            std::vector<int32_t>& out = ch_rows_[c];
            out.clear();
            for (int64_t ci = c * per; ci < std::min(np, (c + 1) * per); ci++)
                for (const int32_t row : c_np[ci])
                    if (hrowpos_[row] < 0 && w.rmark[row] != stamp) { w.rmark[row] = (int32_t)stamp; out.push_back(row); }
        });
        // the join: the first chunk's rows (most of them) placed, the other chunks' lists filtered against them in
        // parallel, then the few rows left joined in chunk order
        for (const int32_t row : ch_rows_[0]) { hrowpos_[row] = (int32_t)univ.size(); univ.push_back(row); }
        if (nch > 1)
            pool_.run(nch - 1, [&](int64_t c, int) {
                std::vector<int32_t>& l = ch_rows_[c + 1];
                size_t k = 0;
                for (const int32_t row : l)
                    if (hrowpos_[row] < 0) l[k++] = row;
                l.resize(k);
            });
        for (int64_t c = 1; c < nch; c++)
            for (const int32_t row : ch_rows_[c])
                if (hrowpos_[row] < 0) { hrowpos_[row] = (int32_t)univ.size(); univ.push_back(row); }
        F_.t_ws_low += now_s() - tl0;
    }
    if (open0 >= 0) {
        const HSn& S = sn_[open0];
        for (int64_t r = S.w; r < S.nr(); r++) {
            const int64_t row = S.ri[r];
            if (hrowpos_[row] < 0) { hrowpos_[row] = (int32_t)univ.size(); univ.push_back((int32_t)row); }
        }
    }
    const int64_t Rl = (int64_t)univ.size() - Ru;
    const double th_ws = now_s();
    F_.t_ws += th_ws - th_p1;
    for (int64_t q = Bp_[j0]; q < Bp_[j1]; q++)   // every entry of the panel's columns lands in the workspace
        if (hrowpos_[Bi_[q]] < 0) { errf(err_, errlen_, "internal: a column entry outside the panel rows"); return FPLUS_ERR_INTERNAL; }
    const int64_t ldw = std::max<int64_t>((Ru + Rl + 1) & ~(int64_t)1, 2);
    F_.rows_upper += Ru;
    F_.rows_lower += Rl;
    F_.rows_lower_max = std::max(F_.rows_lower_max, Rl);
    // the previous panel's output ran meanwhile (it makes the blocks the tasks read, and uses the other staging area):
    // done by now in general
    {
        const int orc = join_out();
        if (orc != FPLUS_OK) return orc;
    }
    up_.begin();
    // ---- update tasks per union supernode and block, with dependency levels: a block's task solves the block's pivot
    // rows the panel's columns reach (read: they wait for every task writing them) and writes the rows below
    const int RT = FPA_ROW_TILE, CTG = FPA_COL_TILE;
    const double fm = vs_ == 1 ? 2.0 : 8.0, es = 8.0 * vs_;
    std::vector<int64_t> ucnt(un.size() + 1, 0);
    for (size_t q = 0; q < ext_s.size(); q++) ucnt[uidx_[ext_s[q]] + 1]++;
    for (size_t u = 0; u < un.size(); u++) ucnt[u + 1] += ucnt[u];
    std::vector<int64_t> ucol(ext_s.size()), ukf(ext_s.size());
    {
        std::vector<int64_t> fill(ucnt.begin(), ucnt.end() - 1);
        for (int64_t ci = 0; ci < np; ci++)
            for (int64_t q = ext_off[ci]; q < ext_off[ci + 1]; q++) {
                const int64_t pos = fill[uidx_[ext_s[q]]]++;
                ucol[pos] = ci;
                ukf[pos] = ext_kf[q];
            }
    }
    struct TaskL { UpdTask t; int32_t level; };   // t.spos0 is set with the levels
    std::vector<TaskL> tl;
    std::vector<int32_t> cols;
    for (size_t u = 0; u < un.size(); u++) {
        const int64_t s = un[u], c0 = ucnt[u], c1 = ucnt[u + 1];
        if (c1 == c0) continue;
        const HSn& S = sn_[s];
        for (int32_t b : S.blocks) {
            const BlockDesc& B = F_.blocks[b];
            const int64_t bb0 = B.c0, bb1 = B.c0 + B.w;
            int64_t kmin = bb1;
            const int64_t col_off = (int64_t)cols.size();
            for (int64_t q = c0; q < c1; q++)
                if (ukf[q] < bb1) { cols.push_back((int32_t)ucol[q]); kmin = std::min(kmin, ukf[q]); }
            const int64_t ncol = (int64_t)cols.size() - col_off;
            if (ncol == 0) continue;
            const int64_t k0 = std::max<int64_t>(kmin, bb0) - bb0;
            TaskL t;
            t.t.block = b; t.t.k0 = (int32_t)k0;
            t.t.col_off = (int32_t)col_off; t.t.ncol = (int32_t)ncol; t.t.xoff = 0; t.t.spos0 = 0; t.t.pad = 0;
            t.level = 0;
            tl.push_back(t);
            F_.gemm_flops += fm * (double)ncol * (double)(B.w - k0) * (double)(B.nrows - B.w);
            F_.gemm_bytes += es * ((double)(B.nrows - k0) * (double)(B.w - k0) + 2.0 * (double)(B.w - k0) * (double)ncol +
                                   2.0 * (double)(B.nrows - B.w) * (double)ncol);
        }
    }
    // the levels.  Each workspace U row is read by one task (the task of the block whose pivot row it is), so a task's
    // level is one past the highest level of the tasks writing into its segment.  The tasks' rows are mapped to their
    // workspace positions in parallel: this checks that a segment's rows are consecutive there and lists, per task, the
    // distinct tasks whose segments its rows below fall in (always later tasks: the task order is topological); then
    // one pass in task order over those lists.
    const int64_t nt = (int64_t)tl.size();
    reader_.assign((size_t)std::max<int64_t>(Ru, 1), -1);
    for (int64_t i = 0; i < nt; i++) {
        const BlockDesc& B = F_.blocks[tl[i].t.block];
        const int64_t k0 = tl[i].t.k0, m = B.w - k0;
        const int32_t sp = hrowpos_[F_.block_rows[tl[i].t.block][k0]];
        if (sp < 0 || sp + m > Ru) { errf(err_, errlen_, "internal: a segment outside the workspace U rows"); return FPLUS_ERR_INTERNAL; }
        tl[i].t.spos0 = sp;
        for (int64_t r = 0; r < m; r++) reader_[sp + r] = (int32_t)i;
    }
    if ((int64_t)succ_.size() < nt) succ_.resize(nt);
    for (std::vector<int64_t>& t : tmark_) if ((int64_t)t.size() < nt) t.resize(nt, 0);
    const int64_t pstamp = (F_.panels + 1) << 32;   // a task's mark: unique over the panels and the panel's tasks
    std::atomic<int> bad{0};
    const double tlp0 = now_s();
    {
        const int64_t TPI = 8;
        pool_.run((nt + TPI - 1) / TPI, [&](int64_t c, int tid) {
            std::vector<int64_t>& mark = tmark_[tid];
            for (int64_t i = c * TPI; i < std::min(nt, (c + 1) * TPI); i++) {
                const int32_t* br = F_.block_rows[tl[i].t.block].data();
                const BlockDesc& B = F_.blocks[tl[i].t.block];
                const int64_t k0 = tl[i].t.k0, w = B.w, nr = B.nrows, stamp = pstamp | (i + 1);
                const int32_t sp = tl[i].t.spos0;
                std::vector<int32_t>& su = succ_[i];
                su.clear();
                for (int64_t r = k0; r < w; r++)
                    if (hrowpos_[br[r]] != sp + (int32_t)(r - k0)) bad.store(1, std::memory_order_relaxed);
                for (int64_t r = w; r < nr; r++) {
                    const int32_t hp = hrowpos_[br[r]];
                    if (hp < 0) { bad.store(2, std::memory_order_relaxed); continue; }
                    if (hp >= Ru) continue;   // a row not pivotal: no task reads it
                    const int32_t t = reader_[hp];
                    if (t < 0 || mark[t] == stamp) continue;
                    mark[t] = stamp;
                    if (t <= i) bad.store(3, std::memory_order_relaxed);
                    su.push_back(t);
                }
            }
        });
    }
    if (bad.load()) {
        const int b = bad.load();
        errf(err_, errlen_, b == 1 ? "internal: a segment's rows are not consecutive in the workspace"
                            : b == 2 ? "internal: update row outside the panel rows"
                                     : "internal: an update task writes into an earlier task's segment");
        return FPLUS_ERR_INTERNAL;
    }
    int32_t maxlevel = 0;
    {
        std::vector<int32_t> lin(nt, 0);
        for (int64_t i = 0; i < nt; i++) {
            const int32_t lev = lin[i] + 1;
            tl[i].level = lev;
            maxlevel = std::max(maxlevel, lev);
            for (const int32_t t : succ_[i]) lin[t] = std::max(lin[t], lev);
        }
    }
    if (deterministic_updates() && maxlevel > 0) {
        // each level split into sub-levels in task order: a task goes one past the last sub-level that wrote any of its
        // rows below its block in this level, so no two tasks of a sub-level share a row and each row is updated in task
        // order -- no sum depends on the order in which a launch's tiles run.  Dependencies stay satisfied: a level's
        // sub-levels all follow the previous level's.
        std::vector<std::vector<int32_t>> bylev((size_t)maxlevel + 1);
        for (int64_t i = 0; i < nt; i++) bylev[(size_t)tl[i].level].push_back((int32_t)i);
        if ((int64_t)det_next_.size() < ldw) { det_next_.resize((size_t)ldw, 0); det_stamp_.resize((size_t)ldw, -1); }
        std::fill(det_stamp_.begin(), det_stamp_.begin() + ldw, -1);
        std::vector<int32_t> newlev((size_t)nt, 0);
        int32_t base = 0;
        for (int32_t L = 1; L <= maxlevel; L++) {
            int32_t nsub = 0;
            for (const int32_t i : bylev[(size_t)L]) {
                const BlockDesc& B = F_.blocks[tl[i].t.block];
                const int32_t* br = F_.block_rows[tl[i].t.block].data();
                int32_t sub = 0;
                for (int64_t r = B.w; r < B.nrows; r++) {
                    const int32_t pos = hrowpos_[br[r]];
                    if (det_stamp_[(size_t)pos] == L) sub = std::max(sub, det_next_[(size_t)pos]);
                }
                for (int64_t r = B.w; r < B.nrows; r++) {
                    const int32_t pos = hrowpos_[br[r]];
                    det_stamp_[(size_t)pos] = L;
                    det_next_[(size_t)pos] = sub + 1;
                }
                newlev[(size_t)i] = base + sub + 1;
                nsub = std::max(nsub, sub + 1);
            }
            base += nsub;
        }
        for (int64_t i = 0; i < nt; i++) tl[i].level = newlev[(size_t)i];
        maxlevel = base;
    }
    F_.t_lev += now_s() - th_ws;
    F_.t_lev_pass += now_s() - tlp0;
    std::stable_sort(tl.begin(), tl.end(), [](const TaskL& a, const TaskL& b) { return a.level < b.level; });
    // tile rows per level: real, 64, or 32 where the level has fewer than four 64-row tiles per compute unit (measured:
    // the 32-row tiles, which read two k chunks ahead, are as fast or faster up to there); complex, always 32 (their
    // tiles multiply with three real products instead of four, which the 64-row tiles' registers do not allow)
    // The policy (tuning knob "update_tile_policy", chosen by fplus_amd.autotune): 0 this rule; 1 every level 32-row
    // tiles; 2 every level 64-row tiles; 3 and 4 the real rule for both scalar types with the bound at 4 and at 16 tiles
    // per compute unit.
    std::vector<int32_t> lev_rt(maxlevel + 2, RT);
    const int policy = update_tile_policy();
    if (policy == 1 || (policy == 0 && vs_ == 2)) std::fill(lev_rt.begin(), lev_rt.end(), RT / 2);
    else if (policy != 2) {
        std::vector<int64_t> nt64(maxlevel + 2, 0);
        for (const TaskL& x : tl) {
            const BlockDesc& B = F_.blocks[x.t.block];
            nt64[x.level] += (int64_t)((B.nrows - x.t.k0 + RT - 1) / RT) * ((x.t.ncol + CTG - 1) / CTG);
        }
        const int64_t fill = (policy == 4 ? 16 : 4) * (int64_t)devctx().ncu;
        for (int32_t L = 0; L <= maxlevel + 1; L++) lev_rt[L] = nt64[L] < fill ? RT / 2 : RT;
    }
    std::vector<UpdTask> tasks(tl.size());
    std::vector<int32_t> lev_gt(maxlevel + 2, 0);   // the first tile of each level's tasks
    std::vector<double> lev_flops(maxlevel + 2, 0.0);   // the TRSM's and the GEMM's algorithmic flops
    std::vector<double> lev_issued(maxlevel + 2, 0.0);
    std::vector<int32_t> lev_maxm(maxlevel + 2, 0);
    // the tiles are written straight into the upload staging, each task's from its offset (in parallel, below)
    std::vector<int64_t> toff(tl.size() + 1, 0);
    for (size_t i = 0; i < tl.size(); i++) {
        const int64_t rt = lev_rt[tl[i].level];
        toff[i + 1] = toff[i] + (int64_t)((F_.blocks[tl[i].t.block].nrows - tl[i].t.k0 + rt - 1) / rt) *
                                    ((tl[i].t.ncol + CTG - 1) / CTG);
    }
    const int64_t ntile = toff[tl.size()];
    UpdTile* gtiles = up_.take<UpdTile>((size_t)ntile);
    int64_t xtot = 0;
    for (size_t i = 0; i < tl.size(); i++) {
        tasks[i] = tl[i].t;
        tasks[i].xoff = xtot;
        xtot += (int64_t)(F_.blocks[tasks[i].block].w - tasks[i].k0) * tasks[i].ncol;
    }
    d_X_.reserve((size_t)std::max<int64_t>(xtot * vs_, 1));
    // a level of one task writes its rows alone (plain loads and stores); with several, a row tile whose rows another
    // task of the level writes too updates them with atomics, the others with plain loads and stores (tuning knob
    // "update_atomic_rows": 0 every tile of such a level atomic).  The rows written twice are found per level, the
    // levels in parallel (each thread its own row marks; a level's tasks stay on one thread).
    std::vector<int> lev_ntask(maxlevel + 2, 0);
    for (const TaskL& x : tl) lev_ntask[x.level]++;
    std::vector<int64_t> rtoff(tl.size() + 1, 0);   // each task's first row tile in cmark
    for (size_t i = 0; i < tl.size(); i++) {
        const int64_t rt = lev_rt[tl[i].level];
        rtoff[i + 1] = rtoff[i] + (F_.blocks[tl[i].t.block].nrows - tl[i].t.k0 + rt - 1) / rt;
    }
    std::vector<uint8_t> cmark((size_t)std::max<int64_t>(rtoff[tl.size()], 1), 1);
    if (update_atomic_rows()) {
        std::fill(cmark.begin(), cmark.end(), 0);
        std::vector<int64_t> lfirst(maxlevel + 2, 0);
        for (const TaskL& x : tl) lfirst[x.level + 1]++;
        for (int32_t L = 0; L <= maxlevel; L++) lfirst[L + 1] += lfirst[L];
        std::vector<int32_t> multi;
        for (int32_t L = 1; L <= maxlevel; L++)
            if (lev_ntask[L] > 1) multi.push_back(L);
        for (int t = 0; t < pool_.size(); t++) {
            if ((int64_t)wstamp_[t].size() < ldw) wstamp_[t].resize((size_t)ldw, -1);
            if ((int64_t)wwho_[t].size() < ldw) wwho_[t].resize((size_t)ldw, 0);
        }
        pool_.run((int64_t)multi.size(), [&](int64_t mi, int tid) {
            const int32_t L = multi[mi];
            int64_t* st = wstamp_[tid].data();
            int64_t* who = wwho_[tid].data();
            const int64_t stamp = ((F_.panels + 1) << 32) | L, rt = lev_rt[L];
            for (int64_t i = lfirst[L]; i < lfirst[L + 1]; i++) {
                const BlockDesc& B = F_.blocks[tl[i].t.block];
                const int32_t* br = F_.block_rows[tl[i].t.block].data();
                const int64_t k0 = tl[i].t.k0;
                for (int64_t r = B.w; r < B.nrows; r++) {
                    const int32_t pos = hrowpos_[br[r]];
                    const int64_t mine = rtoff[i] + (r - k0) / rt;
                    if (st[pos] == stamp) {
                        cmark[(size_t)mine] = 1;
                        cmark[(size_t)who[pos]] = 1;
                    } else {
                        st[pos] = stamp;
                        who[pos] = mine;
                    }
                }
            }
        });
    }
    for (size_t i = 0; i < tl.size(); i++) {
        const UpdTask& T = tasks[i];
        const int32_t L = tl[i].level;
        const BlockDesc& B = F_.blocks[T.block];
        const int64_t m = B.w - T.k0;
        const double tfl = fm * (double)T.ncol * (double)m * ((double)(m - 1) / 2 + (double)(B.nrows - B.w));
        lev_flops[L] += tfl;
        if ((size_t)T.block >= F_.blk_upd_flops.size()) F_.blk_upd_flops.resize((size_t)T.block + 1, 0.0);
        F_.blk_upd_flops[T.block] += tfl;
        {
            // the product's size against what the tiles issue on the matrix cores (rows, columns and k padded)
            const double prod = fm * (double)T.ncol * (double)m * (double)(B.nrows - T.k0);
            const int32_t rt = lev_rt[L];
            // rows and k over the row tiles: the q tiles inside the triangle stop at their last row (k rt, 2 rt, ..),
            // the others run to w (k padded to chunks of 8); a wave skips a part of a tile with no used row or
            // column, so rows count in 32s and columns in a wave's width (16 for 32-row tiles, 32 for 64-row tiles)
            const int64_t q = m / rt, rows_rest = (B.nrows - T.k0 - q * rt + 31) / 32 * 32, cg = rt == 32 ? 16 : 32;
            const double kr = (double)rt * rt * (double)(q * (q + 1) / 2) + (double)rows_rest * (double)((m + 7) / 8 * 8);
            const double issued = fm * kr * (double)((T.ncol + cg - 1) / cg * cg);
            F_.upd_prod_flops += prod;
            F_.upd_mfma_flops += issued;
            lev_issued[L] += issued;
            lev_maxm[L] = std::max<int32_t>(lev_maxm[L], (int32_t)m);
            const int bkt = T.ncol <= 8 ? 0 : T.ncol <= 16 ? 1 : T.ncol <= 32 ? 2 : T.ncol <= 64 ? 3 : 4;
            F_.upd_ncol_flops[bkt] += prod;
        }
        lev_gt[L + 1] = (int32_t)toff[i + 1];
    }
    {
        const int64_t nt = (int64_t)tl.size(), TPI = 16;
        pool_.run((nt + TPI - 1) / TPI, [&](int64_t c, int) {
            for (int64_t i = c * TPI; i < std::min(nt, (c + 1) * TPI); i++) {
                const UpdTask& T = tasks[i];
                const int32_t L = tl[i].level, rt = lev_rt[L];
                const BlockDesc& B = F_.blocks[T.block];
                const int64_t m = B.w - T.k0;
                int64_t gi = toff[i];
                for (int32_t r0 = T.k0; r0 < B.nrows; r0 += rt)
                    for (int32_t cc = 0; cc < T.ncol; cc += CTG) {
                        UpdTile& u = gtiles[gi++];
                        u.iv = B.ivals;
                        u.rows = B.rows;
                        u.xo = T.xoff + (int64_t)cc * m;
                        u.nrows = B.nrows; u.w = B.w; u.k0 = T.k0; u.row0 = r0;
                        u.cfirst = T.col_off + cc; u.nc = std::min<int32_t>(CTG, T.ncol - cc);
                        u.spos0 = T.spos0; u.flags = lev_ntask[L] > 1 && cmark[(size_t)(rtoff[i] + (r0 - T.k0) / rt)] ? FPA_TILE_SHARED : 0;
                        u.block = T.block; u.pad = 0;
                    }
            }
        });
    }
    for (int32_t L = 1; L <= maxlevel + 1; L++) lev_gt[L] = std::max(lev_gt[L], lev_gt[L - 1]);
    if (xcd_tile_order()) {
        // XCD-aware order of each level's tiles: a launch's workgroup i runs on XCD i mod 8 (round-robin dispatch), and
        // the row tiles of one task and column tile all read the same slice of the segment S; giving them positions of
        // one residue mod 8 keeps that slice in one XCD's L2, where otherwise every XCD fetches it from HBM.  Groups go to
        // the least loaded XCD, largest first; an XCD whose list runs out lends its positions to the fullest one.
        constexpr int NX = 8;
        std::vector<UpdTile> tmp;
        std::vector<std::pair<int64_t, int64_t>> grp;   // (size, first tile); a group's tiles are ncc apart
        std::vector<int64_t> gstride;
        size_t ti = 0;
        for (int32_t L = 1; L <= maxlevel; L++) {
            const int64_t g0 = lev_gt[L], g1 = lev_gt[L + 1];
            const size_t t0 = ti;
            while (ti < tl.size() && tl[ti].level == L) ti++;
            if (g1 - g0 < 2 * NX) continue;
            grp.clear();
            gstride.clear();
            for (size_t i = t0; i < ti; i++) {
                const int64_t rt = lev_rt[L];
                const int64_t nrt = (F_.blocks[tasks[i].block].nrows - tasks[i].k0 + rt - 1) / rt;
                const int64_t ncc = (tasks[i].ncol + CTG - 1) / CTG;
                for (int64_t cc = 0; cc < ncc; cc++) { grp.push_back({nrt, toff[i] + cc}); gstride.push_back(ncc); }
            }
            std::vector<int64_t> order(grp.size());
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return grp[x].first > grp[y].first; });
            std::vector<std::vector<int64_t>> bin(NX);
            std::vector<int64_t> load(NX, 0);
            for (const int64_t gi : order) {
                const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
                for (int64_t r = 0; r < grp[gi].first; r++) bin[x].push_back(grp[gi].second + r * gstride[gi]);
                load[x] += grp[gi].first;
            }
            tmp.resize((size_t)(g1 - g0));
            std::vector<size_t> nxt(NX, 0);
            for (int64_t p2 = 0; p2 < g1 - g0; p2++) {
                int x = (int)(p2 % NX);
                if (nxt[x] >= bin[x].size()) {
                    int best = -1;
                    for (int y = 0; y < NX; y++)
                        if (nxt[y] < bin[y].size() && (best < 0 || bin[y].size() - nxt[y] > bin[best].size() - nxt[best])) best = y;
                    x = best;
                }
                tmp[(size_t)p2] = gtiles[bin[x][nxt[x]++]];
            }
            std::copy(tmp.begin(), tmp.end(), gtiles + g0);
        }
    }
    for (int32_t L = 1; L <= maxlevel; L++) {   // levels by their tile count (powers of two), for the statistics
        const int32_t nt = lev_gt[L + 1] - lev_gt[L];
        if (nt <= 0) continue;
        int bk = 0;
        while (bk < 12 && (1 << bk) < nt) bk++;
        F_.lev_hist[bk].levels++;
        F_.lev_hist[bk].tiles += nt;
        F_.lev_hist[bk].issued += lev_issued[L];
        F_.lev_hist[bk].maxm += lev_maxm[L];
    }
    for (int32_t L = 1; L <= maxlevel; L++) F_.shared_levels += lev_ntask[L] > 1;
    for (size_t i = 0; i < tl.size(); i++)
        if (lev_ntask[tl[i].level] > 1) {
            const int64_t ncc = (tl[i].t.ncol + CTG - 1) / CTG;
            F_.shared_level_tiles += (rtoff[i + 1] - rtoff[i]) * ncc;
            for (int64_t q = rtoff[i]; q < rtoff[i + 1]; q++) F_.atomic_tiles += cmark[(size_t)q] * ncc;
        }
    if (rec_) {   // what a refactorization replays: the tasks, the tiles and the levels (in the order they run)
        ReplayPanel P{};
        P.j0 = j0; P.np = np; P.Ru = Ru; P.Rl = Rl; P.ldw = ldw; P.xtot = xtot;
        P.cols_off = (int64_t)rec_->h_cols.size();
        rec_->h_cols.insert(rec_->h_cols.end(), cols.begin(), cols.end());
        P.tasks_off = (int64_t)rec_->h_tasks.size();
        P.ntask = (int64_t)tasks.size();
        rec_->h_tasks.insert(rec_->h_tasks.end(), tasks.begin(), tasks.end());
        const int64_t tb = (int64_t)rec_->h_tiles.size();
        rec_->h_tiles.insert(rec_->h_tiles.end(), gtiles, gtiles + ntile);
        P.lev_off = (int64_t)rec_->levels.size();
        for (int32_t L = 1; L <= maxlevel; L++)
            if (lev_gt[L + 1] > lev_gt[L]) rec_->levels.push_back(ReplayLevel{tb + lev_gt[L], lev_gt[L + 1] - lev_gt[L], lev_rt[L], lev_flops[L]});
        P.nlev = (int64_t)rec_->levels.size() - P.lev_off;

        rec_->max_w = std::max(rec_->max_w, ldw * np * vs_);
        rec_->max_x = std::max(rec_->max_x, xtot * vs_);
        rec_->max_np = std::max(rec_->max_np, np);
        rec_panel_ = P;
    }
    F_.levels += maxlevel;
    F_.tasks += (int64_t)tasks.size();
    F_.t_tasks += now_s() - th_p1;
    F_.t_host += now_s() - th0;
    // ---- device: workspace, scatter, updates, panel LU
    d_W_.reserve((size_t)(ldw * np * vs_));
    d_univ_.reserve(univ.size());
    up_.upload(d_univ_.p, univ.data(), univ.size(), s_);
    k::rowmap_set((int64_t)univ.size(), d_univ_.p, d_rowmap_.p, 0, s_);
    k::fill_f64(d_W_.p, ldw * np * vs_, 0.0, s_);
    d_colmax_.reserve(np);
    k::scatter_cols(vs_, (int32_t)np, d_Bp_, d_Bi_, d_Bx_, j0, d_rowmap_.p, d_W_.p, ldw, d_colmax_.p, s_);
    PanelStatus zero_st;
    std::memset(&zero_st, 0, sizeof(zero_st));
    up_.upload(&d_pout_.p->st, &zero_st, 1, s_);
    if (!tasks.empty()) {
        d_tasks_.reserve(tasks.size());
        up_.upload(d_tasks_.p, tasks.data(), tasks.size(), s_);
        d_cols_.reserve(cols.size());
        up_.upload(d_cols_.p, cols.data(), cols.size(), s_);
        d_tiles_.reserve(ntile > 0 ? (size_t)ntile : 1);
        up_.send(d_tiles_.p, gtiles, (size_t)ntile, s_);
        for (int32_t L = 1; L <= maxlevel; L++) {
            const int32_t g0 = lev_gt[L], g1 = lev_gt[L + 1];
            k::blk_upd(vs_, g1 - g0, d_tiles_.p + g0, d_cols_.p, d_rowmap_.p, d_W_.p, ldw, lev_rt[L], lev_flops[L], d_X_.p,
                       opt_.precision, F_.d_blk_f32.p, F_.d_blk_err.p, s_);
        }
        k::x_back(vs_, (int32_t)tasks.size(), d_tasks_.p, d_cols_.p, d_blocks_p(), d_W_.p, ldw, d_X_.p, s_);
    }
    double* Wl = d_W_.p + Ru * vs_;
    if (np > FPA_PANEL_MAX) { errf(err_, errlen_, "internal: a panel wider than %d columns", FPA_PANEL_MAX); return FPLUS_ERR_INTERNAL; }
    PanelStatus* d_st = &d_pout_.p->st;
    int32_t* d_piv = d_pout_.p->piv;
    // each sub-panel's pivot positions and the status so far come back as soon as it is factored: the host works on
    // a sub-panel's columns while the device factors the next ones
    const int64_t nsp = (np + FPA_PANEL_NB - 1) / FPA_PANEL_NB;
    if ((int64_t)h_pout_.n < nsp) {
        h_pout_.alloc(nsp);
        hip_check(hipHostGetDevicePointer((void**)&h_pout_dev_, h_pout_.p, 0), "hipHostGetDevicePointer");
    }
    sp_ev_.resize(nsp);
    for (int64_t a = 0, sp = 0; a < np; a += FPA_PANEL_NB, sp++) {
        const int32_t nb = (int32_t)std::min<int64_t>(FPA_PANEL_NB, np - a);
        h_pout_.p[sp].st.err = INT32_MIN;   // (the kernel writes the slot; this marks it not written yet)
        k::getf2(vs_, Wl, ldw, (int32_t)Rl, (int32_t)np, (int32_t)a, nb, d_colmax_.p, opt_.pivot_rel, d_piv, d_st,
                 h_pout_dev_ + sp, d_coop_.p, d_coop_cnt_.p, &coop_counted_, opt_.coop_panel_lu, s_);
        k::gemm_sub(vs_, Wl, ldw, (int32_t)Rl, (int32_t)np, (int32_t)a, nb, d_st, s_);
        hip_check(hipEventRecord(sp_ev_.ev[sp], s_), "hipEventRecord");
    }
    // while the device runs the updates: the columns' candidates as lower positions and bitsets (in parallel)
    const int64_t nw = (Rl + 63) / 64;
    {
        if ((int64_t)c_lp_.size() < np) c_lp_.resize(np);
        cbits_.resize((size_t)std::max<int64_t>(np * nw, 1));
        std::atomic<int> badlp{0};
        pool_.run(np, [&](int64_t ci, int) {
            std::vector<int32_t>& lp = c_lp_[ci];
            const std::vector<int32_t>& cn = c_np[ci];
            lp.resize(cn.size());
            uint64_t* cb = cbits_.data() + ci * nw;
            std::fill(cb, cb + nw, (uint64_t)0);
            for (size_t k = 0; k < cn.size(); k++) {
                const int64_t h = (int64_t)hrowpos_[cn[k]] - Ru;
                if (h < 0 || h >= Rl) { badlp.store(1, std::memory_order_relaxed); lp[k] = 0; continue; }
                lp[k] = (int32_t)h;
                cb[h >> 6] |= (uint64_t)1 << (h & 63);
            }
        });
        if (badlp.load()) {
            dev_sync(s_);
            errf(err_, errlen_, "internal: a candidate outside the panel's lower rows");
            return FPLUS_ERR_INTERNAL;
        }
        pivcol_.assign((size_t)std::max<int64_t>(Rl, 1), INT32_MAX);
    }
    const double th1 = now_s();
    double waited = 0;
    // the rows' order as the row swaps leave it: after column ci's swap, position ci holds its pivot row for good
    std::vector<int32_t> lrows(univ.begin() + Ru, univ.end());
    // ---- phase 3 on the host, with the device's pivots (src/factor_sn_panel.inc:162-390)
    const int64_t u_begin = Up_[j0];
    std::vector<URun> pruns;
    std::vector<std::pair<int64_t, int64_t>> touched;   // (supernode, first column in this panel)
    for (int64_t ci = 0; ci < np; ci++) {
        if (ci % FPA_PANEL_NB == 0) {
            const int64_t sp = ci / FPA_PANEL_NB;
            const double tw0 = now_s();
            hip_check(hipEventSynchronize(sp_ev_.ev[sp]), "hipEventSynchronize");
            waited += now_s() - tw0;
            if (sp > 0) F_.t_wait_lu += now_s() - tw0;
            const PanelStatus st = h_pout_.p[sp].st;
            if (st.err == INT32_MIN) {
                dev_sync(s_);
                errf(err_, errlen_, "internal: the panel LU left no status");
                return FPLUS_ERR_INTERNAL;
            }
            if (st.err < 0) {
                dev_sync(s_);
                F_.coop_aborted = true;
                errf(err_, errlen_, "internal: the cooperative panel LU's workgroups did not meet at a barrier");
                return FPLUS_ERR_INTERNAL;
            }
            if (st.err) {
                dev_sync(s_);
                const int64_t j = j0 + st.err - 1;
                errf(err_, errlen_, "Near-zero or zero pivot %.6e at step %ld (threshold %.6e). Matrix is singular.", st.best, (long)j, st.thr);
                return FPLUS_ERR_SINGULAR;
            }
            F_.pivot_ties += st.ties;
        }
        const int32_t pv = h_pout_.p[ci / FPA_PANEL_NB].piv[ci];
        if (pv < ci || pv >= Rl) {
            dev_sync(s_);
            errf(err_, errlen_, "internal: pivot position %d out of range for column %ld", (int)pv, (long)(j0 + ci));
            return FPLUS_ERR_INTERNAL;
        }
        std::swap(lrows[ci], lrows[pv]);
        const int64_t j = j0 + ci, stamp2 = n + 1 + j;
        const double tq0 = now_s();
        // one pass over the column's candidates from the search before the panel: the ones pivotal in this panel (the
        // within-panel search starts from them, in order) apart, and for the ones still not pivotal -- in order: the
        // candidate order is the pivot tie-break order -- the count of the open supernode's rows among them and the
        // others in order (the join's test, below); the search's new candidates follow, in its order (a row is a new one
        // if it is not in the column's bitset)
        int64_t ncand = 0, top = n;
        piv_c_.clear();
        std::vector<int64_t>& npv = miss_;   // reused: the candidates missing from the open supernode, in order
        int64_t in_S = 0, n_miss = 0;
        const int64_t Sw = open_ >= 0 ? sn_[open_].w : INT64_MAX;
        {
            const int32_t* lp = c_lp_[ci].data();
            const int32_t* cr = c_np[ci].data();
            const size_t C = c_np[ci].size();
            for (size_t k = 0; k < C; k++) {
                const int64_t r = cr[k];
                if (pivcol_[lp[k]] < (int32_t)ci) piv_c_.push_back(r);
                else if (bpos_[r] >= Sw) in_S++;
                else npv[n_miss++] = r;
            }
        }
        const uint64_t* cb = cbits_.data() + ci * nw;
        auto in_list = [&](int64_t i) {
            const int64_t h = (int64_t)hrowpos_[i] - Ru;
            return h >= 0 && h < Rl && ((cb[h >> 6] >> (h & 63)) & 1);
        };
        F_.t_p3_split += now_s() - tq0;
        if (open0 >= 0 && sn_[open0].w > we_open && ws0_.smark[open0] != stamp2) {
            // while it stays open, the rows past its pivots that the supernode had when the panel began are all among
            // the column's candidates (its search before the panel read them) and none has become pivotal (a pivot of
            // the panel is either in it, moved above its rows past the pivots, or closed it): the walk finds nothing
            // there and starts at the rows appended since
            const int64_t skip = open_ == open0 ? std::max<int64_t>(0, nr_open - sn_[open0].w) : 0;
            for (int64_t q = ext_off[ci]; q < ext_off[ci + 1]; q++)
                if (ext_s[q] == open0) { top = dfs_in(ws0_, open0, sn_[open0].c0 + we_open, top, stamp2, open_, &ncand, skip, in_list); break; }
        }
        for (int64_t i : piv_c_) {
            const int64_t k = pinv_[i];
            const int64_t s = col_sn_[k];
            if (ws0_.smark[s] == stamp2) { if (k < ws0_.seg0[s]) ws0_.seg0[s] = (int32_t)k; continue; }
            top = dfs_in(ws0_, s, k, top, stamp2, open_, &ncand, 0, in_list);
        }
        for (int64_t q = 0; q < ncand; q++) {
            const int64_t r = ws0_.cand[q];
            if (bpos_[r] >= Sw) in_S++;
            else npv[n_miss++] = r;
        }
        const int64_t pivrow = lrows[ci];
        pivcol_[hrowpos_[pivrow] - Ru] = (int32_t)ci;
        const double tq1 = now_s();
        // U(:, j): pre-panel segments, within-panel segments, the diagonal -- each a run of consecutive rows whose values
        // are consecutive in the workspace column
        int64_t uo = Up_[j];
        auto put_run = [&](int64_t row0, int64_t len, int64_t wrow) {
            if (len <= 0) return;
            Ur_row_.push_back((int32_t)row0);
            Ur_len_.push_back((int32_t)len);
            pruns.push_back(URun{uo, wrow + ci * ldw, (int32_t)row0, (int32_t)len});
            uo += len;
        };
        for (int64_t q = ext_off[ci]; q < ext_off[ci + 1]; q++) {
            const HSn& S = sn_[ext_s[q]];
            const int64_t we = ext_s[q] == open0 ? we_open : S.w, u = uidx_[ext_s[q]];
            put_run(ext_kf[q], S.c0 + we - ext_kf[q], ubase[u] + (ext_kf[q] - ukmin[u]));
        }
        for (int64_t t = top; t < n; t++) {
            const HSn& S = sn_[ws0_.topo[t]];
            const int64_t k0 = ws0_.seg0[ws0_.topo[t]];
            put_run(k0, S.c0 + S.w - k0, Ru + (k0 - j0));
        }
        put_run(j, 1, Ru + ci);
        Up_[j + 1] = uo;
        Urp_[j + 1] = (int64_t)Ur_row_.size();
        const double tq2 = now_s();
        // L(:, j): join the open supernode (exactly or relaxed), else open a new one
        int join = 0;
        if (open_ >= 0) {
            HSn& S = sn_[open_];
            const int64_t extra = (S.nr() - S.w) - in_S;
            if (n_miss + extra == 0) join = 1;
            else if (relax_ok(S, n_miss, extra)) join = 2;
            if (join == 2) {
                S.nzero += (double)n_miss * (double)S.w + (double)extra;
                for (int64_t t = 0; t < n_miss; t++) { bpos_[npv[t]] = (int32_t)S.nr(); S.ri.push_back(npv[t]); }
            }
        }
        if (join) {
            HSn& S = sn_[open_];
            const int64_t w = S.w, pp = bpos_[pivrow];
            if (pp != w) {
                const int64_t other = S.ri[w];
                S.ri[w] = pivrow; S.ri[pp] = other;
                bpos_[pivrow] = (int32_t)w; bpos_[other] = (int32_t)pp;
            }
            S.w = w + 1;
            col_sn_[j] = open_;
            if (touched.empty() || touched.back().first != open_) touched.push_back({open_, j});
        } else {
            if (open_ >= 0) {
                HSn& S = sn_[open_];
                S.srch.assign(S.ri.begin() + S.w, S.ri.end());
                S.closed = true;
                for (int64_t r : S.ri) bpos_[r] = -1;
            }
            const int64_t s = (int64_t)sn_.size();
            sn_.emplace_back();
            HSn& S = sn_.back();
            S.c0 = j; S.w = 1;
            // its rows: the pivot row, then the candidates still not pivotal in order (those from before the panel,
            // then the search's new ones)
            S.ri.reserve(c_np[ci].size() + ncand);
            S.ri.push_back(pivrow); bpos_[pivrow] = 0;
            auto add = [&](int64_t r) {
                if (r == pivrow) return;
                bpos_[r] = (int32_t)S.nr();
                S.ri.push_back(r);
            };
            const int32_t* lp = c_lp_[ci].data();
            for (size_t k = 0; k < c_np[ci].size(); k++)
                if (pivcol_[lp[k]] >= (int32_t)ci) add(c_np[ci][k]);
            for (int64_t q = 0; q < ncand; q++) add(ws0_.cand[q]);
            col_sn_[j] = s;
            open_ = s;
            touched.push_back({s, j});
        }
        pinv_[pivrow] = j; p_[j] = pivrow;
        const double tq3 = now_s();
        // prune the closed supernodes this column reached (pre-panel, then within-panel)
        for (int64_t q = ext_off[ci]; q < ext_off[ci + 1]; q++) if (ext_s[q] != open_) prune(ext_s[q], pivrow);
        for (int64_t t = top; t < n; t++) if (ws0_.topo[t] != open_) prune(ws0_.topo[t], pivrow);
        const double tq4 = now_s();
        F_.t_p3_dfs += tq1 - tq0;
        F_.t_p3_u += tq2 - tq1;
        F_.t_p3_join += tq3 - tq2;
        F_.t_p3_prune += tq4 - tq3;
    }
    F_.t_p3 += now_s() - th1 - waited;
    F_.t_wait += waited;
    F_.t_host += now_s() - th1 - waited;
    // ---- the output on the helper thread, while the next panel's search runs
    out_.j0 = j0;
    out_.j1 = j1;
    out_.Rl = Rl;
    out_.ldw = ldw;
    out_.u_begin = u_begin;
    out_.Wl = Wl;
    if (rec_) {   // the refactorization's workspace rows: the U rows, then the lower rows in their final order
        rec_panel_.rows_off = (int64_t)rec_->h_rows.size();
        rec_->h_rows.insert(rec_->h_rows.end(), univ.begin(), univ.begin() + Ru);
        rec_->h_rows.insert(rec_->h_rows.end(), lrows.begin(), lrows.end());
    }
    for (int32_t row : univ) hrowpos_[row] = -1;   // (the next panel's workspace rows are placed meanwhile)
    out_.univ = std::move(univ);
    out_.lrows = std::move(lrows);
    out_.touched = std::move(touched);
    out_.pruns = std::move(pruns);
    out_pending_ = true;
    side_.submit([this] { out_rc_ = panel_out(); });
    return FPLUS_OK;
}

int Driver::panel_out()
{
    const double th_out = now_s();
    const int64_t j0 = out_.j0, j1 = out_.j1, Rl = out_.Rl, ldw = out_.ldw, u_begin = out_.u_begin;
    double* Wl = out_.Wl;
    const std::vector<int32_t>& univ = out_.univ;
    const std::vector<int32_t>& lrows = out_.lrows;
    const std::vector<std::pair<int64_t, int64_t>>& touched = out_.touched;
    const std::vector<URun>& pruns = out_.pruns;
    // the final positions of the lower rows
    for (int64_t i = 0; i < Rl; i++) lposmap_[lrows[i]] = (int32_t)i;
    // ---- device: the new blocks and the U values out of the workspace
    std::vector<ExtractCol> ex;
    std::vector<int32_t> expos;
    new_blocks_.clear();
    for (size_t t = 0; t < touched.size(); t++) {
        const int64_t s = touched[t].first, b0 = touched[t].second;
        const int64_t b1 = std::min<int64_t>(sn_[s].c0 + sn_[s].w, j1);
        add_block(s, b0, b1, j0, lposmap_, ex, expos);
    }
    d_blocks_upload();
    if (!ex.empty()) {
        d_ex_.reserve(ex.size());
        up_.upload(d_ex_.p, ex.data(), ex.size(), s_);
        d_pos_.reserve(expos.size());
        up_.upload(d_pos_.p, expos.data(), expos.size(), s_);
        d_lrows_.reserve(lrows.size());
        up_.upload(d_lrows_.p, lrows.data(), lrows.size(), s_);
        k::extract_cols(vs_, (int32_t)ex.size(), d_ex_.p, d_pos_.p, d_lrows_.p, Wl, ldw, s_);
    }
    // the new blocks' inverse forms, for the updates of later panels
    std::vector<int2> ijobs;
    std::vector<GemmTile> itiles;
    double iflops = 0;
    for (const int32_t b : new_blocks_) {
        const BlockDesc& B = F_.blocks[b];
        for (int32_t c = 0; c < B.w; c++) ijobs.push_back(make_int2(b, c));
        for (int32_t r0 = 0; r0 < B.nrows - B.w; r0 += 64)
            for (int32_t c0 = 0; c0 < B.w; c0 += 64) itiles.push_back(GemmTile{b, r0, c0, 0});
        iflops += (vs_ == 1 ? 2.0 : 8.0) * (double)B.w * (double)B.w * ((double)B.w / 6 + (double)(B.nrows - B.w) / 2);
    }
    if (!ijobs.empty()) {
        d_ijobs_.reserve(ijobs.size());
        up_.upload(d_ijobs_.p, ijobs.data(), ijobs.size(), s_);
        d_itiles_.reserve(itiles.size() > 0 ? itiles.size() : 1);
        up_.upload(d_itiles_.p, itiles.data(), itiles.size(), s_);
        k::blk_inverse(vs_, (int32_t)ijobs.size(), d_ijobs_.p, (int32_t)itiles.size(), d_itiles_.p, d_blocks_p(), iflops, s_);
    }
    if ((opt_.precision == 2 || opt_.precision == 3) && !new_blocks_.empty()) {
        // the new blocks' features and (precision 2) the selector's marks, on the device: later panels' update tiles read
        // the marks, so the decisions need no trip to the host
        d_newblk_.reserve(new_blocks_.size());
        up_.upload(d_newblk_.p, new_blocks_.data(), new_blocks_.size(), s_);
        k::block_features(vs_, (int32_t)new_blocks_.size(), d_newblk_.p, d_blocks_p(), d_colmax_.p, j0, opt_.sel.d_Dr, opt_.sel.d_Dc,
                          n_, F_.d_featmax.p, F_.d_feat.p, s_);
        if (opt_.precision == 2)
            k::sel_infer((int64_t)new_blocks_.size(), d_newblk_.p, F_.d_feat.p, opt_.sel.d_model, opt_.sel.threshold, nullptr,
                         F_.d_blk_f32.p, s_);
    }
    {
        const int64_t cnt = Up_[j1] - u_begin;
        if (u_begin + cnt > (int64_t)(F_.d_Ux.n / vs_)) {
            DevBuf<double> nb((size_t)std::max<int64_t>(2 * (u_begin + cnt), 1024) * vs_);
            if (u_begin > 0)
                hip_check(hipMemcpyAsync(nb.p, F_.d_Ux.p, u_begin * vs_ * sizeof(double), hipMemcpyDeviceToDevice, s_), "grow U");
            dev_sync(s_);
            F_.d_Ux = std::move(nb);
        }
        if (u_begin + cnt > (int64_t)F_.d_Ui.n) {
            DevBuf<int32_t> nb((size_t)std::max<int64_t>(2 * (u_begin + cnt), 1024));
            if (u_begin > 0)
                hip_check(hipMemcpyAsync(nb.p, F_.d_Ui.p, u_begin * sizeof(int32_t), hipMemcpyDeviceToDevice, s_), "grow U rows");
            dev_sync(s_);
            F_.d_Ui = std::move(nb);
        }
        d_runs_.reserve(pruns.size());
        up_.upload(d_runs_.p, pruns.data(), pruns.size(), s_);
        k::u_runs(vs_, (int32_t)pruns.size(), d_runs_.p, d_W_.p, F_.d_Ux.p, F_.d_Ui.p, s_);
        F_.nnzU = u_begin + cnt;
    }
    k::rowmap_set((int64_t)univ.size(), d_univ_.p, d_rowmap_.p, 1, s_);
    // the device finishes the panel while the host starts the next one (the uploads read page-locked staging)
    up_.end(s_);
    if (rec_) {   // (rec_panel_ is the main thread's again only after it joined this job)
        ReplayPanel& P = rec_panel_;
        P.blk0 = new_blocks_.empty() ? (int64_t)F_.blocks.size() : (int64_t)*std::min_element(new_blocks_.begin(), new_blocks_.end());
        P.blk1 = new_blocks_.empty() ? P.blk0 : (int64_t)*std::max_element(new_blocks_.begin(), new_blocks_.end()) + 1;
        P.ex_off = (int64_t)rec_->h_ex.size();
        P.nex = (int64_t)ex.size();
        rec_->h_ex.insert(rec_->h_ex.end(), ex.begin(), ex.end());
        P.expos_off = (int64_t)rec_->h_expos.size();
        rec_->h_expos.insert(rec_->h_expos.end(), expos.begin(), expos.end());
        P.ij_off = (int64_t)rec_->h_ij.size();
        P.nij = (int64_t)ijobs.size();
        rec_->h_ij.insert(rec_->h_ij.end(), ijobs.begin(), ijobs.end());
        P.it_off = (int64_t)rec_->h_it.size();
        P.nit = (int64_t)itiles.size();
        rec_->h_it.insert(rec_->h_it.end(), itiles.begin(), itiles.end());
        P.iflops = iflops;
        P.runs_off = (int64_t)rec_->h_runs.size();
        P.nruns = (int64_t)pruns.size();
        rec_->h_runs.insert(rec_->h_runs.end(), pruns.begin(), pruns.end());
        P.nb_off = (int64_t)rec_->h_newblk.size();
        P.nnb = (int64_t)new_blocks_.size();
        rec_->h_newblk.insert(rec_->h_newblk.end(), new_blocks_.begin(), new_blocks_.end());
        rec_->panels.push_back(P);
    }
    for (int32_t row : lrows) lposmap_[row] = -1;
    F_.t_out += now_s() - th_out;
    F_.t_host += now_s() - th_out;
    F_.panels++;
    return FPLUS_OK;
}


int Driver::run()
{
    const int64_t n = n_;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    s_ = devctx().stream;
    F_.n = n;
    F_.vs = vs_;
    relax_ = opt_.relax_rows > 0 ? opt_.relax_rows : 0;
    relax_auto_ = opt_.relax_rows < 0;
    if (n >= 700000000) {   // the search's 32-bit marks reach 3 n + P
        errf(err_, errlen_, "matrix order %ld is beyond this build's supernodal factorization", (long)n);
        return FPLUS_ERR_BADARG;
    }
    const int64_t nnzA = Bp_[n];
    if (n <= 4096 && nnzA <= 24 * n) {   // the automatic rule's bounds for small sparse matrices (src/factor_sn.c:644-645)
        rz_w_[0] = 2; rz_w_[1] = 8; rz_w_[2] = 32; rz_z_[0] = 0.3; rz_z_[1] = 0.03; rz_z_[2] = 0.01; rz_cap_ = 8;
    }
    col_sn_.assign(nn, -1); pinv_.assign(nn, -1); p_.assign(nn, -1); bpos_.assign(nn, -1);
    ws0_.init(nn);
    wsT_.resize(pool_.size());
    tmark_.assign(pool_.size(), {});
    wstamp_.assign(pool_.size(), {});
    wwho_.assign(pool_.size(), {});
    for (SWork& w : wsT_) w.init(nn);
    miss_.assign(nn, 0); ustamp_.assign(nn, 0); uidx_.assign(nn, 0);
    hrowpos_.assign(nn, -1); lposmap_.assign(nn, -1);
    sn_.reserve(nn);
    Up_.assign(n + 1, 0);
    Urp_.assign(n + 1, 0);
    Ur_row_.reserve((size_t)(nnzA + n + 16));
    Ur_len_.reserve((size_t)(nnzA + n + 16));
    d_rowmap_.alloc(nn);
    k::fill_i32(d_rowmap_.p, n, -1, s_);
    d_pout_.alloc(1);
    d_coop_.alloc(k::getf2_coop_bytes(vs_) / sizeof(double) + 1);
    d_coop_cnt_.alloc(2);
    k::fill_i32(d_coop_cnt_.p, 2, 0, s_);
    coop_counted_ = 0;
    F_.d_blocks.alloc(nn);
    F_.d_blk_f32.alloc(nn);   // (a block has at least one column)
    k::fill_u8(F_.d_blk_f32.p, (int64_t)nn, opt_.precision == 1 ? 1 : 0, s_);
    F_.precision = opt_.precision;
    if ((opt_.precision == 2 || opt_.precision == 3)) {
        if (opt_.precision == 2 && !opt_.sel.d_model) throw std::invalid_argument("per-block precision without a selector model");
        if (!opt_.sel.d_Dr || !opt_.sel.d_Dc) throw std::invalid_argument("the selector's features need the scalings");
        F_.d_feat.alloc(nn * SEL_NF);
        F_.d_featmax.alloc(3 * nn);
        k::fill_u64(F_.d_featmax.p, (int64_t)(3 * nn), 0ULL, s_);
        if (opt_.precision == 3) {
            F_.d_blk_err.alloc(2 * nn);
            k::fill_u64(F_.d_blk_err.p, (int64_t)(2 * nn), 0ULL, s_);
        }
    }
    if (opt_.record) {
        F_.replay.reset(new Replay());
        rec_ = F_.replay.get();
    }
    F_.val_chunks.clear();
    F_.row_chunks.clear();
    vals_.chunks = &F_.val_chunks;
    vals_.min_chunk = (size_t)1 << 25;
    ivals_.chunks = &iv_chunks_;
    ivals_.min_chunk = (size_t)1 << 25;
    F_.linv_chunks.clear();
    linv_.chunks = &F_.linv_chunks;
    linv_.min_chunk = (size_t)1 << 24;
    F_.uinv_chunks.clear();
    uinv_.chunks = &F_.uinv_chunks;
    uinv_.min_chunk = (size_t)1 << 24;
    rows_.chunks = &F_.row_chunks;
    rows_.min_chunk = (size_t)1 << 24;
    const int64_t P = std::max<int64_t>(1, opt_.panel);
    for (int64_t j0 = 0; j0 < n; j0 += P) {
        int rc = FPLUS_OK;
        try {
            rc = panel(j0, std::min(j0 + P, n));
        } catch (...) {
            if (out_pending_) { out_pending_ = false; try { side_.wait(); } catch (...) {} }
            throw;
        }
        if (rc != FPLUS_OK) {
            if (out_pending_) { out_pending_ = false; try { side_.wait(); } catch (...) {} }
            return rc;
        }
        if (opt_.verbose >= 2 && ((j0 / P) % (n / (20 * P) + 1) == 0))
            printf("[FPLUS]   [%5.1f%%] column %ld/%ld  supernodes=%ld (panel %ld)\n", 100.0 * (double)std::min(j0 + P, n) / (double)n,
                   (long)std::min(j0 + P, n), (long)n, (long)sn_.size(), (long)P);
    }
    {
        const int orc = join_out();
        if (orc != FPLUS_OK) return orc;
    }
    dev_sync(s_);   // the last panel's blocks and U values are out
    {
        // a workgroup of the cooperative panel LU that gave up at a barrier leaves this flag, whatever the others did
        int32_t gave_up = 0;
        d_coop_cnt_.download(&gave_up, 1, s_, 1);
        dev_sync(s_);
        if (gave_up) {
            F_.coop_aborted = true;
            errf(err_, errlen_, "internal: the cooperative panel LU's workgroups did not meet at a barrier");
            return FPLUS_ERR_INTERNAL;
        }
    }
    // the final structure (the inverse forms go with the driver unless a refactorization will replay this one: the
    // solves use the blocks' values only)
    if (rec_) {
        rec_->iv_chunks = std::move(iv_chunks_);
        rec_->upload(s_);
    } else {
        for (BlockDesc& B : F_.blocks) B.ivals = nullptr;
    }
    F_.nsn = (int64_t)sn_.size();
    F_.sn_first.resize(F_.nsn + 1);
    F_.sn_rows.resize(F_.nsn);
    for (int64_t s = 0; s < F_.nsn; s++) { F_.sn_first[s] = sn_[s].c0; F_.sn_rows[s] = std::move(sn_[s].ri); }
    F_.sn_first[F_.nsn] = n;
    F_.p = p_;
    F_.pinv = pinv_;
    F_.Up = std::move(Up_);
    F_.Urp = std::move(Urp_);
    F_.Ur_row = std::move(Ur_row_);
    F_.Ur_len = std::move(Ur_len_);
    // factor flops c * sum_k l_k u_k (l_k below the diagonal of L(:, k), u_k right of the diagonal in U(k, :)); the
    // row counts of U from its runs by differences
    std::vector<int64_t> urow(nn + 1, 0);
    for (size_t r = 0; r < F_.Ur_row.size(); r++) { urow[F_.Ur_row[r]]++; urow[F_.Ur_row[r] + F_.Ur_len[r]]--; }
    for (int64_t i = 1; i < n; i++) urow[i] += urow[i - 1];
    double fl = 0;
    for (int64_t s = 0; s < F_.nsn; s++) {
        const int64_t c0 = F_.sn_first[s], c1 = F_.sn_first[s + 1], nr = (int64_t)F_.sn_rows[s].size();
        for (int64_t k = c0; k < c1; k++) fl += (double)(nr - (k - c0) - 1) * (double)(urow[k] - 1);
    }
    F_.flops = (vs_ == 1 ? 2.0 : 8.0) * fl;
    return FPLUS_OK;
}

}  // namespace

int sn_factor_device(int64_t n, const int64_t* Bp, const int64_t* Bi, const int64_t* d_Bp, const int32_t* d_Bi,
                     const double* d_Bx, int vs, const SnOptions& opt, DevFactor& F, char* err, size_t errlen)
{
    Driver d(n, Bp, Bi, d_Bp, d_Bi, d_Bx, vs, opt, F, err, errlen);
    return d.run();
}

namespace {
template <class T> void up_vec(DevBuf<T>& d, std::vector<T>& h, hipStream_t s)
{
    d.alloc(h.empty() ? 1 : h.size());
    d.upload(h.data(), h.size(), s);
    std::vector<T>().swap(h);
}
}  // namespace

Replay::~Replay()
{
    if (gexec) (void)hipGraphExecDestroy(gexec);
    for (hipEvent_t e : ev_pre) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_done) (void)hipEventDestroy(e);
    if (s_side) (void)hipStreamDestroy(s_side);
}

void Replay::upload(hipStream_t s)
{
    // the levels a pipelined replay may run while the previous panel finishes: those before the first level with a
    // tile of a block the previous panel made (its values and inverse form are written at that panel's end)
    for (size_t p = 0; p < panels.size(); p++) {
        ReplayPanel& P = panels[p];
        P.npre = p == 0 ? P.nlev : 0;
        if (p == 0) continue;
        const int64_t b0 = panels[p - 1].blk0, b1 = panels[p - 1].blk1;
        int64_t L = P.lev_off;
        for (; L < P.lev_off + P.nlev; L++) {
            const ReplayLevel& V = levels[L];
            bool dep = false;
            for (int64_t t = V.tile0; t < V.tile0 + V.ntile && !dep; t++) dep = h_tiles[t].block >= b0 && h_tiles[t].block < b1;
            if (dep) break;
        }
        P.npre = L - P.lev_off;
    }
    up_vec(d_rows, h_rows, s);
    up_vec(d_cols, h_cols, s);
    up_vec(d_expos, h_expos, s);
    up_vec(d_tasks, h_tasks, s);
    up_vec(d_tiles, h_tiles, s);
    up_vec(d_ex, h_ex, s);
    up_vec(d_ij, h_ij, s);
    up_vec(d_it, h_it, s);
    up_vec(d_runs, h_runs, s);
    up_vec(d_newblk, h_newblk, s);
}

size_t Replay::device_bytes() const
{
    size_t b = d_rows.n * 4 + d_cols.n * 4 + d_expos.n * 4 + d_tasks.n * sizeof(UpdTask) + d_tiles.n * sizeof(UpdTile) +
               d_ex.n * sizeof(ExtractCol) + d_ij.n * sizeof(int2) + d_it.n * sizeof(GemmTile) + d_runs.n * sizeof(URun);
    for (const DevBuf<double>& c : iv_chunks) b += c.n * sizeof(double);
    return b;
}

// The recorded factorization replayed on new values: per panel the workspace rows (lower ones in final pivot order),
// the scatter of the panel's columns, the update levels, the solved segments back, the panel LU without row exchanges,
// the new blocks and their inverse forms, U out -- the same launches as the factorization with pivoting, with none of
// its host work and no wait for the device until the end.
int sn_refactor_device(DevFactor& F, const int64_t* d_Bp, const int32_t* d_Bi, const double* d_Bx, double pivot_rel,
                       int static_pivots, int precision, const SelectorArgs& sel, int64_t* replaced, double* min_ratio, char* err,
                       size_t errlen)
{
    if (!F.replay) {
        errf(err, errlen, "the factorization was not recorded for refactorization");
        return FPLUS_ERR_BADARG;
    }
    Replay& R = *F.replay;
    const int vs = F.vs;
    const int64_t n = F.n;
    hipStream_t s = devctx().stream;
    R.d_W.reserve((size_t)R.max_w);
    R.d_X.reserve((size_t)R.max_x);
    R.d_colmax.reserve((size_t)R.max_np);
    if (R.d_rowmap.n < (size_t)std::max<int64_t>(n, 1)) {
        R.d_rowmap.alloc((size_t)std::max<int64_t>(n, 1));
        k::fill_i32(R.d_rowmap.p, n, -1, s);
    }
    if (!R.d_st.p) R.d_st.alloc(1);
    RefactorStatus st0;
    st0.err_col = 0;
    st0.nreplaced = 0;
    st0.min_ratio_bits = 0x7FF0000000000000ULL;   // +inf
    st0.ticket = 0;
    st0.pad = 0;
    R.d_st.upload(&st0, 1, s);
    if (precision == 1) k::fill_u8(F.d_blk_f32.p, (int64_t)F.blocks.size(), 1, s);
    else k::fill_u8(F.d_blk_f32.p, (int64_t)F.blocks.size(), 0, s);
    if ((precision == 2 || precision == 3)) {
        if (precision == 2 && !sel.d_model) { errf(err, errlen, "per-block precision needs a selector"); return FPLUS_ERR_BADARG; }
        const size_t nn = (size_t)std::max<int64_t>(n, 1);
        if (F.d_feat.n < nn * SEL_NF) F.d_feat.alloc(nn * SEL_NF);
        if (F.d_featmax.n < 3 * nn) {
            F.d_featmax.alloc(3 * nn);
            k::fill_u64(F.d_featmax.p, (int64_t)(3 * nn), 0ULL, s);
        }
        if (precision == 3) {
            if (F.d_blk_err.n < 2 * nn) F.d_blk_err.alloc(2 * nn);
            k::fill_u64(F.d_blk_err.p, (int64_t)(2 * nn), 0ULL, s);
        }
    }
    const BlockDesc* blocks = F.d_blocks.p;
    // one workspace set: the panel workspace W, the solved segments X, the column maxima, the row map
    struct Buf {
        double *W, *X, *colmax;
        int32_t* rowmap;
    };
    auto setup = [&](const ReplayPanel& P, const Buf& B, hipStream_t st) {
        const int32_t* rows = R.d_rows.p + P.rows_off;
        const int64_t nrow = P.Ru + P.Rl, ldw = P.ldw;
        k::rowmap_set(nrow, rows, B.rowmap, 0, st);
        k::fill_f64(B.W, ldw * P.np * vs, 0.0, st);
        k::scatter_cols(vs, (int32_t)P.np, d_Bp, d_Bi, d_Bx, P.j0, B.rowmap, B.W, ldw, B.colmax, st);
    };
    auto levels = [&](const ReplayPanel& P, int64_t L0, int64_t L1, const Buf& B, hipStream_t st) {
        const int32_t* cols = R.d_cols.p + P.cols_off;
        for (int64_t L = L0; L < L1; L++) {
            const ReplayLevel& V = R.levels[L];
            k::blk_upd(vs, V.ntile, R.d_tiles.p + V.tile0, cols, B.rowmap, B.W, P.ldw, V.rt, V.flops, B.X, precision,
                       F.d_blk_f32.p, F.d_blk_err.p, st);
        }
    };
    auto tail = [&](const ReplayPanel& P, const Buf& B, hipStream_t st) {
        const int32_t* rows = R.d_rows.p + P.rows_off;
        const int64_t nrow = P.Ru + P.Rl, ldw = P.ldw;
        const int32_t* cols = R.d_cols.p + P.cols_off;
        if (P.ntask) k::x_back(vs, (int32_t)P.ntask, R.d_tasks.p + P.tasks_off, cols, blocks, B.W, ldw, B.X, st);
        double* Wl = B.W + P.Ru * vs;
        for (int64_t a = 0; a < P.np; a += FPA_PANEL_NB) {
            const int32_t nb = (int32_t)std::min<int64_t>(FPA_PANEL_NB, P.np - a);
            k::lu_nopiv(vs, Wl, ldw, (int32_t)P.Rl, (int32_t)P.np, (int32_t)a, nb, B.colmax, pivot_rel, static_pivots, P.j0,
                        R.d_st.p, st);
            k::gemm_sub(vs, Wl, ldw, (int32_t)P.Rl, (int32_t)P.np, (int32_t)a, nb, nullptr, st);
        }
        if (P.nex) k::extract_cols(vs, (int32_t)P.nex, R.d_ex.p + P.ex_off, R.d_expos.p + P.expos_off, rows + P.Ru, Wl, ldw, st);
        if (P.nij)
            k::blk_inverse(vs, (int32_t)P.nij, R.d_ij.p + P.ij_off, (int32_t)P.nit, R.d_it.p + P.it_off, blocks, P.iflops, st);
        if ((precision == 2 || precision == 3) && P.nnb) {
            k::block_features(vs, (int32_t)P.nnb, R.d_newblk.p + P.nb_off, blocks, B.colmax, P.j0, sel.d_Dr, sel.d_Dc, n,
                              F.d_featmax.p, F.d_feat.p, st);
            if (precision == 2)
                k::sel_infer(P.nnb, R.d_newblk.p + P.nb_off, F.d_feat.p, sel.d_model, sel.threshold, nullptr, F.d_blk_f32.p, st);
        }
        if (P.nruns) k::u_runs(vs, (int32_t)P.nruns, R.d_runs.p + P.runs_off, B.W, F.d_Ux.p, F.d_Ui.p, st);
        k::rowmap_set(nrow, rows, B.rowmap, 1, st);
    };
    const Buf B0{R.d_W.p, R.d_X.p, R.d_colmax.p, R.d_rowmap.p};
    auto replay = [&] {
        for (const ReplayPanel& P : R.panels) {
            setup(P, B0, s);
            levels(P, P.lev_off, P.lev_off + P.nlev, B0, s);
            tail(P, B0, s);
        }
    };
    // the pipelined replay: panel p + 1's set-up and its levels that read no block of panel p run on a second stream
    // (with the second workspace set) while panel p finishes on the first; panel p + 1 waits for panel p - 1's end
    // (its workspace set free, the blocks it reads made), and its remaining levels wait for its first ones
    auto replay_pipelined = [&] {
        const size_t np = R.panels.size();
        if (!R.s_side) hip_check(hipStreamCreateWithFlags(&R.s_side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
        while (R.ev_pre.size() < np + 1) {
            hipEvent_t a, b;
            hip_check(hipEventCreateWithFlags(&a, hipEventDisableTiming), "hipEventCreateWithFlags");
            hip_check(hipEventCreateWithFlags(&b, hipEventDisableTiming), "hipEventCreateWithFlags");
            R.ev_pre.push_back(a);
            R.ev_done.push_back(b);
        }
        const Buf B1{R.d_W2.p, R.d_X2.p, R.d_colmax2.p, R.d_rowmap2.p};
        hipStream_t side = R.s_side;
        hip_check(hipEventRecord(R.ev_pre[np], s), "hipEventRecord");   // the set-up above is done
        hip_check(hipStreamWaitEvent(side, R.ev_pre[np], 0), "hipStreamWaitEvent");
        const ReplayPanel& P0 = R.panels[0];
        setup(P0, B0, s);
        levels(P0, P0.lev_off, P0.lev_off + P0.npre, B0, s);
        for (size_t p = 0; p < np; p++) {
            const Buf& Bp = (p & 1) ? B1 : B0;
            if (p + 1 < np) {
                const ReplayPanel& Q = R.panels[p + 1];
                const Buf& Bq = ((p + 1) & 1) ? B1 : B0;
                if (p >= 1) hip_check(hipStreamWaitEvent(side, R.ev_done[p - 1], 0), "hipStreamWaitEvent");
                setup(Q, Bq, side);
                levels(Q, Q.lev_off, Q.lev_off + Q.npre, Bq, side);
                hip_check(hipEventRecord(R.ev_pre[p + 1], side), "hipEventRecord");
            }
            const ReplayPanel& P = R.panels[p];
            if (p >= 1) hip_check(hipStreamWaitEvent(s, R.ev_pre[p], 0), "hipStreamWaitEvent");
            levels(P, P.lev_off + P.npre, P.lev_off + P.nlev, Bp, s);
            tail(P, Bp, s);
            hip_check(hipEventRecord(R.ev_done[p], s), "hipEventRecord");
        }
    };
    // the panel loop as launched, or through the graph captured from it (no host work and no allocation inside it:
    // every buffer is reserved above, the status reset and read around it)
    const bool use_graph = refactor_graph() && !kprof().on;
    if (use_graph) {
        uint64_t thr_bits, pr_bits;
        std::memcpy(&thr_bits, &sel.threshold, sizeof(float));
        thr_bits &= 0xffffffffull;
        std::memcpy(&pr_bits, &pivot_rel, 8);
        const std::vector<uint64_t> key = {
            (uint64_t)precision, (uint64_t)static_pivots, pr_bits, thr_bits, (uint64_t)sel.d_model, (uint64_t)sel.d_Dr,
            (uint64_t)sel.d_Dc, (uint64_t)d_Bp, (uint64_t)d_Bi, (uint64_t)d_Bx, (uint64_t)R.d_W.p, (uint64_t)R.d_X.p,
            (uint64_t)R.d_colmax.p, (uint64_t)R.d_rowmap.p, (uint64_t)R.d_st.p, (uint64_t)F.d_blk_f32.p,
            (uint64_t)F.d_blk_err.p, (uint64_t)F.d_feat.p, (uint64_t)F.d_featmax.p, (uint64_t)F.d_Ux.p, (uint64_t)F.d_Ui.p,
            (uint64_t)blocks};
        if (!R.gexec || R.gkey != key) {
            if (R.gexec) {
                (void)hipGraphExecDestroy(R.gexec);
                R.gexec = nullptr;
            }
            hipGraph_t graph = nullptr;
            hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            try {
                replay();
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            hip_check(hipStreamEndCapture(s, &graph), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&R.gexec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(e, "hipGraphInstantiate");
            R.gkey = key;
        }
        hip_check(hipGraphLaunch(R.gexec, s), "hipGraphLaunch");
    } else if (refactor_pipeline() && R.panels.size() > 1) {
        R.d_W2.reserve((size_t)R.max_w);
        R.d_X2.reserve((size_t)R.max_x);
        R.d_colmax2.reserve((size_t)R.max_np);
        if (R.d_rowmap2.n < (size_t)std::max<int64_t>(n, 1)) {
            R.d_rowmap2.alloc((size_t)std::max<int64_t>(n, 1));
            k::fill_i32(R.d_rowmap2.p, n, -1, s);
        }
        replay_pipelined();
    } else {
        replay();
    }
    RefactorStatus st;
    R.d_st.download(&st, 1, s);
    dev_sync(s);
    F.precision = precision;
    if (replaced) *replaced = st.nreplaced;
    double mr;
    std::memcpy(&mr, &st.min_ratio_bits, 8);
    if (min_ratio) *min_ratio = mr;
    if (st.err_col) {
        errf(err, errlen, "refactorization in the recorded pivot order: the pivot of column %ld fell below pivot_rel times its "
             "column's largest entry; factor with pivoting", (long)(st.err_col - 1));
        return FPLUS_ERR_SINGULAR;
    }
    return FPLUS_OK;
}

}  // namespace fpa
