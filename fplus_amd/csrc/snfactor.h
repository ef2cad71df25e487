// snfactor.h -- the supernodal partial-pivoting factorization: host symbolic driver + device numerics.
#pragma once
#include "device_api.h"
#include "factor_dev.h"

#include <cstdint>
#include <memory>
#include <vector>

namespace fpa {

// What a same-pattern refactorization replays: every panel's device work of a factorization with pivoting, recorded as
// it ran (the workspace rows with the lower ones in their final pivot order, the update tasks and tiles by dependency
// level, the new blocks' columns and inverse-form jobs, the runs of U), kept on the device with the blocks' inverse
// forms, so that new values with the same pattern factor in the same pivot order with no host work.
struct ReplayLevel {
    int64_t tile0;       // into the tiles
    int32_t ntile, rt;   // tiles, tile rows
    double flops;
};
struct ReplayPanel {
    int64_t j0, np, Ru, Rl, ldw, xtot;
    int64_t rows_off;              // Ru + Rl workspace rows: the U rows, then the lower rows in final (pivot) order
    int64_t cols_off, tasks_off, ntask, lev_off, nlev;
    int64_t ex_off, nex, expos_off;
    int64_t ij_off, nij, it_off, nit;
    double iflops;
    int64_t runs_off, nruns;
    int64_t nb_off, nnb;           // the panel's new blocks (their features and precision marks)
    int64_t blk0, blk1;            // the blocks the panel made at its end (ids blk0 .. blk1 - 1)
    int64_t npre;                  // the update levels before the first one that reads a block of the previous panel
};
struct Replay {
    std::vector<ReplayPanel> panels;
    std::vector<ReplayLevel> levels;
    // filled while recording, uploaded once at the end of the factorization, then released
    std::vector<int32_t> h_rows, h_cols, h_expos;
    std::vector<UpdTask> h_tasks;
    std::vector<UpdTile> h_tiles;
    std::vector<ExtractCol> h_ex;
    std::vector<int2> h_ij;
    std::vector<GemmTile> h_it;
    std::vector<URun> h_runs;
    std::vector<int32_t> h_newblk;
    DevBuf<int32_t> d_rows, d_cols, d_expos, d_newblk;
    DevBuf<UpdTask> d_tasks;
    DevBuf<UpdTile> d_tiles;
    DevBuf<ExtractCol> d_ex;
    DevBuf<int2> d_ij;
    DevBuf<GemmTile> d_it;
    DevBuf<URun> d_runs;
    std::vector<DevBuf<double>> iv_chunks;   // the blocks' inverse forms (rewritten by every refactorization)
    int64_t max_w = 1, max_x = 1, max_np = 1;  // workspace, solved-segment buffer (scalars), widest panel
    // a refactorization's workspace, kept across refactorizations; the pipelined replay's second set and its stream
    DevBuf<double> d_W, d_X, d_colmax, d_W2, d_X2, d_colmax2;
    DevBuf<int32_t> d_rowmap, d_rowmap2;
    hipStream_t s_side = nullptr;
    std::vector<hipEvent_t> ev_pre, ev_done;
    DevBuf<RefactorStatus> d_st;
    // the captured replay (refactor_graph) and what it was captured with
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> gkey;
    ~Replay();
    void upload(hipStream_t s);
    size_t device_bytes() const;
};

// the factorization as the device holds it, plus the host structure the solves and the CSC export need
struct DevFactor {
    int64_t n = 0;
    int vs = 1;                          // doubles per scalar: 1 real, 2 complex
    int64_t nsn = 0;
    // host structure
    std::vector<int64_t> p;              // p[j] = pivot row of column j (rows of the matrix factored)
    std::vector<int64_t> pinv;
    std::vector<int64_t> sn_first;       // supernode s: columns sn_first[s] .. sn_first[s+1]-1
    std::vector<std::vector<int64_t>> sn_rows;   // final rows of each supernode (pivot rows first)
    std::vector<BlockDesc> blocks;       // host mirror of the block table (device pointers)
    std::vector<std::vector<int32_t>> block_rows;   // host copy of every block's rows
    std::vector<int64_t> Up;             // U in CSC, column j: earlier supernodes' segments, the diagonal last
    // U's row indices as runs of consecutive rows (pivot order): column j holds runs Urp[j] .. Urp[j+1]-1, run r the
    // rows Ur_row[r] .. Ur_row[r] + Ur_len[r] - 1, in storage order (a segment is a run: a supernode's pivot rows)
    std::vector<int64_t> Urp;
    std::vector<int32_t> Ur_row, Ur_len;
    // device
    DevBuf<BlockDesc> d_blocks;
    std::vector<DevBuf<double>> val_chunks;
    std::vector<DevBuf<int32_t>> row_chunks;
    std::vector<DevBuf<double>> linv_chunks;   // the blocks' L11^-1, for the solves
    std::vector<DevBuf<double>> uinv_chunks;   // and their U11^-1
    DevBuf<double> d_Ux;                 // U values, in Ui order
    DevBuf<int32_t> d_Ui;                // U row indices on the device, expanded from the runs panel by panel
    int64_t nnzU = 0;
    // statistics
    double flops = 0, gemm_flops = 0, gemm_bytes = 0;
    int64_t panels = 0, levels = 0, tasks = 0;
    double t_host = 0, t_wait = 0;
    double t_wait_lu = 0;   // of the wait: for the panel LU's sub-panels after the first (the rest is the updates)
    double t_p1 = 0, t_tasks = 0, t_p3 = 0, t_out = 0;   // host phases: search, tasks and levels, symbolic with the pivots, blocks and U out
    double t_p3_dfs = 0, t_p3_u = 0, t_p3_join = 0, t_p3_prune = 0;   // the symbolic with the pivots, by step
    double t_p3_split = 0;   // of the search: the pass over the column's candidates
    double t_ws = 0, t_lev = 0;   // within the tasks and levels: the union order and workspace rows, the tasks' levels
    double t_ws_low = 0, t_lev_pass = 0;   // of which: the rows not yet pivotal; the sequential pass of the levels
    // the block updates' products: their flops, what the tiles issue on the matrix cores, the flops by task width
    // (columns <= 8, 16, 32, 64, more)
    double upd_prod_flops = 0, upd_mfma_flops = 0, upd_ncol_flops[5] = {0, 0, 0, 0, 0};
    std::vector<double> blk_upd_flops;   // per block (source of updates): its update products' algorithmic flops
    // the dependency levels by tile count (bucket b: at most 2^b tiles, the last one more): levels, tiles, flops issued
    // on the matrix cores, and the sum of each level's longest segment
    struct LevHist { int64_t levels = 0, tiles = 0; double issued = 0, maxm = 0; } lev_hist[13];
    int64_t rows_upper = 0, rows_lower = 0, rows_lower_max = 0;   // workspace rows summed over the panels
    int64_t shared_levels = 0;   // dependency levels whose tasks can write the same rows (their updates are atomic)
    int64_t pivot_ties = 0;      // columns whose pivot search met an exact tie for the largest modulus
    int64_t shared_level_tiles = 0, atomic_tiles = 0;   // the update tiles of such levels, and those that write with atomics
    int64_t max_block_w = 0;     // the widest block
    bool coop_aborted = false;   // the cooperative panel LU gave up at a barrier (the factorization failed; retry without it)
    bool coop_retried = false;   // this factorization is the retry with the one-workgroup panel LU
    DevBuf<uint8_t> d_blk_f32;   // per block: its updates of later columns on the FP32 matrix cores (precision mode 2)
    DevBuf<float> d_feat;        // per block: the selector's features (precision 2 and 3)
    DevBuf<unsigned long long> d_featmax;   // per block: 3 maxima the features read (zero between uses)
    DevBuf<unsigned long long> d_blk_err;   // per block: the logged FP32 differences and FP64 magnitudes (precision 3)
    std::unique_ptr<Replay> replay;   // kept when the factorization records what a refactorization replays
    int precision = 0;           // the updates' precision of the last (re)factorization: 0 FP64, 1 FP32, 2 per block
};

// what the precision selector needs while a factorization runs: the model (SEL_NW floats on the device) and its
// threshold, and the scalings of the analysed matrix (for the blocks' features)
struct SelectorArgs {
    const float* d_model = nullptr;
    float threshold = 0.5f;
    const double* d_Dr = nullptr;
    const double* d_Dc = nullptr;
};

struct SnOptions {
    int64_t relax_rows = 32;
    double pivot_rel = 1.4901161193847656e-08;
    int64_t panel = 256;
    int verbose = 0;
    bool coop_panel_lu = true;   // false: the one-workgroup panel LU for every sub-panel
    int precision = 0;           // the block updates: 0 FP64, 1 FP32, 2 per block by DevFactor::d_blk_f32 (set by the
                                 // precision selector as the blocks are made), 3 FP64 with the FP32 differences logged
    bool record = false;         // keep what a same-pattern refactorization replays (DevFactor::replay)
    SelectorArgs sel;            // precision 2 and 3
};

// the order of the update tiles within a level: XCD-aware (the row tiles reading one segment slice on one XCD; the
// default) or as built (tuning knob "xcd_tile_order")
bool xcd_tile_order();
void set_xcd_tile_order(bool on);
// refactorizations replayed through a captured HIP graph (tuning knob "refactor_graph"): the panel loop's launches are
// captured once per set of buffers and arguments and the graph is launched for later refactorizations
bool refactor_graph();
// refactorizations replayed as a pipeline of two panels (tuning knob "refactor_pipeline", default on): a panel's update
// levels that read no block of the previous panel run on a second stream while the previous panel finishes
bool refactor_pipeline();
void set_refactor_pipeline(bool on);
void set_refactor_graph(bool on);
// the update tile height per dependency level (tuning knob "update_tile_policy", 0 to 4; see snfactor.cpp)
int update_tile_policy();
bool set_update_tile_policy(int p);
// on a dependency level of several update tasks, atomics only for the row tiles with rows another task of the level
// writes too (tuning knob "update_atomic_rows", on by default), else for every tile of the level
bool update_atomic_rows();
void set_update_atomic_rows(bool on);
// deterministic update levels (tuning knob "deterministic_updates", off by default): each dependency level split so that
// no two of its tasks write one workspace row, every row then updated by its tasks in task order (bit-reproducible
// factors, more launches); for factorizations made afterwards
bool deterministic_updates();
void set_deterministic_updates(bool on);

// a same-pattern refactorization: the values d_Bx (the recorded factorization's analysed, equilibrated pattern) factored
// in the recorded pivot order by replaying its device work; static_pivots 0: a pivot below pivot_rel * colmax fails it
// (FPLUS_ERR_SINGULAR, the caller factors with pivoting), 1: such pivots are replaced (counted in *replaced).
// precision as in SnOptions.  min_ratio: the smallest |pivot| / colmax.
int sn_refactor_device(DevFactor& F, const int64_t* d_Bp, const int32_t* d_Bi, const double* d_Bx, double pivot_rel,
                       int static_pivots, int precision, const SelectorArgs& sel, int64_t* replaced, double* min_ratio, char* err,
                       size_t errlen);

// factor the n x n CSC matrix whose values are on the device (d_Bx, vs doubles per entry); the host holds its
// pattern (Bp, Bi); d_Bp / d_Bi are the same pattern on the device.  Throws GpuError / std::bad_alloc; returns
// FPLUS_OK or FPLUS_ERR_SINGULAR (message in err).
int sn_factor_device(int64_t n, const int64_t* Bp, const int64_t* Bi, const int64_t* d_Bp, const int32_t* d_Bi,
                     const double* d_Bx, int vs, const SnOptions& opt, DevFactor& F, char* err, size_t errlen);

}  // namespace fpa
