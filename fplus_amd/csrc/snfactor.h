// snfactor.h -- the supernodal partial-pivoting factorization: host symbolic driver + device numerics.
#pragma once
#include "device_api.h"
#include "factor_dev.h"

#include <cstdint>
#include <memory>
#include <vector>

namespace fpa {

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
    // the dependency levels by tile count (bucket b: at most 2^b tiles, the last one more): levels, tiles, flops issued
    // on the matrix cores, and the sum of each level's longest segment
    struct LevHist { int64_t levels = 0, tiles = 0; double issued = 0, maxm = 0; } lev_hist[13];
    int64_t rows_upper = 0, rows_lower = 0, rows_lower_max = 0;   // workspace rows summed over the panels
    int64_t shared_levels = 0;   // dependency levels whose tasks can write the same rows (their updates are atomic)
    int64_t max_block_w = 0;     // the widest block
    bool coop_aborted = false;   // the cooperative panel LU gave up at a barrier (the factorization failed; retry without it)
    bool coop_retried = false;   // this factorization is the retry with the one-workgroup panel LU
};

struct SnOptions {
    int64_t relax_rows = 32;
    double pivot_rel = 1.4901161193847656e-08;
    int64_t panel = 256;
    int verbose = 0;
    bool coop_panel_lu = true;   // false: the one-workgroup panel LU for every sub-panel
};

// factor the n x n CSC matrix whose values are on the device (d_Bx, vs doubles per entry); the host holds its
// pattern (Bp, Bi); d_Bp / d_Bi are the same pattern on the device.  Throws GpuError / std::bad_alloc; returns
// FPLUS_OK or FPLUS_ERR_SINGULAR (message in err).
int sn_factor_device(int64_t n, const int64_t* Bp, const int64_t* Bi, const int64_t* d_Bp, const int32_t* d_Bi,
                     const double* d_Bx, int vs, const SnOptions& opt, DevFactor& F, char* err, size_t errlen);

}  // namespace fpa
