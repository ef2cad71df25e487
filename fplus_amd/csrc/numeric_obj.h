// numeric_obj.h -- the numeric object behind fplus_numeric*: the device factorization, the composed orders and
// scalings, what the solves, a refactorization and the refined solves keep.
#pragma once
#include "fplus.h"
#include "fplus_amd.h"
#include "host_common.h"
#include "snfactor.h"
#include "symbolic.h"

#include <memory>
#include <mutex>
#include <vector>

using namespace fpa;

struct fplus_numeric {
    int64_t n = 0;
    int is_complex = 0;
    // what a refactorization and the refinement need: the analysed pattern, the orders, the options, the scalings in
    // the analysed order (device)
    std::shared_ptr<PatternCache> cache;
    std::vector<int64_t> rp, cp;
    fplus_options opt{};
    fplus_amd_options aopt{};
    DevBuf<double> d_Drp, d_Dcp;
    int64_t refactor_replaced = 0;   // the last refactorization's replaced pivots
    double refactor_min_ratio = 0;   // and its smallest |pivot| / colmax
    std::vector<int64_t> p, q;
    std::vector<double> Dr, Dc;
    DevFactor F;
    int ruiz_iters = 0;
    // solve data on the device
    DevBuf<int64_t> d_p, d_q, d_Up;
    DevBuf<int32_t> d_pinv;
    DevBuf<double> d_Dr, d_Dc;
    std::vector<int32_t> lev_lf, lev_ub, lev_utf, lev_ltb;   // level offsets into the block lists
    DevBuf<int32_t> d_blk_lf, d_blk_ub, d_blk_utf, d_blk_ltb;
    std::vector<int32_t> blk_lf, blk_ub, blk_utf, blk_ltb;
    // L y = b: every block's rows below have one contribution slot each; a row gathers the slots that hold its updates
    DevBuf<int64_t> d_slot_off, d_incp, d_incs;
    // the pivot rows of every L level's blocks (the rows the level gathers), level offsets
    std::vector<int32_t> lev_gr;
    DevBuf<int32_t> d_grows;
    std::vector<double> bytes_lf_gather;
    int64_t nslots = 0;
    std::vector<int32_t> lev_lft, lev_ubt, lev_utt, lev_ltt;   // level offsets into the push / dot-product tiles
    DevBuf<int2> d_lft, d_utt, d_ltt;
    DevBuf<int2> d_ujobs;                                       // (block, column) jobs of the blocks' U11^-1
    int64_t n_ujobs = 0;
    double uinv_flops = 0;
    bool valid = true;   // false after a failed refactorization: the factors are not usable until the next success
    DevBuf<int4> d_ubt;        // U x = y dot-product tiles: block, first piece, pieces, partial slot
    std::vector<int32_t> lev_ubr;
    DevBuf<int4> d_ubr;        // and the partial sums' reduction tiles: block, first of 16 rows, first slot, slots
    int64_t ub_slots_max = 0;  // the most partial slots of one level
    std::vector<double> bytes_ub_prod;
    std::vector<int32_t> lev_lpt, lev_upt, lev_utp, lev_ltp;   // level offsets into the triangle-product tiles
    DevBuf<int2> d_lpt, d_upt, d_utp, d_ltp;                    // (block, first of 64 rows), per sweep
    std::vector<double> bytes_lf_pull, bytes_lf_push, bytes_ub, bytes_utf, bytes_ltb;
    // U x = y beyond the blocks: a block's pieces are the runs of later columns over its rows, each covering the block's
    // rows a .. w-1 (a = pa[p]), in column order; piece p holds U(c0 + t, pj[p]) at Ux[pu[p] + t]
    DevBuf<int64_t> d_pu;
    DevBuf<int32_t> d_pj, d_pa;
    // the sweeps with A as persistent launches (solve_dev.h: lsweep_df / usweep_df): their item lists, per block the
    // blocks each waits for, the U groups' first row, and the U groups' rows in all
    DevBuf<int4> d_dfl_items, d_dfu_items;
    int32_t dfl_nitems = 0, dfu_nitems = 0;
    DevBuf<int32_t> d_ldep_off, d_ldep, d_udep_off, d_udep;
    DevBuf<int64_t> d_ubase;
    int64_t df_part = 0;
    double df_bytes_l = 0, df_bytes_u = 0;
    // solve workspace, reused across calls (solves on one numeric object are serialized)
    mutable std::mutex solve_mu;
    mutable DevBuf<double> d_w, d_y, d_contrib, d_b, d_ub_part;
    mutable DevBuf<int32_t> d_dfcnt;
    // the CSC export, built on first access
    mutable std::mutex mu;
    mutable bool exported = false;
    mutable fplus_factors f{};
    // the refined solves: the equilibrated matrix Dr B Dc in the analysed order (CSC values, as the factorization saw
    // them), the row pivot order p_lu, and their workspace
    DevBuf<double> d_Abar;
    DevBuf<int64_t> d_plu;
    mutable DevBuf<double> ir_V, ir_w, ir_bb, ir_u, ir_r, ir_part, ir_d, ir_c, ir_sc;
    mutable DevBuf<unsigned long long> ir_norm;
    // a factorization by the row kernel (rook / full pivoting, or drop_tol > 0): its factors in f (CSC, exported from
    // the start) and the solve plan of their patterns; everything above that belongs to the device factorization stays
    // empty, and the extensions that need it refuse such a handle
    bool row = false;
    fplus_solve_plan* rplan = nullptr;
    // timings of the last factorization
    double t_ruiz = 0, t_factor = 0, t_total = 0, t_prepare = 0;
    double t_prep_part[6] = {0, 0, 0, 0, 0, 0};   // the preparation by step: L levels, U pieces, U levels, slots, tiles, uploads
};


namespace fpa {
// the extensions that need the device factorization refuse a row-kernel handle with this
inline int row_handle_error(char* err, size_t errlen)
{
    errf(err, errlen, "this factorization comes from the row kernel (rook / full pivoting or drop_tol > 0): fplus_solve and the "
                      "factor accessors take it, the device-factorization extensions do not");
    return FPLUS_ERR_BADARG;
}
// the selector's device model and threshold with the scalings its features read (selector.cpp)
SelectorArgs selector_args(const fplus_amd_selector* sel, const double* d_Dr, const double* d_Dc);
// the sweeps on w (n x R row-major, pivot order, R <= solve_rhs_per_pass) in place: trans 0 L U x = w, the result in
// column order; 1 / 2 with U^T L^T (conjugated for 2)
void device_sweeps(const fplus_numeric* num, double* w, int32_t R, int trans, hipStream_t s);
// the sweeps' scratch: the intermediate vector y (n x R), the forward sweep's contribution slots and the backward
// sweep's partial sums (sizes in doubles); their contents need no clearing between calls (every kernel writes what it
// reads before reading it)
struct SweepScratch {
    double* y = nullptr;
    double* contrib = nullptr;
    double* part = nullptr;
    int32_t* cnt = nullptr;   // the persistent sweeps' counters (cleared by the sweeps themselves)
    size_t n_y = 0, n_contrib = 0, n_part = 0, n_cnt = 0;   // (n_cnt in ints)
};
// the sweeps with A as persistent launches (tuning knob "solve_dataflow"), else (the default) one launch per level and
// step
bool solve_dataflow();
void set_solve_dataflow(bool on);
SweepScratch sweep_scratch_size(const fplus_numeric* num, int32_t R);
// the no-allocation solve's array (noalloc.hip): its size, and the bytes a call may use (0: all of it)
size_t noalloc_scratch_bytes();
bool set_noalloc_limit(size_t bytes);
void device_sweeps(const fplus_numeric* num, double* w, int32_t R, int trans, const SweepScratch& z, hipStream_t s);
// solve for nrhs right-hand sides in device memory b (column-major, ld), in passes of solve_rhs_per_pass columns
void solve_on_device(const fplus_numeric* num, double* b, int64_t nrhs, int64_t ldb, int tr, hipStream_t s);
}  // namespace fpa
