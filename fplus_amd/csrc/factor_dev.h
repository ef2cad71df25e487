// factor_dev.h -- device side of the supernodal factorization: block / task descriptors and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define FPA_ROW_TILE 64    // rows per update GEMM tile
#define FPA_COL_TILE 64    // panel columns per update GEMM tile
#define FPA_PANEL_NB 16    // sub-panel width of the dense panel LU

namespace fpa {

inline int col_tile(int vs) { return vs == 1 ? 64 : 32; }   // panel columns per tile of the panel-LU GEMM

// A block: the columns [c0, c0 + w) of one supernode computed in one panel, stored column-major over its rows (ld =
// nrows): the pivot rows of its columns first, in column order, then the supernode's other rows as they stood at the
// end of that panel.  Values are the L factor below the unit diagonal (the entries above it are U values, unused by
// the updates).  rows hold row indices of the permuted matrix.  ivals, same shape, is the inverse form the updates use:
// [L11^-1 ; L21 L11^-1] with L11 the unit lower triangle of the pivot rows and L21 the rows below (needed while
// factoring only); linv keeps L11^-1 alone (w x w, ld = w) for the solves, and uinv U11^-1 (w x w, ld = w), U11 the
// upper triangle of the pivot rows with the diagonal.
struct BlockDesc {
    double* vals;      // nrows * w scalars (a scalar is 1 or 2 doubles)
    double* ivals;     // nrows * w scalars
    double* linv;      // w * w scalars
    double* uinv;      // w * w scalars
    int32_t* rows;     // nrows
    int32_t nrows, w;
    int32_t c0, sn;
};

// one block applied to some panel columns: the block's pivot rows [k0, w) (the segment the columns reach) are solved
// with the block's unit lower triangle and the rows [w, nrows) updated by the rectangle below, both as one product with
// the block's inverse form (blk_upd)
struct UpdTask {
    int32_t block, k0;
    int32_t col_off, ncol;   // panel-local column indices at cols[col_off .. col_off + ncol)
    int64_t xoff;            // the solved segment, compact: X[(k - k0) + c * (w - k0)] at xbuf + xoff scalars
    int32_t spos0, pad;      // the W row of the segment's first row (the segment's rows are consecutive in W)
};

// a tile of a new block's inverse form (blk_inverse): rows [w + row0, ...) and columns [col0, col0 + 64) of block `task`
struct GemmTile {
    int32_t task, row0, col0, pad;
};

// a tile of a task's update (blk_upd) with all its workgroup needs in one 64-byte read: rows [row0, row0 + tile rows)
// of the block, the task's columns cols[cfirst .. cfirst + nc) (nc <= 64)
#define FPA_TILE_SHARED 1   // another task of the launch may write the tile's rows below the triangle: atomic updates
struct alignas(16) UpdTile {
    const double* iv;      // the block's inverse form (ld = nrows)
    const int32_t* rows;   // the block's rows
    int64_t xo;            // the task's solved segment from the tile's first column on, in scalars from the launch's X:
                           // X[xo + (row - k0) + c * (w - k0)]
    int32_t nrows, w, k0, row0;
    int32_t cfirst, nc;
    int32_t spos0;         // the W row of the segment's first row (the segment's rows are consecutive in W)
    int32_t flags;         // FPA_TILE_SHARED
    int32_t block;         // the block (its entry of the per-block precision marks)
    int32_t pad;
};
static_assert(sizeof(UpdTile) == 64, "UpdTile is one 64-byte read");

// one run of U entries out of the panel workspace: Ui[uoff + t] = row0 + t, Ux[uoff + t] = W[woff + t], t < len
struct URun {
    int64_t uoff, woff;
    int32_t row0, len;
};

struct PanelStatus {
    int32_t err;         // 0, or 1 + the panel-local column of a singular pivot
    int32_t ties;        // with err 0: the sub-panel's columns whose largest candidate modulus two rows or more held
    double best, thr;    // the failing column's best modulus and threshold
};

// the panel LU's status and pivot positions, together: one copy brings both back to the host
#define FPA_PANEL_MAX 256   // the widest panel
struct PanelOut {
    PanelStatus st;
    int32_t piv[FPA_PANEL_MAX];
};

// a refactorization's status: the first column (1 + global index) whose pivot fell below pivot_rel * colmax with
// pivots not replaced, the pivots replaced (static pivoting), and min |pivot| / colmax over the columns (double bits)
struct RefactorStatus {
    int32_t err_col;
    int32_t nreplaced;
    unsigned long long min_ratio_bits;
    uint32_t ticket;     // lu_nopiv: its workgroups' arrivals after reading the pivot block (0 between launches)
    uint32_t pad;
};

// one block column copied out of the panel workspace: dst[r] = Wl[pos[pos_off + r] + ci * ldw], r < nrows; for a
// block's first column also its row indices, rdst[r] = lrows[pos[pos_off + r]] (lrows: the row at each lower position)
struct ExtractCol {
    double* dst;
    int32_t* rdst;     // or null
    int64_t pos_off;
    int32_t nrows, ci;
};

namespace k {
void extract_cols(int vs, int32_t ncols, const ExtractCol* d, const int32_t* pos, const int32_t* lrows, const double* Wl,
                  int64_t ldw, hipStream_t s);
void rowmap_set(int64_t cnt, const int32_t* rows, int32_t* rowmap, int reset, hipStream_t s);
void scatter_cols(int vs, int32_t ncol, const int64_t* Bp, const int32_t* Bi, const double* Bx, int64_t j0,
                  const int32_t* rowmap, double* W, int64_t ldw, double* colmax, hipStream_t s);
void u_runs(int vs, int32_t nruns, const URun* runs, const double* W, double* Ux, int32_t* Ui, hipStream_t s);
// the inverse forms of new blocks: jobs (block, column) for L11^-1, then the tiles of the rows below
void blk_inverse(int vs, int32_t njobs, const int2* jobs, int32_t ntile, const GemmTile* tiles, const BlockDesc* blocks,
                 double flops, hipStream_t s);
// U11^-1 of blocks, jobs (block, column): only the solves read it, so it is made for every block at once
void blk_uinv(int vs, int64_t njobs, const int2* jobs, const BlockDesc* blocks, double flops, hipStream_t s);
// the tiles of one dependency level: each task's solved segment into its slice of X, its rows below subtracted from W
// through rowmap (atomically for tiles marked shared); tile_rows: 64, or 32 (tiles of half the rows, for levels with
// few tiles)
// X: the solved segments' buffer the tiles' xo index; prec_mode 0: every tile's product in FP64, 1: every tile's in
// FP32, 2: by its block's mark blk_f32[tile.block] (nonzero: FP32), set on the device (e.g. by the precision selector);
// 3: FP64 applied, and per block the largest difference of the FP32 product and the largest FP64 value max-accumulated
// into blk_err[2 b], blk_err[2 b + 1] (the selector's training logs); 4: every tile's as three BF16 products of the
// operands' hi/lo split with FP32 accumulation
void blk_upd(int vs, int32_t ntile, const UpdTile* tiles, const int32_t* cols, const int32_t* rowmap, double* W, int64_t ldw,
             int tile_rows, double flops, double* X, int prec_mode, const uint8_t* blk_f32, unsigned long long* blk_err,
             hipStream_t s);
// the solved segments of all tasks back into W's rows
void x_back(int vs, int32_t ntask, const UpdTask* tasks, const int32_t* cols, const BlockDesc* blocks, double* W, int64_t ldw,
            const double* xbuf, hipStream_t s);
// dense partial-pivot LU of the panel's lower section (nr x P, column-major, ld), sub-panel [a, a + nb), and the
// sub-panel's pivot rows in the columns right of it solved with its unit lower triangle: a cooperative launch over the rows (coop: getf2_coop_bytes(vs) of scratch; counter: one int, zero at first and then only touched
// here, with *counted the arrivals it holds), one workgroup for very tall panels
// hout: the sub-panel's slot in page-locked host memory (a device address of it), where the kernel writes the pivots
// of the sub-panel's columns and the status (err 0 once the sub-panel is factored): no copy back
// counter: two ints, zero at first: [0] the barrier counter, [1] set (and never cleared) when a workgroup gave up at a
// barrier; allow_coop false: the one-workgroup kernel for every sub-panel
void getf2(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const double* colmax,
           double pivot_rel, int32_t* piv, PanelStatus* st, PanelOut* hout, void* coop, int32_t* counter, int64_t* counted,
           bool allow_coop, hipStream_t s);
size_t getf2_coop_bytes(int vs);
int getf2_coop_max_wg();
void set_getf2_coop_max_wg(int cap);   // at most 512; 0: the one-workgroup kernel for every sub-panel
// (getf2's own second step after its one-workgroup form)
void trsm_sub(int vs, double* Wl, int64_t ld, int32_t P, int32_t a, int32_t nb, const PanelStatus* st, hipStream_t s);
void gemm_sub(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const PanelStatus* st,
              hipStream_t s);
// sub-panel [a, a + nb) of the panel LU with the rows already in pivot order (no exchanges), and the sub-panel's rows in
// the columns right of it; mode 0: a small pivot is recorded in st, 1: replaced and counted (j0: the panel's first column)
void lu_nopiv(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const double* colmax,
              double pivot_rel, int mode, int64_t j0, RefactorStatus* st, hipStream_t s);
}  // namespace k

}  // namespace fpa
