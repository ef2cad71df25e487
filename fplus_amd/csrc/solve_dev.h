// solve_dev.h -- device triangular solves on the supernode blocks and U (kernel launchers).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor_dev.h"

namespace fpa {
namespace k {
// x: n x R row-major (x[i * R + r]), indexed by pivot order; R <= solve_rhs_per_pass(vs).  blk: the blocks of one
// level.  bytes: the algorithmic bytes of the launch (for the optional per-kernel timing).
int solve_rhs_per_pass(int vs);
// rows[k] (the pivot rows of a level's blocks) subtract the contributions in their slots: x_i -= sum over incs[incp[i] ..]
void lfwd_gather(int vs, int32_t nrow, const int32_t* rows, const int64_t* incp, const int64_t* incs, const double* contrib,
                 double* x, int32_t R, double bytes, hipStream_t s);
// the unit-lower triangles of a level's blocks as products with L11^-1 (tiles: block, first of 64 rows): y = L11^-1 x on
// the blocks' rows; then the rows below (tiles: block, first of 64 rows below) into the contribution slots from y
void lfwd_pull(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* x, double* y, int32_t R,
               double bytes, hipStream_t s);
void lfwd_push(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int64_t* slot_off, const double* y,
               double* contrib, int32_t R, double bytes, hipStream_t s);
// U x = y per level.  ub_dots tiles (block, first piece, pieces, slot): block row t gets, in slot `slot` of part
// (part[(slot * 256 + t) * 16 + r] scalars), the sum over the tile's pieces p with pa[p] <= t of U(t, pj[p]) =
// Ux[pu[p] + t] times x[pj[p]] (x: the solved rows of later blocks).  ub_red tiles (block, first of 16 rows, first slot,
// slots): y (the lower sweep's result) minus the block's slots.  ub_prod (tiles: block, first of 64 rows): x = U11^-1 y
// on the blocks' rows.
constexpr int UB_PC = 32;   // pieces per ub_dots tile, staged at once
void ub_dots(int vs, int32_t ntile, const int4* tiles, const BlockDesc* blocks, const int32_t* pa, const int32_t* pj,
             const int64_t* pu, const double* Ux, const double* x, double* part, int32_t R, double bytes, hipStream_t s);
void ub_red(int vs, int32_t ntile, const int4* tiles, const BlockDesc* blocks, const double* part, double* x, int32_t R,
            hipStream_t s);
void ub_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* y, double* x, int32_t R,
             double bytes, hipStream_t s);
// the transposed sweeps per level: the dot products (tiles: block, column) with the solved rows of other
// blocks (U^T: y of earlier blocks, subtracted in x; L^T: x of later blocks, subtracted in y), then the blocks' triangles
// as products with the kept inverses, transposed and conjugated with conj (tiles: block, first of 64 rows):
// ut_prod y = U11^-T x, lt_prod x = L11^-T y on the blocks' rows
void utdots(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
            const double* Ux, const double* y, double* x, int32_t R, int conj, double bytes, hipStream_t s);
void ltdots(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int32_t* pinv_rows, const double* x,
            double* y, int32_t R, int conj, double bytes, hipStream_t s);
void ut_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* x, double* y, int32_t R, int conj,
             double bytes, hipStream_t s);
void lt_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* y, double* x, int32_t R, int conj,
             double bytes, hipStream_t s);
// The sweeps with A as one persistent launch each (dataflow; solve_dev.hip describes the items and their waits).
// items: (type << 29 | block, a, b, c); L: type 0 gather tile (block, first row), 1 triangle tile (block, first row),
// 2 push tile (block, first row below); U: 0 dot-product group (block, first piece, pieces, group), 1 reduction (block,
// groups), 2 triangle tile (block, first row, 1 if the block has groups).  dep_off / dep: per block the blocks it waits for (L: those writing
// its pivot rows; U: those of its pieces' columns).  cnt: df_counters(nb) ints, zero at the start of a sweep pair;
// cnt[2] nonzero afterwards when a wait gave up (the results are then incomplete).  ubase: per block its groups' first
// row in part (groups of w rows of R scalars: group g's row t at (ubase[b] + g w + t) R).
inline int64_t df_counters(int64_t nb) { return 4 + 6 * nb; }
void lsweep_df(int vs, int32_t nitems, const int4* items, const BlockDesc* blocks, const int32_t* dep_off, const int32_t* dep,
               const int64_t* incp, const int64_t* incs, const int64_t* slot_off, double* w, double* y, double* contrib,
               int32_t* cnt, int32_t nb, int32_t R, double bytes, hipStream_t s);
void usweep_df(int vs, int32_t nitems, const int4* items, const BlockDesc* blocks, const int32_t* dep_off, const int32_t* dep,
               const int32_t* pa, const int32_t* pj, const int64_t* pu, const double* Ux, const int64_t* ubase, double* x,
               double* y, double* part, int32_t* cnt, int32_t nb, int32_t R, double bytes, hipStream_t s);
// w[k * R + r] = D[perm[k]] * b[perm[k] + r * ldb]  /  b[perm[k] + r * ldb] = D[perm[k]] * w[k * R + r]
void gather_scale(int vs, int64_t n, int32_t R, const int64_t* perm, const double* D, const double* b, int64_t ldb, double* w,
                  hipStream_t s);
void scatter_scale(int vs, int64_t n, int32_t R, const int64_t* perm, const double* D, const double* w, double* b, int64_t ldb,
                   hipStream_t s);
}  // namespace k
}  // namespace fpa
