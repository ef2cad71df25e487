// solve_dev.h -- device triangular solves on the supernode blocks and U (kernel launchers).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor_dev.h"

namespace fpa {
namespace k {
// x: n x nrhs column-major (ld = n), indexed by pivot order.  blk: block ids of one level.  pinv_rows maps a
// row of the factored matrix to its pivot position.
void lfwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int32_t* pinv_rows, double* x,
          int64_t ldx, int32_t nrhs, hipStream_t s);
void ubwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
          const double* Ux, double* x, int64_t ldx, int32_t nrhs, hipStream_t s);
void utfwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
           const double* Ux, double* x, int64_t ldx, int32_t nrhs, int conj, hipStream_t s);
void ltbwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int32_t* pinv_rows, double* x,
           int64_t ldx, int32_t nrhs, int conj, hipStream_t s);
// w[k, r] = D[perm[k]] * b[perm[k], r]  /  x[perm[k], r] = D[perm[k]] * w[k, r]
void gather_scale(int vs, int64_t n, int32_t nrhs, const int64_t* perm, const double* D, const double* b, int64_t ldb,
                  double* w, hipStream_t s);
void scatter_scale(int vs, int64_t n, int32_t nrhs, const int64_t* perm, const double* D, const double* w, double* b,
                   int64_t ldb, hipStream_t s);
}  // namespace k
}  // namespace fpa
