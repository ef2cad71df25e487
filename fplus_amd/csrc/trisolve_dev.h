// trisolve_dev.h -- device sweeps on CSC factors given by the caller (fplus_tri_solve, fplus_tri_solve_plan): launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpa {

constexpr int TS_GROUP = 64;   // indices per group (one workgroup walks a group's chain of dependent indices)
constexpr int TS_RHS = 16;     // right-hand sides per pass

namespace k {
// One level of a sweep.  Index k's value is  x_k = (x_k - sum_q tval[q] * x[tidx[q]]) [/ diag[k]]  over its terms
// q = tp[k] .. tp[k+1]-1 in that order; a term whose x value is zero is skipped when skipzero is set.  A group g covers
// the indices grp[g].x .. grp[g].x + grp[g].y - 1, walked ascending (dir > 0) or descending; its indices' terms
// either point outside the group (indices of earlier levels) or at indices of the group earlier in the walk.  The first
// nfirst[k] terms of k point outside the group.  x: n x R row-major, R <= TS_RHS.  diag: null for a unit triangle.
void tsweep(int vs, int32_t ngrp, const int32_t* glist, const int2* grp, const int64_t* tp, const int32_t* nfirst,
            const int32_t* tidx, const double* tval, const double* diag, double* x, int32_t R, int dir, int skipzero,
            double bytes, hipStream_t s);
// rows[m] (m < nrows) with terms beyond the diagonal: out[m * R + r] = x_rows[m] - sum_q tval[q] * x[tidx[q]] over
// q = tp[m] .. tp[m+1]-1 (zero x skipped), then put() writes out back to x
void lupper(int vs, int64_t nrows, const int32_t* rows, const int64_t* tp, const int32_t* tidx, const double* tval,
            const double* x, double* out, int32_t R, hipStream_t s);
void lupper_put(int vs, int64_t nrows, const int32_t* rows, const double* out, double* x, int32_t R, hipStream_t s);
// dst[o] = src[map[o]], conjugated when conj (complex entries)
void gather_conj(int vs, int64_t cnt, const int64_t* map, const double* src, double* dst, int conj, hipStream_t s);
// w[k * R + r] = b[k + r * ldb]  /  b[k + r * ldb] = w[k * R + r]
void rhs_in(int vs, int64_t n, int32_t R, const double* b, int64_t ldb, double* w, hipStream_t s);
void rhs_out(int vs, int64_t n, int32_t R, const double* w, double* b, int64_t ldb, hipStream_t s);
}  // namespace k
}  // namespace fpa
