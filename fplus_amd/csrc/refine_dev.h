// refine_dev.h -- kernels of the refined solves (refine_dev.hip).  Vectors are n x R row-major, R <= 16.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpa {
namespace k {
// out[k, r] = bb[k, r] - (Abar u)[plu[k], r] (bb null: the product alone), Abar by rows through rp / cj / pos (the
// entries' positions in the CSC value array); bytes: algorithmic bytes (timing)
void spmv_eq(int vs, int64_t n, int32_t R, const int64_t* plu, const int64_t* rp, const int32_t* cj, const int64_t* pos,
             const double* Abar, const double* u, const double* bb, double* out, double bytes, hipStream_t s);
// the largest absolute column sum of the CSC values Abar (bits of a non-negative double, max-accumulated into *out)
void norm1(int vs, int64_t n, const int64_t* Bp, const double* Abar, unsigned long long* out, hipStream_t s);
// d[j, r] = sum_k conj(V_j[k, r]) w[k, r] for j < J (V_j at V + j vstride doubles); part: dots_blocks() * J * R scalars
int dots_blocks();
void dots(int vs, int64_t n, int32_t R, int32_t J, const double* V, int64_t vstride, const double* w, double* part, double* d,
          hipStream_t s);
// w[k, r] += sgn sum_{j < J} c[j, r] V_j[k, r]
void combine(int vs, int64_t n, int32_t R, int32_t J, const double* V, int64_t vstride, const double* c, double sgn, double* w,
             hipStream_t s);
// out[k, r] = w[k, r] sc[r]
void scale_cols(int vs, int64_t n, int32_t R, const double* w, const double* sc, double* out, hipStream_t s);
}  // namespace k
}  // namespace fpa
