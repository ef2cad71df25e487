// selector_dev.h -- the learned precision selector's device side (selector_dev.hip).
//
// Features of a block (SEL_NF floats): 0 log2 w, 1 log2(rows below + 1), 2 log2 rows, 3 log10 max |inverse form|,
// 4 log10 max |L21|, 5 log10(max |U_jj| / min |U_jj|), 6 log10 min_j |U_jj| / colmax_j, 7 log10(max Dc / min Dc) over its
// columns, 8 log10(max Dr / min Dr) over its pivot rows, 9 log10 max |U_jj|, 10 first column / n, 11 log10 max |L11^-1|.
// The model: W1 [64][12], b1 [64], W2 [64][64], b2 [64], W3 [2][64], b3 [2] (SEL_NP trainable floats), then the features'
// means and scales [12] each (SEL_NW floats in all).  Class 1: the block's updates may run in FP32.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "factor_dev.h"

#define SEL_NF 12
#define SEL_H 64
#define SEL_NP (SEL_H * SEL_NF + SEL_H + SEL_H * SEL_H + SEL_H + 2 * SEL_H + 2)
#define SEL_NW (SEL_NP + 2 * SEL_NF)

namespace fpa {
namespace k {
// features of the blocks blks[0 .. nb) into feat[block * SEL_NF ..]; colmax: the panel's column maxima (panel-local from
// column j0); Dr, Dc: the scalings of the analysed matrix; fmax: 3 per block, zero on entry and left zero
void block_features(int vs, int32_t nb, const int32_t* blks, const BlockDesc* blocks, const double* colmax, int64_t j0,
                    const double* Dr, const double* Dc, int64_t n, unsigned long long* fmax, float* feat, hipStream_t s);
// P(class 1) of samples idx[t] (or t when idx is null) of feat: into prob[t] and/or mark[idx[t]] = P > thr
void sel_infer(int64_t ns, const int32_t* idx, const float* feat, const float* model, float thr, float* prob, uint8_t* mark,
               hipStream_t s);
// one training step on the minibatch order[first .. first + 256) of the ns samples (labels 0 / 1)
void sel_train_step(int64_t ns, const int32_t* order, int64_t first, const float* feat, const float* label, float* model,
                    float* mom, float lr, float beta, float wdecay, float* loss_acc, hipStream_t s);
}  // namespace k
}  // namespace fpa
