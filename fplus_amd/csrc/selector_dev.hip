// selector_dev.hip -- the learned precision selector on the device: per-block features, the MLP (12 -> 64 -> 64 -> 2,
// FP32) that decides per block whether its updates of later columns may run on the FP32 matrix cores, and its batched
// training (forward, softmax cross-entropy, backward, SGD with momentum) in one workgroup per step.
#include "dev_common.h"
#include "prof.h"
#include "selector_dev.h"

#include <cfloat>

namespace fpa {
namespace {

constexpr int TB = 256;

__device__ inline float lg10(double v) { return (float)fmin(40.0, fmax(-40.0, log10(fmax(v, 1e-300)))); }

template <class T> __device__ inline double mag(T v) { return Sc<T>::absval(v); }

// the three largest moduli of a block -- its inverse form (nr x w), its L21 (rows w.. of its values) and L11^-1 (w x w)
// -- over a grid of (block, column group): group g takes the columns g, g + FG, ...; the maxima meet in fmax[3 b ..]
// (bmax: non-negative doubles as bits, max-accumulated; zero on entry)
constexpr int FG = 64;   // workgroups per block (column groups): the block's values and inverse form are read once
template <class T>
__global__ __launch_bounds__(TB) void k_feat_max(const int32_t* __restrict__ blks, const BlockDesc* __restrict__ blocks,
                                                  unsigned long long* __restrict__ bmax)
{
    const int b = blks[blockIdx.x];
    const BlockDesc B = blocks[b];
    const T* V = reinterpret_cast<const T*>(B.vals);
    const T* IV = reinterpret_cast<const T*>(B.ivals);
    const T* LI = reinterpret_cast<const T*>(B.linv);
    const int32_t nr = (int32_t)B.nrows, w = (int32_t)B.w;
    double m0 = 0, m1 = 0, m2 = 0;
    for (int32_t c = blockIdx.y; c < w; c += FG) {
        const T* ivc = IV + (int64_t)c * nr;
        const T* vc = V + (int64_t)c * nr;
        // four rows per thread per pass, their loads issued before the maxima (the column is read once)
        for (int32_t r0 = threadIdx.x; r0 < nr; r0 += 4 * TB) {
            T a[4], v[4];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int32_t r = min(r0 + j * TB, nr - 1);
                a[j] = ivc[r];
                v[j] = vc[r];
            }
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int32_t r = r0 + j * TB;
                if (r < nr) {
                    m0 = fmax(m0, mag(a[j]));
                    if (r >= w) m1 = fmax(m1, mag(v[j]));
                }
            }
        }
        for (int32_t r = threadIdx.x; r < w; r += TB) m2 = fmax(m2, mag(LI[(int64_t)c * w + r]));
    }
    for (int off = 32; off > 0; off >>= 1) {
        m0 = fmax(m0, __shfl_xor(m0, off));
        m1 = fmax(m1, __shfl_xor(m1, off));
        m2 = fmax(m2, __shfl_xor(m2, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(bmax + 3 * (int64_t)b, m0);
        atomic_max_nonneg(bmax + 3 * (int64_t)b + 1, m1);
        atomic_max_nonneg(bmax + 3 * (int64_t)b + 2, m2);
    }
}

// one workgroup per new block: the block's features (fields in selector_dev.h) from the maxima of k_feat_max (read,
// from bmax, then reset to zero for the next factorization), its pivots (U11's diagonal, on the diagonal of its values) against the
// columns' largest entries (colmax, panel-local from j0), and the row and column scalings in the analysed order
template <class T>
__global__ __launch_bounds__(TB) void k_features(int32_t nb, const int32_t* __restrict__ blks, const BlockDesc* __restrict__ blocks,
                                                  const double* __restrict__ colmax, int64_t j0, const double* __restrict__ Dr,
                                                  const double* __restrict__ Dc, int64_t n, unsigned long long* __restrict__ bmax,
                                                  float* __restrict__ feat)
{
    const int b = blks[blockIdx.x];
    const BlockDesc B = blocks[b];
    const T* V = reinterpret_cast<const T*>(B.vals);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int32_t nr = (int32_t)B.nrows, w = (int32_t)B.w;
    // 0 max |U_jj|, 1 min |U_jj| (negated), 2 min |U_jj| / colmax (negated), 3 max Dc, 4 max Dr, 5 min Dc (negated),
    // 6 min Dr (negated)
    double m[7] = {0, -DBL_MAX, -DBL_MAX, 0, 0, -DBL_MAX, -DBL_MAX};
    for (int32_t c = tid; c < w; c += TB) {
        const double d = mag(V[(int64_t)c * nr + c]);
        m[0] = fmax(m[0], d);
        m[1] = fmax(m[1], -d);
        const double cm = colmax[B.c0 - j0 + c];
        m[2] = fmax(m[2], -(cm > 0 ? d / cm : 1.0));
        const double dc = Dc[B.c0 + c], dr = Dr[B.rows[c]];
        m[3] = fmax(m[3], dc);
        m[4] = fmax(m[4], dr);
        m[5] = fmax(m[5], -dc);
        m[6] = fmax(m[6], -dr);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int q = 0; q < 7; q++) m[q] = fmax(m[q], __shfl_xor(m[q], off));
    __shared__ double red[7][TB / 64];
    if (lane == 0)
        for (int q = 0; q < 7; q++) red[q][wv] = m[q];
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 7; q++)
            for (int k = 1; k < TB / 64; k++) red[q][0] = fmax(red[q][0], red[q][k]);
        double mx[3];
        for (int q = 0; q < 3; q++) {
            const unsigned long long bits = bmax[3 * (int64_t)b + q];
            mx[q] = __longlong_as_double((long long)bits);
            bmax[3 * (int64_t)b + q] = 0ULL;
        }
        float* f = feat + (int64_t)b * SEL_NF;
        f[0] = log2f((float)w);
        f[1] = log2f((float)(nr - w + 1));
        f[2] = log2f((float)nr);
        f[3] = lg10(mx[0]);
        f[4] = lg10(mx[1]);
        f[5] = lg10(red[0][0] / fmax(-red[1][0], 1e-300));
        f[6] = lg10(-red[2][0]);
        f[7] = lg10(red[3][0] / fmax(-red[5][0], 1e-300));
        f[8] = lg10(red[4][0] / fmax(-red[6][0], 1e-300));
        f[9] = lg10(red[0][0]);
        f[10] = (float)((double)B.c0 / (double)(n > 0 ? n : 1));
        f[11] = lg10(mx[2]);
    }
}

// the MLP on one sample x (normalised inside): logits o[2], and the hidden activations when asked for
struct MlpView {
    const float* W1;   // [64][12]
    const float* b1;   // [64]
    const float* W2;   // [64][64]
    const float* b2;
    const float* W3;   // [2][64]
    const float* b3;
    const float* mu;   // feature means [12]
    const float* sd;   // feature scales [12]
};
__device__ inline MlpView view(const float* m)
{
    MlpView v;
    v.W1 = m; v.b1 = v.W1 + SEL_H * SEL_NF; v.W2 = v.b1 + SEL_H; v.b2 = v.W2 + SEL_H * SEL_H; v.W3 = v.b2 + SEL_H;
    v.b3 = v.W3 + 2 * SEL_H; v.mu = v.b3 + 2; v.sd = v.mu + SEL_NF;
    return v;
}
__device__ inline void forward(const MlpView& v, const float* xin, float* x, float* h1, float* h2, float* o)
{
    for (int i = 0; i < SEL_NF; i++) x[i] = (xin[i] - v.mu[i]) / v.sd[i];
    for (int j = 0; j < SEL_H; j++) {
        float s = v.b1[j];
        for (int i = 0; i < SEL_NF; i++) s += v.W1[j * SEL_NF + i] * x[i];
        h1[j] = fmaxf(s, 0.0f);
    }
    for (int j = 0; j < SEL_H; j++) {
        float s = v.b2[j];
        for (int i = 0; i < SEL_H; i++) s += v.W2[j * SEL_H + i] * h1[i];
        h2[j] = fmaxf(s, 0.0f);
    }
    for (int k = 0; k < 2; k++) {
        float s = v.b3[k];
        for (int i = 0; i < SEL_H; i++) s += v.W3[k * SEL_H + i] * h2[i];
        o[k] = s;
    }
}

// P(FP32 is safe) for samples of the features array (one thread per sample; the weights from LDS)
__global__ __launch_bounds__(TB) void k_infer(int64_t ns, const int32_t* __restrict__ idx, const float* __restrict__ feat,
                                               const float* __restrict__ model, float thr, float* __restrict__ prob,
                                               uint8_t* __restrict__ mark)
{
    __shared__ float mdl[SEL_NW];
    for (int i = threadIdx.x; i < SEL_NW; i += TB) mdl[i] = model[i];
    __syncthreads();
    const int64_t t = blockIdx.x * (int64_t)TB + threadIdx.x;
    if (t >= ns) return;
    const int64_t s = idx ? idx[t] : t;
    const MlpView v = view(mdl);
    float x[SEL_NF], h1[SEL_H], h2[SEL_H], o[2];
    forward(v, feat + s * SEL_NF, x, h1, h2, o);
    const float p = 1.0f / (1.0f + expf(o[0] - o[1]));   // softmax probability of class 1 (FP32 safe)
    if (prob) prob[t] = p;
    if (mark) mark[s] = p > thr ? 1 : 0;
}

// one SGD-with-momentum step on the minibatch samples[first .. first + TB) (indices into the shuffled order): each
// thread one sample, the gradients summed in LDS, the weights updated in global memory by the workgroup
__global__ __launch_bounds__(TB) void k_train_step(int64_t ns, const int32_t* __restrict__ order, int64_t first,
                                                    const float* __restrict__ feat, const float* __restrict__ label,
                                                    float* __restrict__ model, float* __restrict__ mom, float lr, float beta,
                                                    float wdecay, float* __restrict__ loss_acc)
{
    __shared__ float mdl[SEL_NW];
    __shared__ float grad[SEL_NP];
    __shared__ float lsum[TB];
    for (int i = threadIdx.x; i < SEL_NW; i += TB) mdl[i] = model[i];
    for (int i = threadIdx.x; i < SEL_NP; i += TB) grad[i] = 0.0f;
    __syncthreads();
    const int64_t t = first + threadIdx.x;
    const MlpView v = view(mdl);
    float loss = 0.0f;
    int cnt = 0;
    if (t < ns) {
        const int64_t s = order[t];
        float x[SEL_NF], h1[SEL_H], h2[SEL_H], o[2];
        forward(v, feat + s * SEL_NF, x, h1, h2, o);
        const float mx = fmaxf(o[0], o[1]);
        const float e0 = expf(o[0] - mx), e1 = expf(o[1] - mx), z = e0 + e1;
        const float y = label[s];   // 1: FP32 safe
        const float p1 = e1 / z;
        loss = -(y * logf(fmaxf(p1, 1e-12f)) + (1.0f - y) * logf(fmaxf(1.0f - p1, 1e-12f)));
        cnt = 1;
        const float d0 = e0 / z - (1.0f - y), d1 = p1 - y;   // dL/do
        float* gW1 = grad;
        float* gb1 = gW1 + SEL_H * SEL_NF;
        float* gW2 = gb1 + SEL_H;
        float* gb2 = gW2 + SEL_H * SEL_H;
        float* gW3 = gb2 + SEL_H;
        float* gb3 = gW3 + 2 * SEL_H;
        atomicAdd(&gb3[0], d0);
        atomicAdd(&gb3[1], d1);
        float dh2[SEL_H];
        for (int i = 0; i < SEL_H; i++) {
            atomicAdd(&gW3[i], d0 * h2[i]);
            atomicAdd(&gW3[SEL_H + i], d1 * h2[i]);
            dh2[i] = h2[i] > 0.0f ? d0 * v.W3[i] + d1 * v.W3[SEL_H + i] : 0.0f;
        }
        float dh1[SEL_H];
        for (int i = 0; i < SEL_H; i++) dh1[i] = 0.0f;
        for (int j = 0; j < SEL_H; j++) {
            if (dh2[j] == 0.0f) continue;
            atomicAdd(&gb2[j], dh2[j]);
            for (int i = 0; i < SEL_H; i++) {
                if (h1[i] != 0.0f) atomicAdd(&gW2[j * SEL_H + i], dh2[j] * h1[i]);
                dh1[i] += dh2[j] * v.W2[j * SEL_H + i];
            }
        }
        for (int j = 0; j < SEL_H; j++) {
            if (h1[j] <= 0.0f) continue;
            atomicAdd(&gb1[j], dh1[j]);
            for (int i = 0; i < SEL_NF; i++) atomicAdd(&gW1[j * SEL_NF + i], dh1[j] * x[i]);
        }
    }
    lsum[threadIdx.x] = loss;
    __syncthreads();
    const int64_t bs = ns - first < TB ? ns - first : TB;
    const float inv = 1.0f / (float)(bs > 0 ? bs : 1);
    for (int i = threadIdx.x; i < SEL_NP; i += TB) {
        const float g = grad[i] * inv + wdecay * mdl[i];
        const float mm = beta * mom[i] + g;
        mom[i] = mm;
        model[i] = mdl[i] - lr * mm;
    }
    if (threadIdx.x == 0) {
        float s = 0.0f;
        for (int i = 0; i < TB; i++) s += lsum[i];
        atomicAdd(loss_acc, s);
    }
    (void)cnt;
}

}  // namespace

namespace k {
void block_features(int vs, int32_t nb, const int32_t* blks, const BlockDesc* blocks, const double* colmax, int64_t j0,
                    const double* Dr, const double* Dc, int64_t n, unsigned long long* fmax, float* feat, hipStream_t s)
{
    if (nb <= 0) return;
    ProfScope ps("sel_features", 0.0, s);
    if (vs == 1) {
        hipLaunchKernelGGL(k_feat_max<double>, dim3(nb, FG), dim3(TB), 0, s, blks, blocks, fmax);
        hipLaunchKernelGGL(k_features<double>, dim3(nb), dim3(TB), 0, s, nb, blks, blocks, colmax, j0, Dr, Dc, n, fmax, feat);
    } else {
        hipLaunchKernelGGL(k_feat_max<Zd>, dim3(nb, FG), dim3(TB), 0, s, blks, blocks, fmax);
        hipLaunchKernelGGL(k_features<Zd>, dim3(nb), dim3(TB), 0, s, nb, blks, blocks, colmax, j0, Dr, Dc, n, fmax, feat);
    }
}
void sel_infer(int64_t ns, const int32_t* idx, const float* feat, const float* model, float thr, float* prob, uint8_t* mark,
               hipStream_t s)
{
    if (ns <= 0) return;
    ProfScope ps("sel_infer", 0.0, s);
    hipLaunchKernelGGL(k_infer, dim3((unsigned)((ns + TB - 1) / TB)), dim3(TB), 0, s, ns, idx, feat, model, thr, prob, mark);
}
void sel_train_step(int64_t ns, const int32_t* order, int64_t first, const float* feat, const float* label, float* model,
                    float* mom, float lr, float beta, float wdecay, float* loss_acc, hipStream_t s)
{
    hipLaunchKernelGGL(k_train_step, dim3(1), dim3(TB), 0, s, ns, order, first, feat, label, model, mom, lr, beta, wdecay, loss_acc);
}
}  // namespace k
}  // namespace fpa
