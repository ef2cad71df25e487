// refine_dev.hip -- the device side of the refined solves (GMRES-based iterative refinement in FP64): the residual with
// the equilibrated matrix, blocked dot products over several right-hand sides, and the Gram-Schmidt and solution
// updates.  Vectors are n x R row-major (v[k * R + r]), R <= 16 right-hand sides, as the triangular sweeps take them.
#include "dev_common.h"
#include "prof.h"
#include "refine_dev.h"

namespace fpa {
namespace {

constexpr int TB = 256;

inline unsigned gridn(int64_t n, int per)
{
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

// out[k, r] = bb[k, r] - sum_q Abar(i, j_q) u[j_q, r] for row i = plu[k] of the equilibrated matrix (CSR: rp, cj, the
// entries' positions pos in the CSC value array Abar); bb null: out = the product itself
template <class T>
__global__ void k_spmv(int64_t n, int32_t R, const int64_t* __restrict__ plu, const int64_t* __restrict__ rp,
                       const int32_t* __restrict__ cj, const int64_t* __restrict__ pos, const double* __restrict__ Abar,
                       const double* __restrict__ u, const double* __restrict__ bb, double* __restrict__ out)
{
    const T* A = reinterpret_cast<const T*>(Abar);
    const T* U = reinterpret_cast<const T*>(u);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * R; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = t / R;
        const int r = (int)(t - k * R);
        const int64_t i = plu[k];
        T acc = Sc<T>::zero();
        for (int64_t q = rp[i]; q < rp[i + 1]; q++) acc = Sc<T>::fma_(acc, A[pos[q]], U[(int64_t)cj[q] * R + r]);
        if (bb) acc = Sc<T>::sub(reinterpret_cast<const T*>(bb)[t], acc);
        reinterpret_cast<T*>(out)[t] = acc;
    }
}

// the largest column sum of |Abar| (CSC), as the bits of a non-negative double
template <class T>
__global__ void k_norm1(int64_t n, const int64_t* __restrict__ Bp, const double* __restrict__ Abar, unsigned long long* out)
{
    const T* A = reinterpret_cast<const T*>(Abar);
    double m = 0.0;
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
        double sj = 0.0;
        for (int64_t q = Bp[j]; q < Bp[j + 1]; q++) sj += Sc<T>::absval(A[q]);
        m = fmax(m, sj);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomic_max_nonneg(out, m);
}

template <class T> __device__ inline T conj_mul(T a, T b) { return Sc<T>::mul(Sc<T>::conj(a), b); }

// partial sums of the dot products d[j, r] = sum_k conj(V_j[k, r]) w[k, r], j < J: block b sums the rows
// b, b + G, b + 2 G, ... (G the grid) into part[(b * J + j) * R + r]; the threads of a block split the rows TB / R ways
// and meet in LDS in a fixed order
template <class T>
__global__ __launch_bounds__(TB) void k_dots_part(int64_t n, int32_t R, int32_t J, const double* __restrict__ V, int64_t vstride,
                                                   const double* __restrict__ w, double* __restrict__ part)
{
    __shared__ T red[TB];
    const int tid = threadIdx.x, r = tid % R, sub = tid / R, nsub = TB / R;
    const int64_t G = gridDim.x;
    const T* W = reinterpret_cast<const T*>(w);
    for (int j = 0; j < J; j++) {
        const T* Vj = reinterpret_cast<const T*>(V + (int64_t)j * vstride);   // (vstride counts doubles)
        T acc = Sc<T>::zero();
        if (sub < nsub)
            for (int64_t k = blockIdx.x + (int64_t)sub * G; k < n; k += (int64_t)nsub * G)
                acc = Sc<T>::add(acc, conj_mul(Vj[k * R + r], W[k * R + r]));
        red[tid] = acc;
        __syncthreads();
        if (tid < R) {
            T s = Sc<T>::zero();
            for (int q = 0; q < nsub; q++) s = Sc<T>::add(s, red[q * R + tid]);
            reinterpret_cast<T*>(part)[((int64_t)blockIdx.x * J + j) * R + tid] = s;
        }
        __syncthreads();
    }
}

// d[j, r] = the sum over the G blocks' partial sums, in block order
template <class T>
__global__ void k_dots_sum(int32_t G, int32_t R, int32_t J, const double* __restrict__ part, double* __restrict__ d)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= J * R) return;
    const T* P = reinterpret_cast<const T*>(part);
    T s = Sc<T>::zero();
    for (int b = 0; b < G; b++) s = Sc<T>::add(s, P[(int64_t)b * J * R + t]);
    reinterpret_cast<T*>(d)[t] = s;
}

// w[k, r] += sgn * sum_{j < J} c[j, r] V_j[k, r]
template <class T>
__global__ void k_combine(int64_t n, int32_t R, int32_t J, const double* __restrict__ V, int64_t vstride, const double* __restrict__ c,
                          double sgn, double* __restrict__ w)
{
    const T* C = reinterpret_cast<const T*>(c);
    T* Wv = reinterpret_cast<T*>(w);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * R; t += (int64_t)gridDim.x * blockDim.x) {
        const int r = (int)(t % R);
        T acc = Sc<T>::zero();
        for (int j = 0; j < J; j++) acc = Sc<T>::fma_(acc, C[j * R + r], reinterpret_cast<const T*>(V + (int64_t)j * vstride)[t]);
        Wv[t] = Sc<T>::add(Wv[t], Sc<T>::scale(acc, sgn));
    }
}

// out[k, r] = w[k, r] * s[r] (real scales)
template <class T>
__global__ void k_scale_cols(int64_t n, int32_t R, const double* __restrict__ w, const double* __restrict__ sc, double* __restrict__ out)
{
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n * R; t += (int64_t)gridDim.x * blockDim.x)
        reinterpret_cast<T*>(out)[t] = Sc<T>::scale(reinterpret_cast<const T*>(w)[t], sc[t % R]);
}

}  // namespace

namespace k {
void spmv_eq(int vs, int64_t n, int32_t R, const int64_t* plu, const int64_t* rp, const int32_t* cj, const int64_t* pos,
             const double* Abar, const double* u, const double* bb, double* out, double bytes, hipStream_t s)
{
    if (n <= 0) return;
    ProfScope ps("ir_spmv", bytes, s);
    if (vs == 1) hipLaunchKernelGGL(k_spmv<double>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, plu, rp, cj, pos, Abar, u, bb, out);
    else hipLaunchKernelGGL(k_spmv<Zd>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, plu, rp, cj, pos, Abar, u, bb, out);
}
void norm1(int vs, int64_t n, const int64_t* Bp, const double* Abar, unsigned long long* out, hipStream_t s)
{
    if (n <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_norm1<double>, dim3(gridn(n, TB)), dim3(TB), 0, s, n, Bp, Abar, out);
    else hipLaunchKernelGGL(k_norm1<Zd>, dim3(gridn(n, TB)), dim3(TB), 0, s, n, Bp, Abar, out);
}
int dots_blocks() { return 512; }
void dots(int vs, int64_t n, int32_t R, int32_t J, const double* V, int64_t vstride, const double* w, double* part, double* d,
          hipStream_t s)
{
    if (J <= 0) return;
    if (R < 1 || R > 16) throw std::invalid_argument("dots: 1 to 16 right-hand sides");
    const int G = dots_blocks();
    ProfScope ps("ir_dots", (double)(J + 1) * n * R * 8.0 * vs, s);
    if (vs == 1) {
        hipLaunchKernelGGL(k_dots_part<double>, dim3(G), dim3(TB), 0, s, n, R, J, V, vstride, w, part);
        hipLaunchKernelGGL(k_dots_sum<double>, dim3(gridn(J * R, TB)), dim3(TB), 0, s, G, R, J, part, d);
    } else {
        hipLaunchKernelGGL(k_dots_part<Zd>, dim3(G), dim3(TB), 0, s, n, R, J, V, vstride, w, part);
        hipLaunchKernelGGL(k_dots_sum<Zd>, dim3(gridn(J * R, TB)), dim3(TB), 0, s, G, R, J, part, d);
    }
}
void combine(int vs, int64_t n, int32_t R, int32_t J, const double* V, int64_t vstride, const double* c, double sgn, double* w,
             hipStream_t s)
{
    if (J <= 0 || n <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_combine<double>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, J, V, vstride, c, sgn, w);
    else hipLaunchKernelGGL(k_combine<Zd>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, J, V, vstride, c, sgn, w);
}
void scale_cols(int vs, int64_t n, int32_t R, const double* w, const double* sc, double* out, hipStream_t s)
{
    if (n <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_scale_cols<double>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, w, sc, out);
    else hipLaunchKernelGGL(k_scale_cols<Zd>, dim3(gridn(n * R, TB)), dim3(TB), 0, s, n, R, w, sc, out);
}
}  // namespace k
}  // namespace fpa
