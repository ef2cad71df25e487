// solve_dev.hip -- the triangular solves on the device factors (gfx950).
//
// L is held as supernode blocks (factor_dev.h), U in CSC with its rows in pivot order and the diagonal last in every
// column.  A solve walks the blocks level by level (a level's blocks are independent); one workgroup per block:
//   lfwd   L y = b   the block's unit-lower triangle, then the rows below subtracted (push, atomic);
//   ubwd   U x = y   columns of the block last to first, x_j = y_j / U_jj, U(:, j) subtracted (push);
//   utfwd  U^T y = b (U^H with conj) dot products down U's columns (pull), the block's own rows in order;
//   ltbwd  L^T x = y (L^H) dot products down the block's columns (pull), the block's triangle last to first.
// Blocks are at most 256 columns wide (the panel width), so the triangular part keeps one row per thread.
#include "dev_common.h"
#include "device_api.h"
#include "solve_dev.h"

namespace fpa {
namespace {

constexpr int TB = 256;

template <class T> __device__ inline T cj(T v, int c) { return c ? Sc<T>::conj(v) : v; }

template <class T>
__global__ __launch_bounds__(TB) void k_lfwd(const int32_t* __restrict__ blk, const BlockDesc* __restrict__ blocks,
                                              const int32_t* __restrict__ pinv_rows, double* __restrict__ xd, int64_t ldx,
                                              int32_t nrhs)
{
    __shared__ T ys[TB];
    const BlockDesc B = blocks[blk[blockIdx.x]];
    const T* V = reinterpret_cast<const T*>(B.vals);
    T* x = reinterpret_cast<T*>(xd);
    const int t = threadIdx.x, w = B.w;
    for (int32_t r = 0; r < nrhs; r++) {
        T* xr = x + (int64_t)r * ldx;
        T my = t < w ? xr[B.c0 + t] : Sc<T>::zero();
        for (int c = 0; c < w; c++) {
            if (t == c) ys[c] = my;
            __syncthreads();
            if (t > c && t < w) my = Sc<T>::fms(my, V[(int64_t)c * B.nrows + t], ys[c]);
        }
        if (t < w) xr[B.c0 + t] = my;
        __syncthreads();
        for (int32_t i = w + t; i < B.nrows; i += TB) {
            T acc = Sc<T>::zero();
            for (int c = 0; c < w; c++) acc = Sc<T>::fma_(acc, V[(int64_t)c * B.nrows + i], ys[c]);
            atomic_sub(xr + pinv_rows[B.rows[i]], acc);
        }
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(TB) void k_ubwd(const int32_t* __restrict__ blk, const BlockDesc* __restrict__ blocks,
                                              const int64_t* __restrict__ Up, const int32_t* __restrict__ Ui,
                                              const double* __restrict__ Uxd, double* __restrict__ xd, int64_t ldx, int32_t nrhs)
{
    __shared__ T ys[TB];
    const BlockDesc B = blocks[blk[blockIdx.x]];
    const T* Ux = reinterpret_cast<const T*>(Uxd);
    T* x = reinterpret_cast<T*>(xd);
    const int t = threadIdx.x, w = B.w;
    const int64_t c0 = B.c0;
    for (int32_t r = 0; r < nrhs; r++) {
        T* xr = x + (int64_t)r * ldx;
        if (t < w) ys[t] = xr[c0 + t];
        __syncthreads();
        for (int c = w - 1; c >= 0; c--) {
            const int64_t q0 = Up[c0 + c], qd = Up[c0 + c + 1] - 1;
            if (t == 0) ys[c] = Sc<T>::div(ys[c], Ux[qd]);
            __syncthreads();
            const T xc = ys[c];
            for (int64_t q = q0 + t; q < qd; q += TB) {
                const int64_t i = Ui[q];
                if (i >= c0) ys[i - c0] = Sc<T>::fms(ys[i - c0], Ux[q], xc);
                else atomic_sub(xr + i, Sc<T>::mul(Ux[q], xc));
            }
            __syncthreads();
        }
        if (t < w) xr[c0 + t] = ys[t];
        __syncthreads();
    }
}

template <class T>
__device__ inline T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) {
        if constexpr (Sc<T>::vs == 1) v += __shfl_xor(v, off);
        else { v.x += __shfl_xor(v.x, off); v.y += __shfl_xor(v.y, off); }
    }
    return v;
}

template <class T>
__global__ __launch_bounds__(TB) void k_utfwd(const int32_t* __restrict__ blk, const BlockDesc* __restrict__ blocks,
                                               const int64_t* __restrict__ Up, const int32_t* __restrict__ Ui,
                                               const double* __restrict__ Uxd, double* __restrict__ xd, int64_t ldx,
                                               int32_t nrhs, int conj)
{
    __shared__ T ys[TB];
    const BlockDesc B = blocks[blk[blockIdx.x]];
    const T* Ux = reinterpret_cast<const T*>(Uxd);
    T* x = reinterpret_cast<T*>(xd);
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6, w = B.w;
    const int64_t c0 = B.c0;
    for (int32_t r = 0; r < nrhs; r++) {
        T* xr = x + (int64_t)r * ldx;
        // the parts of the dot products over rows before the block: one wavefront per column
        for (int c = wid; c < w; c += TB / 64) {
            const int64_t q0 = Up[c0 + c], qd = Up[c0 + c + 1] - 1;
            T acc = Sc<T>::zero();
            for (int64_t q = q0 + lane; q < qd; q += 64) {
                const int64_t i = Ui[q];
                if (i < c0) acc = Sc<T>::fma_(acc, cj(Ux[q], conj), xr[i]);
            }
            acc = wave_sum(acc);
            if (lane == 0) ys[c] = Sc<T>::sub(xr[c0 + c], acc);
        }
        __syncthreads();
        // the block's own rows, in column order (wavefront 0)
        if (wid == 0) {
            for (int c = 0; c < w; c++) {
                const int64_t q0 = Up[c0 + c], qd = Up[c0 + c + 1] - 1;
                T acc = Sc<T>::zero();
                for (int64_t q = q0 + lane; q < qd; q += 64) {
                    const int64_t i = Ui[q];
                    if (i >= c0) acc = Sc<T>::fma_(acc, cj(Ux[q], conj), ys[i - c0]);
                }
                acc = wave_sum(acc);
                if (lane == 0) ys[c] = Sc<T>::div(Sc<T>::sub(ys[c], acc), cj(Ux[qd], conj));
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            }
        }
        __syncthreads();
        if (t < w) xr[c0 + t] = ys[t];
        __syncthreads();
    }
}

template <class T>
__global__ __launch_bounds__(TB) void k_ltbwd(const int32_t* __restrict__ blk, const BlockDesc* __restrict__ blocks,
                                               const int32_t* __restrict__ pinv_rows, double* __restrict__ xd, int64_t ldx,
                                               int32_t nrhs, int conj)
{
    __shared__ T ys[TB];
    const BlockDesc B = blocks[blk[blockIdx.x]];
    const T* V = reinterpret_cast<const T*>(B.vals);
    T* x = reinterpret_cast<T*>(xd);
    const int t = threadIdx.x, lane = t & 63, wid = t >> 6, w = B.w;
    for (int32_t r = 0; r < nrhs; r++) {
        T* xr = x + (int64_t)r * ldx;
        for (int c = wid; c < w; c += TB / 64) {
            T acc = Sc<T>::zero();
            for (int32_t i = w + lane; i < B.nrows; i += 64)
                acc = Sc<T>::fma_(acc, cj(V[(int64_t)c * B.nrows + i], conj), xr[pinv_rows[B.rows[i]]]);
            acc = wave_sum(acc);
            if (lane == 0) ys[c] = Sc<T>::sub(xr[B.c0 + c], acc);
        }
        __syncthreads();
        // the triangle, last column first: x_c = ys_c - sum_{rr > c} L(rr, c) x_rr (thread rr holds x_rr)
        T my = t < w ? ys[t] : Sc<T>::zero();
        for (int c = w - 1; c >= 0; c--) {
            // every thread rr > c contributes L(rr, c) x_rr; a block reduction into ys[c]
            T part = (t > c && t < w) ? Sc<T>::mul(cj(V[(int64_t)c * B.nrows + t], conj), my) : Sc<T>::zero();
            part = wave_sum(part);
            __shared__ T red[TB / 64];
            if (lane == 0) red[wid] = part;
            __syncthreads();
            if (t == c) {
                T s = Sc<T>::zero();
                for (int k2 = 0; k2 < TB / 64; k2++) s = Sc<T>::add(s, red[k2]);
                my = Sc<T>::sub(my, s);
            }
            __syncthreads();
        }
        if (t < w) xr[B.c0 + t] = my;
        __syncthreads();
    }
}

template <class T>
__global__ void k_gather_scale(int64_t n, int32_t nrhs, const int64_t* __restrict__ perm, const double* __restrict__ D,
                               const double* __restrict__ bd, int64_t ldb, double* __restrict__ wd)
{
#pragma clang fp contract(off)
    const T* b = reinterpret_cast<const T*>(bd);
    T* w = reinterpret_cast<T*>(wd);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t src = perm[k];
        const double d = D[src];
        for (int32_t r = 0; r < nrhs; r++) w[k + (int64_t)r * n] = Sc<T>::scale(b[src + (int64_t)r * ldb], d);
    }
}

template <class T>
__global__ void k_scatter_scale(int64_t n, int32_t nrhs, const int64_t* __restrict__ perm, const double* __restrict__ D,
                                const double* __restrict__ wd, double* __restrict__ bd, int64_t ldb)
{
#pragma clang fp contract(off)
    const T* w = reinterpret_cast<const T*>(wd);
    T* b = reinterpret_cast<T*>(bd);
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t dst = perm[k];
        const double d = D[dst];
        for (int32_t r = 0; r < nrhs; r++) b[dst + (int64_t)r * ldb] = Sc<T>::scale(w[k + (int64_t)r * n], d);
    }
}

inline unsigned gridn(int64_t n)
{
    int64_t g = (n + TB - 1) / TB;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

namespace k {
void lfwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int32_t* pinv_rows, double* x,
          int64_t ldx, int32_t nrhs, hipStream_t s)
{
    if (nblk <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_lfwd<double>, dim3(nblk), dim3(TB), 0, s, blk, blocks, pinv_rows, x, ldx, nrhs);
    else hipLaunchKernelGGL(k_lfwd<Zd>, dim3(nblk), dim3(TB), 0, s, blk, blocks, pinv_rows, x, ldx, nrhs);
}
void ubwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
          const double* Ux, double* x, int64_t ldx, int32_t nrhs, hipStream_t s)
{
    if (nblk <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_ubwd<double>, dim3(nblk), dim3(TB), 0, s, blk, blocks, Up, Ui, Ux, x, ldx, nrhs);
    else hipLaunchKernelGGL(k_ubwd<Zd>, dim3(nblk), dim3(TB), 0, s, blk, blocks, Up, Ui, Ux, x, ldx, nrhs);
}
void utfwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
           const double* Ux, double* x, int64_t ldx, int32_t nrhs, int conj, hipStream_t s)
{
    if (nblk <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_utfwd<double>, dim3(nblk), dim3(TB), 0, s, blk, blocks, Up, Ui, Ux, x, ldx, nrhs, conj);
    else hipLaunchKernelGGL(k_utfwd<Zd>, dim3(nblk), dim3(TB), 0, s, blk, blocks, Up, Ui, Ux, x, ldx, nrhs, conj);
}
void ltbwd(int vs, int32_t nblk, const int32_t* blk, const BlockDesc* blocks, const int32_t* pinv_rows, double* x,
           int64_t ldx, int32_t nrhs, int conj, hipStream_t s)
{
    if (nblk <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_ltbwd<double>, dim3(nblk), dim3(TB), 0, s, blk, blocks, pinv_rows, x, ldx, nrhs, conj);
    else hipLaunchKernelGGL(k_ltbwd<Zd>, dim3(nblk), dim3(TB), 0, s, blk, blocks, pinv_rows, x, ldx, nrhs, conj);
}
void gather_scale(int vs, int64_t n, int32_t nrhs, const int64_t* perm, const double* D, const double* b, int64_t ldb,
                  double* w, hipStream_t s)
{
    if (n <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_gather_scale<double>, dim3(gridn(n)), dim3(TB), 0, s, n, nrhs, perm, D, b, ldb, w);
    else hipLaunchKernelGGL(k_gather_scale<Zd>, dim3(gridn(n)), dim3(TB), 0, s, n, nrhs, perm, D, b, ldb, w);
}
void scatter_scale(int vs, int64_t n, int32_t nrhs, const int64_t* perm, const double* D, const double* w, double* b,
                   int64_t ldb, hipStream_t s)
{
    if (n <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_scatter_scale<double>, dim3(gridn(n)), dim3(TB), 0, s, n, nrhs, perm, D, w, b, ldb);
    else hipLaunchKernelGGL(k_scatter_scale<Zd>, dim3(gridn(n)), dim3(TB), 0, s, n, nrhs, perm, D, w, b, ldb);
}
}  // namespace k
}  // namespace fpa
