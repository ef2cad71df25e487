// trisolve_dev.hip -- triangular sweeps on CSC factors given by the caller (gfx950).
//
// The plain column sweeps of reference src/solve.c:11-60 push every column's update into the rows below it.  Here every
// index pulls: its terms (the entries that update it) are listed in the order in which the column sweep would apply
// them, and one thread per (index, right-hand side) applies them in that order.  Each value therefore sees the same
// sequence of roundings as in the column sweep, and this file is compiled without floating-point contraction
// (-ffp-contract=off), so the results equal the column sweep's bit for bit (complex division: Smith's formula, the
// same operations as the C99 division of gcc's runtime for finite values).
//
// Indices are grouped into runs of at most TS_GROUP consecutive indices; the groups of one level are independent and
// run as one launch, a workgroup per group.  Inside a group the workgroup steps through the indices in sweep order:
// at step t index t is final; after a barrier the later indices apply their terms that have become ready.
#include "dev_common.h"
#include "device_api.h"
#include "prof.h"
#include "trisolve_dev.h"

namespace fpa {
namespace {

constexpr int TB = 256;
constexpr int PP = TS_GROUP * TS_RHS / TB;   // (index, right-hand side) pairs per thread

template <class T>
__global__ __launch_bounds__(TB) void k_tsweep(const int32_t* __restrict__ glist, const int2* __restrict__ grp,
                                                const int64_t* __restrict__ tp, const int32_t* __restrict__ nfirst,
                                                const int32_t* __restrict__ tidx, const double* __restrict__ tvald,
                                                const double* __restrict__ diagd, double* __restrict__ xd, int32_t R,
                                                int dir, int skipzero)
{
    __shared__ T xs[TS_GROUP * TS_RHS];
    const int2 G = grp[glist[blockIdx.x]];
    const int32_t c0 = G.x, w = G.y;
    const T* tv = reinterpret_cast<const T*>(tvald);
    const T* dg = reinterpret_cast<const T*>(diagd);
    T* x = reinterpret_cast<T*>(xd);
    // step position of the index k of this group
    auto posof = [&](int32_t k) { return dir > 0 ? k - c0 : c0 + w - 1 - k; };
    auto indexof = [&](int t) { return dir > 0 ? c0 + t : c0 + w - 1 - t; };
    T acc[PP];
    int64_t q[PP], qe[PP];
    int tpos[PP];
#pragma unroll
    for (int s = 0; s < PP; s++) {
        const int p = threadIdx.x + s * TB, t = p / R, r = p - t * R;
        tpos[s] = t < w ? t : -1;
        acc[s] = Sc<T>::zero();
        q[s] = qe[s] = 0;
        if (t >= w) continue;
        const int32_t k = indexof(t);
        acc[s] = x[(int64_t)k * R + r];
        q[s] = tp[k];
        qe[s] = tp[k + 1];
        // the terms that point outside the group: their values are final
        const int64_t qf = q[s] + nfirst[k];
        for (; q[s] < qf; q[s]++) {
            const T xv = x[(int64_t)tidx[q[s]] * R + r];
            if (skipzero && Sc<T>::iszero(xv)) continue;
            acc[s] = Sc<T>::fms(acc[s], tv[q[s]], xv);
        }
    }
    for (int t = 0; t < w; t++) {
        // index t: every remaining term is ready (outside the group, or earlier in the group), then it is final
#pragma unroll
        for (int s = 0; s < PP; s++) {
            if (tpos[s] != t) continue;
            const int r = threadIdx.x + s * TB - t * R;
            for (; q[s] < qe[s]; q[s]++) {
                const int32_t j = tidx[q[s]];
                const uint32_t jl = (uint32_t)(j - c0);
                const T xv = jl < (uint32_t)w ? xs[posof(j) * R + r] : x[(int64_t)j * R + r];
                if (skipzero && Sc<T>::iszero(xv)) continue;
                acc[s] = Sc<T>::fms(acc[s], tv[q[s]], xv);
            }
            const int32_t k = indexof(t);
            const T v = dg ? Sc<T>::div(acc[s], dg[k]) : acc[s];
            xs[t * R + r] = v;
            x[(int64_t)k * R + r] = v;
        }
        __syncthreads();
        // the later indices apply the terms that are ready now
#pragma unroll
        for (int s = 0; s < PP; s++) {
            if (tpos[s] <= t) continue;
            const int r = threadIdx.x + s * TB - tpos[s] * R;
            for (; q[s] < qe[s]; q[s]++) {
                const int32_t j = tidx[q[s]];
                const uint32_t jl = (uint32_t)(j - c0);
                T xv;
                if (jl < (uint32_t)w) {
                    const int pj = posof(j);
                    if (pj > t) break;
                    xv = xs[pj * R + r];
                } else {
                    xv = x[(int64_t)j * R + r];
                }
                if (skipzero && Sc<T>::iszero(xv)) continue;
                acc[s] = Sc<T>::fms(acc[s], tv[q[s]], xv);
            }
        }
    }
}

template <class T>
__global__ void k_lupper(int64_t nrows, const int32_t* __restrict__ rows, const int64_t* __restrict__ tp,
                         const int32_t* __restrict__ tidx, const double* __restrict__ tvald, const double* __restrict__ xd,
                         double* __restrict__ outd, int32_t R)
{
    const T* tv = reinterpret_cast<const T*>(tvald);
    const T* x = reinterpret_cast<const T*>(xd);
    T* out = reinterpret_cast<T*>(outd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nrows * R; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = idx / R, r = idx % R;
        T acc = x[(int64_t)rows[m] * R + r];
        for (int64_t q = tp[m]; q < tp[m + 1]; q++) {
            const T xv = x[(int64_t)tidx[q] * R + r];
            if (Sc<T>::iszero(xv)) continue;
            acc = Sc<T>::fms(acc, tv[q], xv);
        }
        out[idx] = acc;
    }
}

template <class T>
__global__ void k_lupper_put(int64_t nrows, const int32_t* __restrict__ rows, const double* __restrict__ outd,
                             double* __restrict__ xd, int32_t R)
{
    const T* out = reinterpret_cast<const T*>(outd);
    T* x = reinterpret_cast<T*>(xd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < nrows * R; idx += (int64_t)gridDim.x * blockDim.x)
        x[(int64_t)rows[idx / R] * R + idx % R] = out[idx];
}

__global__ void k_gather_conj(int vs, int64_t cnt, const int64_t* __restrict__ map, const double* __restrict__ src,
                              double* __restrict__ dst, int conj)
{
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < cnt; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t m = map[o];
        if (vs == 1) {
            dst[o] = src[m];
        } else {
            dst[2 * o] = src[2 * m];
            dst[2 * o + 1] = conj ? -src[2 * m + 1] : src[2 * m + 1];
        }
    }
}

template <class T>
__global__ void k_rhs_in(int64_t n, int32_t R, const double* __restrict__ bd, int64_t ldb, double* __restrict__ wd)
{
    const T* b = reinterpret_cast<const T*>(bd);
    T* w = reinterpret_cast<T*>(wd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n * R; idx += (int64_t)gridDim.x * blockDim.x)
        w[idx] = b[idx / R + (idx % R) * ldb];
}

template <class T>
__global__ void k_rhs_out(int64_t n, int32_t R, const double* __restrict__ wd, double* __restrict__ bd, int64_t ldb)
{
    const T* w = reinterpret_cast<const T*>(wd);
    T* b = reinterpret_cast<T*>(bd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n * R; idx += (int64_t)gridDim.x * blockDim.x)
        b[idx / R + (idx % R) * ldb] = w[idx];
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
void tsweep(int vs, int32_t ngrp, const int32_t* glist, const int2* grp, const int64_t* tp, const int32_t* nfirst,
            const int32_t* tidx, const double* tval, const double* diag, double* x, int32_t R, int dir, int skipzero,
            double bytes, hipStream_t s)
{
    if (ngrp <= 0) return;
    if (R < 1 || R > TS_RHS) throw GpuError("tsweep: right-hand sides per pass out of range");
    ProfScope ps("tsweep", bytes, s);
    if (vs == 1)
        hipLaunchKernelGGL(k_tsweep<double>, dim3(ngrp), dim3(TB), 0, s, glist, grp, tp, nfirst, tidx, tval, diag, x, R, dir, skipzero);
    else
        hipLaunchKernelGGL(k_tsweep<Zd>, dim3(ngrp), dim3(TB), 0, s, glist, grp, tp, nfirst, tidx, tval, diag, x, R, dir, skipzero);
}
void lupper(int vs, int64_t nrows, const int32_t* rows, const int64_t* tp, const int32_t* tidx, const double* tval,
            const double* x, double* out, int32_t R, hipStream_t s)
{
    if (nrows <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_lupper<double>, dim3(gridn(nrows * R)), dim3(TB), 0, s, nrows, rows, tp, tidx, tval, x, out, R);
    else hipLaunchKernelGGL(k_lupper<Zd>, dim3(gridn(nrows * R)), dim3(TB), 0, s, nrows, rows, tp, tidx, tval, x, out, R);
}
void lupper_put(int vs, int64_t nrows, const int32_t* rows, const double* out, double* x, int32_t R, hipStream_t s)
{
    if (nrows <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_lupper_put<double>, dim3(gridn(nrows * R)), dim3(TB), 0, s, nrows, rows, out, x, R);
    else hipLaunchKernelGGL(k_lupper_put<Zd>, dim3(gridn(nrows * R)), dim3(TB), 0, s, nrows, rows, out, x, R);
}
void gather_conj(int vs, int64_t cnt, const int64_t* map, const double* src, double* dst, int conj, hipStream_t s)
{
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_gather_conj, dim3(gridn(cnt)), dim3(TB), 0, s, vs, cnt, map, src, dst, conj);
}
void rhs_in(int vs, int64_t n, int32_t R, const double* b, int64_t ldb, double* w, hipStream_t s)
{
    if (n <= 0 || R <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_rhs_in<double>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, b, ldb, w);
    else hipLaunchKernelGGL(k_rhs_in<Zd>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, b, ldb, w);
}
void rhs_out(int vs, int64_t n, int32_t R, const double* w, double* b, int64_t ldb, hipStream_t s)
{
    if (n <= 0 || R <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_rhs_out<double>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, w, b, ldb);
    else hipLaunchKernelGGL(k_rhs_out<Zd>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, w, b, ldb);
}
}  // namespace k
}  // namespace fpa
