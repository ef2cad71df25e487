// factor_dev.hip -- the numeric kernels of the supernodal factorization (gfx950).
//
// A panel of P columns lives in a compact workspace W (column-major, ld = ldw): its rows are the rows the panel's
// columns touch -- first the rows already pivotal (the U segments), then the rows not yet pivotal -- found by the host
// search; rowmap[row] gives a row's position.  The kernels here
//   * scatter the panel's columns of A into W (scatter_cols);
//   * apply the earlier supernodes to the panel, slab by slab: the unit-lower solve of a slab's pivot rows
//     (slab_solve, the TRSM) and the rectangle below subtracted through rowmap (slab_gemm, the GEMM with the
//     relative-index scatter) -- one launch pair per dependency level, every slab of a level in one launch;
//   * factor the rows not yet pivotal by dense partial pivoting (getf2 per sub-panel of nb columns, then the
//     trailing TRSM and GEMM), LAPACK-style with whole-row swaps, so the pivot rows end at the top in column order.
// Entries of W outside a column's sparse pattern are exact zeros and stay so, which makes the dense panel LU produce
// the sparse factorization's values and pivots (a row outside the pattern has modulus 0 and cannot win).
#include "dev_common.h"
#include "device_api.h"
#include "factor_dev.h"
#include "prof.h"

namespace fpa {
namespace {

constexpr int TB = 256;

template <class T> __device__ inline T ld_val(const double* p, int64_t i) { return reinterpret_cast<const T*>(p)[i]; }
template <class T> __device__ inline void st_val(double* p, int64_t i, T v) { reinterpret_cast<T*>(p)[i] = v; }

__global__ void k_rowmap_set(int64_t cnt, const int32_t* __restrict__ rows, int32_t* __restrict__ rowmap, int reset)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < cnt; i += (int64_t)gridDim.x * blockDim.x)
        rowmap[rows[i]] = reset ? -1 : (int32_t)i;
}

// one wavefront per column: the entries placed through rowmap, the column's largest modulus (for the pivot threshold)
template <class T>
__global__ void k_scatter_cols(int32_t ncol, const int64_t* __restrict__ Bp, const int32_t* __restrict__ Bi,
                               const double* __restrict__ Bx, int64_t j0, const int32_t* __restrict__ rowmap,
                               double* __restrict__ W, int64_t ldw, double* __restrict__ colmax)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t ci = blockIdx.x * (TB / 64) + wid;
    if (ci >= ncol) return;
    const int64_t j = j0 + ci;
    double cm = 0.0;
    for (int64_t q = Bp[j] + lane; q < Bp[j + 1]; q += 64) {
        const T v = ld_val<T>(Bx, q);
        const int32_t r = rowmap[Bi[q]];
        st_val<T>(W, (int64_t)r + ci * ldw, v);
        const double a = Sc<T>::absval(v);
        if (a > cm) cm = a;
    }
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off));
    if (lane == 0) colmax[ci] = cm;
}

constexpr int SLAB = FPA_SLAB;   // columns per slab (the K of the update GEMM)
constexpr int RT = FPA_ROW_TILE; // rows per GEMM tile
// panel columns per tile: 64 real, 32 complex (LDS within 64 KB)
template <class T> struct CTile { static constexpr int v = 64; };
template <> struct CTile<Zd> { static constexpr int v = 32; };
constexpr int PNB = FPA_PANEL_NB;  // sub-panel width of the dense panel LU

// the unit-lower solve of a slab's pivot rows, in place in W, for up to CT panel columns
template <class T>
__global__ __launch_bounds__(TB) void k_slab_solve(const SlabTask* __restrict__ tasks, const int32_t* __restrict__ cols,
                                                    const BlockDesc* __restrict__ blocks, const int32_t* __restrict__ rowmap,
                                                    double* __restrict__ W, int64_t ldw, int32_t max_ctiles)
{
    constexpr int CT = CTile<T>::v;
    __shared__ T Ls[SLAB][SLAB + 1];
    __shared__ T Xs[SLAB][CT + 1];
    __shared__ int32_t rpos[SLAB];
    __shared__ int32_t cidx[CT];
    const int32_t t = blockIdx.x / max_ctiles, ct = blockIdx.x % max_ctiles;
    const SlabTask tk = tasks[t];
    const int32_t c0 = ct * CT;
    if (c0 >= tk.ncol) return;
    const int32_t nc = min(CT, tk.ncol - c0);
    const BlockDesc B = blocks[tk.block];
    const int32_t kk = tk.k1 - tk.k0;
    const T* V = reinterpret_cast<const T*>(B.vals);
    for (int i = threadIdx.x; i < kk; i += TB) rpos[i] = rowmap[B.rows[tk.k0 + i]];
    for (int i = threadIdx.x; i < nc; i += TB) cidx[i] = cols[tk.col_off + c0 + i];
    for (int i = threadIdx.x; i < kk * kk; i += TB) {
        const int r = i % kk, c = i / kk;
        Ls[r][c] = V[(int64_t)(tk.k0 + c) * B.nrows + tk.k0 + r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kk * nc; i += TB) {
        const int r = i % kk, c = i / kk;
        Xs[r][c] = ld_val<T>(W, (int64_t)rpos[r] + (int64_t)cidx[c] * ldw);
    }
    __syncthreads();
    // forward substitution: column c of X handled by the threads with tid % CT == c, rows split four ways
    const int c = threadIdx.x % CT, part = threadIdx.x / CT;
    for (int a = 0; a < kk - 1; a++) {
        if (c < nc) {
            const T xa = Xs[a][c];
            for (int r = a + 1 + part; r < kk; r += TB / CT) Xs[r][c] = Sc<T>::fms(Xs[r][c], Ls[r][a], xa);
        }
        __syncthreads();
    }
    for (int i = threadIdx.x; i < kk * nc; i += TB) {
        const int r = i % kk, cc = i / kk;
        st_val<T>(W, (int64_t)rpos[r] + (int64_t)cidx[cc] * ldw, Xs[r][cc]);
    }
}

// the rectangle below a slab: W[rows below] -= L[rows below, slab] * X[slab pivot rows], atomically (several slabs of
// one level may update the same rows)
template <class T>
__global__ __launch_bounds__(TB) void k_slab_gemm(const GemmTile* __restrict__ tiles, const SlabTask* __restrict__ tasks,
                                                   const int32_t* __restrict__ cols, const BlockDesc* __restrict__ blocks,
                                                   const int32_t* __restrict__ rowmap, double* __restrict__ W, int64_t ldw)
{
    constexpr int CT = CTile<T>::v, TC = CT / 16;
    __shared__ T Ls[RT][SLAB + 1];
    __shared__ T Xs[SLAB][CT + 1];
    __shared__ int32_t rpos[RT];
    __shared__ int32_t cidx[CT];
    const GemmTile tl = tiles[blockIdx.x];
    const SlabTask tk = tasks[tl.task];
    const BlockDesc B = blocks[tk.block];
    const int32_t kk = tk.k1 - tk.k0;
    const int32_t nr = min(RT, B.nrows - tl.row0), nc = min(CT, tk.ncol - tl.col0);
    const T* V = reinterpret_cast<const T*>(B.vals);
    for (int i = threadIdx.x; i < nr; i += TB) rpos[i] = rowmap[B.rows[tl.row0 + i]];
    for (int i = threadIdx.x; i < nc; i += TB) cidx[i] = cols[tk.col_off + tl.col0 + i];
    for (int i = threadIdx.x; i < nr * kk; i += TB) {
        const int r = i % nr, c = i / nr;
        Ls[r][c] = V[(int64_t)(tk.k0 + c) * B.nrows + tl.row0 + r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kk * nc; i += TB) {
        const int r = i % kk, c = i / kk;
        Xs[r][c] = ld_val<T>(W, (int64_t)rowmap[B.rows[tk.k0 + r]] + (int64_t)cidx[c] * ldw);
    }
    __syncthreads();
    // a 4 x TC register tile per thread over the RT x CT output tile
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * TC;
    T acc[4][TC];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TC; j++) acc[i][j] = Sc<T>::zero();
    for (int a = 0; a < kk; a++) {
        T l[4], x[TC];
#pragma unroll
        for (int i = 0; i < 4; i++) l[i] = Ls[tr + i][a];
#pragma unroll
        for (int j = 0; j < TC; j++) x[j] = Xs[a][tc + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < TC; j++) acc[i][j] = Sc<T>::fma_(acc[i][j], l[i], x[j]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TC; j++)
            if (tr + i < nr && tc + j < nc)
                atomic_sub(reinterpret_cast<T*>(W) + (int64_t)rpos[tr + i] + (int64_t)cidx[tc + j] * ldw, acc[i][j]);
}

// ---- dense partial-pivot LU of the panel's lower section

constexpr int LU_TB = 1024;

template <class T>
__global__ __launch_bounds__(LU_TB) void k_getf2(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a,
                                                  int32_t nb, const double* __restrict__ colmax, double pivot_rel,
                                                  int32_t* __restrict__ piv, PanelStatus* __restrict__ st)
{
    __shared__ double bestv[LU_TB / 64];
    __shared__ int32_t bestr[LU_TB / 64];
    __shared__ int32_t pr;
    T* Wl = reinterpret_cast<T*>(Wd);
    if (st->err) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t jend = min(a + nb, P);
    for (int32_t j = a; j < jend; j++) {
        // argmax of the modulus over rows j .. nr-1, the lowest row on ties
        double bv = -1.0;
        int32_t br = -1;
        for (int32_t r = j + threadIdx.x; r < nr; r += LU_TB) {
            const double v = Sc<T>::absval(Wl[r + j * ld]);
            if (v > bv) { bv = v; br = r; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int32_t orr = __shfl_xor(br, off);
            if (ov > bv || (ov == bv && orr >= 0 && (br < 0 || orr < br))) { bv = ov; br = orr; }
        }
        if (lane == 0) { bestv[wid] = bv; bestr[wid] = br; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double v = -1.0;
            int32_t r = -1;
            for (int w = 0; w < LU_TB / 64; w++)
                if (bestv[w] > v || (bestv[w] == v && bestr[w] >= 0 && (r < 0 || bestr[w] < r))) { v = bestv[w]; r = bestr[w]; }
            const double thr = pivot_rel * colmax[j];
            if (r < 0 || v <= 0.0 || v < thr) {
                st->err = j + 1;
                st->best = v > 0.0 ? v : 0.0;
                st->thr = thr;
                pr = -1;
            } else {
                pr = r;
                piv[j] = r;
            }
        }
        __syncthreads();
        const int32_t r = pr;
        if (r < 0) return;
        if (r != j)
            for (int32_t c = threadIdx.x; c < P; c += LU_TB) {
                const T t0 = Wl[j + c * ld];
                Wl[j + c * ld] = Wl[r + c * ld];
                Wl[r + c * ld] = t0;
            }
        __syncthreads();
        const T pv = Wl[j + j * ld];
        for (int32_t rr = j + 1 + threadIdx.x; rr < nr; rr += LU_TB) {
            const T l = Sc<T>::div(Wl[rr + j * ld], pv);
            Wl[rr + j * ld] = l;
            for (int32_t c = j + 1; c < jend; c++) Wl[rr + c * ld] = Sc<T>::fms(Wl[rr + c * ld], l, Wl[j + c * ld]);
        }
        __syncthreads();
    }
}

// rows [a, a + nb) of the columns right of the sub-panel: the unit-lower solve with the sub-panel's pivot block
template <class T>
__global__ __launch_bounds__(TB) void k_trsm_sub(double* __restrict__ Wd, int64_t ld, int32_t P, int32_t a, int32_t nb,
                                                  const PanelStatus* __restrict__ st)
{
    __shared__ T Ls[PNB][PNB + 1];
    if (st->err) return;
    T* Wl = reinterpret_cast<T*>(Wd);
    for (int i = threadIdx.x; i < nb * nb; i += TB) {
        const int r = i % nb, c = i / nb;
        Ls[r][c] = Wl[(a + r) + (int64_t)(a + c) * ld];
    }
    __syncthreads();
    const int32_t c = a + nb + blockIdx.x * TB + threadIdx.x;
    if (c >= P) return;
    T x[PNB];
#pragma unroll
    for (int i = 0; i < PNB; i++) x[i] = i < nb ? Wl[(a + i) + (int64_t)c * ld] : Sc<T>::zero();
#pragma unroll
    for (int i = 1; i < PNB; i++)
#pragma unroll
        for (int k = 0; k < i; k++) x[i] = Sc<T>::fms(x[i], Ls[i][k], x[k]);
    for (int i = 0; i < nb; i++) Wl[(a + i) + (int64_t)c * ld] = x[i];
}

// W[a+nb:nr, a+nb:P] -= W[a+nb:nr, a:a+nb] * W[a:a+nb, a+nb:P]
template <class T>
__global__ __launch_bounds__(TB) void k_gemm_sub(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a,
                                                  int32_t nb, const PanelStatus* __restrict__ st)
{
    constexpr int CT = CTile<T>::v, TC = CT / 16;
    __shared__ T As[RT][PNB + 1];
    __shared__ T Bs[PNB][CT + 1];
    if (st->err) return;
    T* Wl = reinterpret_cast<T*>(Wd);
    const int32_t r0 = a + nb + blockIdx.x * RT, c0 = a + nb + blockIdx.y * CT;
    if (r0 >= nr || c0 >= P) return;
    const int32_t mr = min(RT, nr - r0), mc = min(CT, P - c0);
    for (int i = threadIdx.x; i < mr * nb; i += TB) {
        const int r = i % mr, c = i / mr;
        As[r][c] = Wl[(r0 + r) + (int64_t)(a + c) * ld];
    }
    for (int i = threadIdx.x; i < nb * mc; i += TB) {
        const int r = i % nb, c = i / nb;
        Bs[r][c] = Wl[(a + r) + (int64_t)(c0 + c) * ld];
    }
    __syncthreads();
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * TC;
    T acc[4][TC];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TC; j++) acc[i][j] = Sc<T>::zero();
    for (int k = 0; k < nb; k++) {
        T l[4], x[TC];
#pragma unroll
        for (int i = 0; i < 4; i++) l[i] = As[tr + i][k];
#pragma unroll
        for (int j = 0; j < TC; j++) x[j] = Bs[k][tc + j];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < TC; j++) acc[i][j] = Sc<T>::fma_(acc[i][j], l[i], x[j]);
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TC; j++)
            if (tr + i < mr && tc + j < mc) {
                T* p = Wl + (r0 + tr + i) + (int64_t)(c0 + tc + j) * ld;
                *p = Sc<T>::sub(*p, acc[i][j]);
            }
}

template <class T>
__global__ void k_extract_cols(const ExtractCol* __restrict__ d, const int32_t* __restrict__ pos, const double* __restrict__ Wl,
                               int64_t ldw)
{
    const ExtractCol e = d[blockIdx.x];
    const T* src = reinterpret_cast<const T*>(Wl) + (int64_t)e.ci * ldw;
    T* dst = reinterpret_cast<T*>(e.dst);
    for (int32_t r = threadIdx.x; r < e.nrows; r += TB) dst[r] = src[pos[e.pos_off + r]];
}

inline unsigned gridn(int64_t n, int per)
{
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

}  // namespace

namespace k {
void extract_cols(int vs, int32_t ncols, const ExtractCol* d, const int32_t* pos, const double* Wl, int64_t ldw,
                  hipStream_t s)
{
    if (ncols <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_extract_cols<double>, dim3(ncols), dim3(TB), 0, s, d, pos, Wl, ldw);
    else hipLaunchKernelGGL(k_extract_cols<Zd>, dim3(ncols), dim3(TB), 0, s, d, pos, Wl, ldw);
}
void rowmap_set(int64_t cnt, const int32_t* rows, int32_t* rowmap, int reset, hipStream_t s)
{
    if (cnt > 0) hipLaunchKernelGGL(k_rowmap_set, dim3(gridn(cnt, TB)), dim3(TB), 0, s, cnt, rows, rowmap, reset);
}
void scatter_cols(int vs, int32_t ncol, const int64_t* Bp, const int32_t* Bi, const double* Bx, int64_t j0,
                  const int32_t* rowmap, double* W, int64_t ldw, double* colmax, hipStream_t s)
{
    if (ncol <= 0) return;
    const unsigned g = (unsigned)((ncol + TB / 64 - 1) / (TB / 64));
    if (vs == 1) hipLaunchKernelGGL(k_scatter_cols<double>, dim3(g), dim3(TB), 0, s, ncol, Bp, Bi, Bx, j0, rowmap, W, ldw, colmax);
    else hipLaunchKernelGGL(k_scatter_cols<Zd>, dim3(g), dim3(TB), 0, s, ncol, Bp, Bi, Bx, j0, rowmap, W, ldw, colmax);
}
void slab_solve(int vs, int32_t ntask, const SlabTask* tasks, const int32_t* cols, const BlockDesc* blocks,
                const int32_t* rowmap, double* W, int64_t ldw, int32_t max_ctiles, hipStream_t s)
{
    if (ntask <= 0) return;
    ProfScope ps("slab_solve", 0.0, s);
    const unsigned g = (unsigned)(ntask * max_ctiles);
    if (vs == 1) hipLaunchKernelGGL(k_slab_solve<double>, dim3(g), dim3(TB), 0, s, tasks, cols, blocks, rowmap, W, ldw, max_ctiles);
    else hipLaunchKernelGGL(k_slab_solve<Zd>, dim3(g), dim3(TB), 0, s, tasks, cols, blocks, rowmap, W, ldw, max_ctiles);
}
void slab_gemm(int vs, int32_t ntile, const GemmTile* tiles, const SlabTask* tasks, const int32_t* cols,
               const BlockDesc* blocks, const int32_t* rowmap, double* W, int64_t ldw, double flops, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("slab_gemm", flops, s);
    if (vs == 1) hipLaunchKernelGGL(k_slab_gemm<double>, dim3(ntile), dim3(TB), 0, s, tiles, tasks, cols, blocks, rowmap, W, ldw);
    else hipLaunchKernelGGL(k_slab_gemm<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, tasks, cols, blocks, rowmap, W, ldw);
}
void getf2(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const double* colmax,
           double pivot_rel, int32_t* piv, PanelStatus* st, hipStream_t s)
{
    ProfScope ps("getf2", (vs == 1 ? 2.0 : 8.0) * (double)nr * nb * nb / 2, s);
    if (vs == 1) hipLaunchKernelGGL(k_getf2<double>, dim3(1), dim3(LU_TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, piv, st);
    else hipLaunchKernelGGL(k_getf2<Zd>, dim3(1), dim3(LU_TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, piv, st);
}
void trsm_sub(int vs, double* Wl, int64_t ld, int32_t P, int32_t a, int32_t nb, const PanelStatus* st, hipStream_t s)
{
    const int32_t nc = P - a - nb;
    if (nc <= 0) return;
    ProfScope ps("trsm_sub", 0.0, s);
    const unsigned g = (unsigned)((nc + TB - 1) / TB);
    if (vs == 1) hipLaunchKernelGGL(k_trsm_sub<double>, dim3(g), dim3(TB), 0, s, Wl, ld, P, a, nb, st);
    else hipLaunchKernelGGL(k_trsm_sub<Zd>, dim3(g), dim3(TB), 0, s, Wl, ld, P, a, nb, st);
}
void gemm_sub(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const PanelStatus* st,
              hipStream_t s)
{
    const int32_t mr = nr - a - nb, mc = P - a - nb;
    if (mr <= 0 || mc <= 0) return;
    ProfScope ps("gemm_sub", (vs == 1 ? 2.0 : 8.0) * mr * (double)mc * nb, s);
    const int ct = vs == 1 ? CTile<double>::v : CTile<Zd>::v;
    const dim3 g((unsigned)((mr + RT - 1) / RT), (unsigned)((mc + ct - 1) / ct));
    if (vs == 1) hipLaunchKernelGGL(k_gemm_sub<double>, g, dim3(TB), 0, s, Wl, ld, nr, P, a, nb, st);
    else hipLaunchKernelGGL(k_gemm_sub<Zd>, g, dim3(TB), 0, s, Wl, ld, nr, P, a, nb, st);
}
}  // namespace k

}  // namespace fpa
