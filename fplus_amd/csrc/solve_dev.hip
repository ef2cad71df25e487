// solve_dev.hip -- the triangular solves on the device factors (gfx950).
//
// L is held as supernode blocks (factor_dev.h), U in CSC (rows in pivot order, the diagonal last in every column); a
// U column is a list of runs of consecutive rows, which the solves with A read per block.  The right-hand sides live row-major during the sweeps, x[i * R + r] for
// R right-hand sides, so one entry of a factor meets all R values of a row in one contiguous read.  A sweep walks the
// blocks level by level (a level's blocks are independent), one launch per level and step, and every sweep pulls: each
// value is written by one thread, no atomics, so the results are the same on every run.  The sweeps use a second vector
// y: L y = b leaves its result in y, U x = y its result in x; U^T y = b leaves its result in y, L^T x = y in x.
//   lfwd_gather L y = b   a level's pivot rows gather the contributions the blocks before them left in their slots
//                         (one wavefront per row, in x);
//   lfwd_pull   y on the block's rows = L11^-1 x (the inverse kept from the factorization);
//   lfwd_push   the block's rows below: contribution[slot + r] = L(r, block) y(block) (the slot is the row's own);
//   ub_dots     U x = y   a block's pivot rows: dot products along their U rows beyond the block (the runs of later
//   ub_red                U columns over the block's rows, in column order) with x, subtracted from y;
//   ub_prod     x on the block's rows = U11^-1 y;
//   utdots      U^T y = b (U^H): dot products down U's columns with the solved rows in y, subtracted in x;
//   ut_prod     y on the block's rows = U11^-T x (conjugated for U^H);
//   ltdots      L^T x = y (L^H): dot products down the block's columns with the solved rows in x, subtracted in y;
//   lt_prod     x on the block's rows = L11^-T y.
// Blocks are at most 256 columns wide (the panel width): a block's input for all right-hand sides fits in LDS.
#include "dev_common.h"
#include "device_api.h"
#include "prof.h"
#include "solve_dev.h"

namespace fpa {
namespace {

constexpr int TB = 256;
// right-hand sides per pass (a block's w x R values in at most 64 KB of LDS)
template <class T> struct RM { static constexpr int v = 16; };

template <class T> __device__ inline T cj(T v, int c) { return c ? Sc<T>::conj(v) : v; }

template <class T>
__device__ inline T wave_sum(T v)
{
    for (int off = 32; off > 0; off >>= 1) {
        if constexpr (Sc<T>::vs == 1) v += __shfl_xor(v, off);
        else { v.x += __shfl_xor(v.x, off); v.y += __shfl_xor(v.y, off); }
    }
    return v;
}

// L y = b, part 0: the pivot rows of a level's blocks subtract the contributions waiting in their slots, one wavefront
// per row: lane = 4 g + q, group g (of 16) summing the row's slots g, g + 16, ... for the right-hand sides q, q + 4,
// ...; the groups' sums combined in a fixed tree
template <class T>
__global__ __launch_bounds__(TB) void k_lfwd_gather(int32_t nrow, const int32_t* __restrict__ rows, const int64_t* __restrict__ incp,
                                                     const int64_t* __restrict__ incs, const double* __restrict__ contribd,
                                                     double* __restrict__ xd, int32_t R)
{
    constexpr int RQ = RM<T>::v / 4;
    const T* contrib = reinterpret_cast<const T*>(contribd);
    T* x = reinterpret_cast<T*>(xd);
    const int lane = threadIdx.x & 63, g = lane >> 2, q = lane & 3;
    const int32_t it = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    if (it >= nrow) return;
    const int64_t i = rows[it], q0 = incp[i], q1 = incp[i + 1];
    T acc[RQ];
#pragma unroll
    for (int u = 0; u < RQ; u++) acc[u] = Sc<T>::zero();
    for (int64_t p = q0 + g; p < q1; p += 16) {
        const int64_t sl = incs[p] * R;
#pragma unroll
        for (int u = 0; u < RQ; u++) {
            const int r = q + 4 * u;
            if (r < R) acc[u] = Sc<T>::add(acc[u], contrib[sl + r]);
        }
    }
#pragma unroll
    for (int u = 0; u < RQ; u++) {
        T v = acc[u];
        for (int off = 32; off >= 4; off >>= 1) {
            if constexpr (Sc<T>::vs == 1) v += __shfl_xor(v, off);
            else { v.x += __shfl_xor(v.x, off); v.y += __shfl_xor(v.y, off); }
        }
        const int r = q + 4 * u;
        if (g == 0 && r < R) x[i * R + r] = Sc<T>::sub(x[i * R + r], v);
    }
}

// The blocks' diagonal triangles in the sweeps, as products with the inverses kept from the factorization:
// out(block rows) = Op in(block rows), Op = L11^-1 (lower, unit diagonal) or U11^-1 (upper) for the sweeps with A, and
// their transposes (conjugated for the sweeps with A^H: conj) for those with A^T; the inverses are w x w column-major
// (ld = w), so Op(row, c) = M[row + c w], or M[c + row w] transposed.  UPPER: Op is upper triangular.  One workgroup
// per 64 rows of a block: lane = row; the four waves take a quarter each of the columns the rows reach (lower: up to
// the chunk's last row, upper: from its first row), with the block's input for those columns and all right-hand sides
// in LDS; the quarters' sums are added in wave order -- a fixed order that does not depend on the number of right-hand
// sides.  The output goes to another vector: the block's input is still read by its other row chunks.
template <class T, bool UPPER, bool TRANS>
__global__ __launch_bounds__(TB) void k_tri_prod(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                  const double* __restrict__ ind, double* __restrict__ outd, int32_t R,
                                                  int conj)
{
    constexpr int RC = RM<T>::v;
    __shared__ T xs[TB * RC];
    const int2 tl = tiles[blockIdx.x];   // block, first row
    const BlockDesc B = blocks[tl.x];
    // (global pointers: a generic load would make each LDS wait below wait for the matrix loads in flight too)
    const gptr<const T> M = as_global(reinterpret_cast<const T*>(UPPER != TRANS ? B.uinv : B.linv));
    const T* in = reinterpret_cast<const T*>(ind);
    T* out = reinterpret_cast<T*>(outd);
    const int w = B.w, r0 = tl.y;
    const int ca = UPPER ? r0 : 0, cb = UPPER ? w : min(w, r0 + 64), ncol = cb - ca;
    for (int idx = threadIdx.x; idx < ncol * R; idx += TB) {
        const int c = idx / R, r = idx - c * R;
        xs[c * RC + r] = in[((int64_t)B.c0 + ca + c) * R + r];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6, row = r0 + lane;
    const int per = (ncol + 3) / 4, qa = ca + q * per, qb = min(cb, qa + per);
    T acc[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = Sc<T>::zero();
    if (row < w) {
        const int lo = UPPER ? max(qa, row) : qa, hi = UPPER ? qb : min(qb, row + 1);
#pragma unroll 8
        for (int c = lo; c < hi; c++) {
            const T v = TRANS ? cj(gload(M, c + (int64_t)row * w), conj) : gload(M, row + (int64_t)c * w);
#pragma unroll
            for (int r = 0; r < RC; r++)
                if (r < R) acc[r] = Sc<T>::fma_(acc[r], v, xs[(c - ca) * RC + r]);
        }
    }
    __syncthreads();   // the LDS now takes the partial sums of waves 1 .. 3
    if (q > 0) {
#pragma unroll
        for (int r = 0; r < RC; r++) xs[((q - 1) * 64 + lane) * RC + r] = acc[r];
    }
    __syncthreads();
    if (q == 0 && row < w) {
        for (int qq = 0; qq < 3; qq++) {
#pragma unroll
            for (int r = 0; r < RC; r++) acc[r] = Sc<T>::add(acc[r], xs[(qq * 64 + lane) * RC + r]);
        }
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) out[((int64_t)B.c0 + row) * R + r] = acc[r];
    }
}

// L y = b, part 2: the rows below a block, 64 rows per workgroup (lane = row), the four waves a quarter of the block's
// columns each (the block's y for all columns and right-hand sides in LDS), the quarters' sums added in wave order;
// slot_off[b] is the block's first slot
template <class T>
__global__ __launch_bounds__(TB) void k_lfwd_push(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                   const int64_t* __restrict__ slot_off, const double* __restrict__ yd,
                                                   double* __restrict__ contribd, int32_t R)
{
    constexpr int RC = RM<T>::v;
    __shared__ T ys[TB * RC];
    const int2 tl = tiles[blockIdx.x];   // block, first row below the block (0-based)
    const BlockDesc B = blocks[tl.x];
    const gptr<const T> V = as_global(reinterpret_cast<const T*>(B.vals));   // a global pointer, as in k_tri_prod
    const T* y = reinterpret_cast<const T*>(yd);
    T* contrib = reinterpret_cast<T*>(contribd);
    const int w = B.w;
    for (int idx = threadIdx.x; idx < w * R; idx += TB) {
        const int c = idx / R, r = idx - c * R;
        ys[c * RC + r] = y[((int64_t)B.c0 + c) * R + r];
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int32_t row = w + tl.y + lane;
    const bool own = row < B.nrows;
    const int per = (w + 3) / 4, qa = q * per, qb = min(w, qa + per);
    T acc[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = Sc<T>::zero();
    if (own) {
#pragma unroll 8
        for (int c = qa; c < qb; c++) {
            const T v = gload(V, (int64_t)c * B.nrows + row);
#pragma unroll
            for (int r = 0; r < RC; r++)
                if (r < R) acc[r] = Sc<T>::fma_(acc[r], v, ys[c * RC + r]);
        }
    }
    __syncthreads();
    if (q > 0) {
#pragma unroll
        for (int r = 0; r < RC; r++) ys[((q - 1) * 64 + lane) * RC + r] = acc[r];
    }
    __syncthreads();
    if (q == 0 && own) {
        for (int qq = 0; qq < 3; qq++) {
#pragma unroll
            for (int r = 0; r < RC; r++) acc[r] = Sc<T>::add(acc[r], ys[(qq * 64 + lane) * RC + r]);
        }
        const int64_t slot = slot_off[tl.x] + row - w;
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) contrib[slot * R + r] = acc[r];
    }
}

// U x = y, part 1: the products of a block's pivot rows with their U rows beyond the block, one share of the block's
// pieces per workgroup.  A piece is a later column j's run over the block's rows [a, w), the pieces in column order: thread t
// owns block row t and adds U(t, j) x_j for every piece with a <= t, so consecutive rows read consecutive U entries; the
// pieces go in batches of UB_PC, whose x_j for all right-hand sides sit in LDS.  A share's sums go to its own slot of
// `part`; ub_red subtracts a block's slots from x in slot order (a fixed order, the same for any number of right-hand
// sides).
template <class T>
__global__ __launch_bounds__(TB) void k_ub_dots(const int4* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                 const int32_t* __restrict__ pa, const int32_t* __restrict__ pj,
                                                 const int64_t* __restrict__ pu, const double* __restrict__ Uxd,
                                                 const double* __restrict__ xd, double* __restrict__ partd, int32_t R)
{
    constexpr int RC = RM<T>::v, PC = k::UB_PC;
    __shared__ int32_t sa[PC];
    __shared__ int64_t su[PC];
    __shared__ T sx[PC * RC];
    const int4 tl = tiles[blockIdx.x];   // block, first piece, pieces, slot
    const int w = blocks[tl.x].w, t = threadIdx.x;
    const T* Ux = reinterpret_cast<const T*>(Uxd);
    const T* x = reinterpret_cast<const T*>(xd);
    T acc[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = Sc<T>::zero();
    for (int32_t b0 = 0; b0 < tl.z; b0 += PC) {
        const int npc = min(PC, tl.z - b0), p0 = tl.y + b0;
        __syncthreads();   // the previous batch is done with the LDS
        if (t < npc) {
            sa[t] = pa[p0 + t];
            su[t] = pu[p0 + t];
        }
        for (int idx = t; idx < npc * R; idx += TB) {
            const int q = idx / R, r = idx - q * R;
            sx[q * RC + r] = x[(int64_t)pj[p0 + q] * R + r];
        }
        __syncthreads();
        if (t < w) {
#pragma unroll 4
            for (int q = 0; q < npc; q++) {
                const T u = sa[q] <= t ? Ux[su[q] + t] : Sc<T>::zero();
#pragma unroll
                for (int r = 0; r < RC; r++)
                    if (r < R) acc[r] = Sc<T>::fma_(acc[r], u, sx[q * RC + r]);
            }
        }
    }
    if (t >= w) return;
    T* part = reinterpret_cast<T*>(partd) + ((int64_t)tl.w * TB + t) * RC;
#pragma unroll
    for (int r = 0; r < RC; r++)
        if (r < R) part[r] = acc[r];
}

// U x = y, between the parts: x minus a block's partial sums, 16 rows of the block per workgroup (thread: row
// row0 + tid / 16, right-hand side tid % 16), the slots in order
template <class T>
__global__ __launch_bounds__(TB) void k_ub_red(const int4* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                const double* __restrict__ partd, double* __restrict__ xd, int32_t R)
{
    constexpr int RC = RM<T>::v;
    const int4 tl = tiles[blockIdx.x];   // block, first row, first slot, slots
    const int t = tl.y + (int)threadIdx.x / RC, r = (int)threadIdx.x % RC;
    const BlockDesc B = blocks[tl.x];
    if (t >= B.w || r >= R) return;
    const T* part = reinterpret_cast<const T*>(partd) + ((int64_t)tl.z * TB + t) * RC + r;
    T s0 = Sc<T>::zero(), s1 = Sc<T>::zero();
    int c = 0;
    for (; c + 1 < tl.w; c += 2) {   // two chains, joined at the end: the order is fixed by the slot count alone
        s0 = Sc<T>::add(s0, part[(int64_t)c * TB * RC]);
        s1 = Sc<T>::add(s1, part[(int64_t)(c + 1) * TB * RC]);
    }
    if (c < tl.w) s0 = Sc<T>::add(s0, part[(int64_t)c * TB * RC]);
    T* xp = reinterpret_cast<T*>(xd) + ((int64_t)B.c0 + t) * R + r;
    *xp = Sc<T>::sub(*xp, Sc<T>::add(s0, s1));
}

// the 256 threads' sums for all right-hand sides added in a fixed order (lanes by a shuffle tree, then the four
// wavefronts in order through LDS); thread 0 ends with the total
template <class T>
__device__ inline void block_sum(T (&acc)[RM<T>::v], T* red)
{
    constexpr int RC = RM<T>::v;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = wave_sum(acc[r]);
    if (lane == 0 && wid > 0) {
#pragma unroll
        for (int r = 0; r < RC; r++) red[(wid - 1) * RC + r] = acc[r];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < TB / 64 - 1; q++) {
#pragma unroll
            for (int r = 0; r < RC; r++) acc[r] = Sc<T>::add(acc[r], red[q * RC + r]);
        }
    }
}

// U^T y = b, part 1: the dot products of a level's block columns over the rows before their block, one workgroup per
// column (tiles: block, column): x_c -= sum_{i < c0} conj?(U(i, c)) y_i, y the solved rows of earlier blocks
template <class T>
__global__ __launch_bounds__(TB) void k_utdots(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                const int64_t* __restrict__ Up, const int32_t* __restrict__ Ui,
                                                const double* __restrict__ Uxd, const double* __restrict__ yd,
                                                double* __restrict__ xd, int32_t R, int conj)
{
    constexpr int RC = RM<T>::v;
    __shared__ T red[(TB / 64 - 1) * RC];
    const int2 tl = tiles[blockIdx.x];
    const BlockDesc B = blocks[tl.x];
    const T* Ux = reinterpret_cast<const T*>(Uxd);
    const T* y = reinterpret_cast<const T*>(yd);
    T* x = reinterpret_cast<T*>(xd);
    const int c = tl.y;
    const int64_t c0 = B.c0, q0 = Up[c0 + c], qd = Up[c0 + c + 1] - 1;
    T acc[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = Sc<T>::zero();
#pragma unroll 2
    for (int64_t q = q0 + threadIdx.x; q < qd; q += TB) {
        const int64_t i = Ui[q];
        if (i >= c0) continue;
        const T u = cj(Ux[q], conj);
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) acc[r] = Sc<T>::fma_(acc[r], u, y[i * R + r]);
    }
    block_sum<T>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) x[(c0 + c) * R + r] = Sc<T>::sub(x[(c0 + c) * R + r], acc[r]);
    }
}

// L^T x = y, part 1: the dot products of a level's block columns over the block's rows below, one workgroup per
// column: y_c -= sum_{rows i below} conj?(L(i, c)) x_pinv(i), x the solved rows of later blocks
template <class T>
__global__ __launch_bounds__(TB) void k_ltdots(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                const int32_t* __restrict__ pinv_rows, const double* __restrict__ xd,
                                                double* __restrict__ yd, int32_t R, int conj)
{
    constexpr int RC = RM<T>::v;
    __shared__ T red[(TB / 64 - 1) * RC];
    const int2 tl = tiles[blockIdx.x];
    const BlockDesc B = blocks[tl.x];
    const T* V = reinterpret_cast<const T*>(B.vals);
    const T* x = reinterpret_cast<const T*>(xd);
    T* y = reinterpret_cast<T*>(yd);
    const int c = tl.y, w = B.w;
    T acc[RC];
#pragma unroll
    for (int r = 0; r < RC; r++) acc[r] = Sc<T>::zero();
#pragma unroll 2
    for (int32_t i = w + threadIdx.x; i < B.nrows; i += TB) {
        const T v = cj(V[(int64_t)c * B.nrows + i], conj);
        const int64_t pi = pinv_rows[B.rows[i]];
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) acc[r] = Sc<T>::fma_(acc[r], v, x[pi * R + r]);
    }
    block_sum<T>(acc, red);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int r = 0; r < RC; r++)
            if (r < R) y[(B.c0 + c) * R + r] = Sc<T>::sub(y[(B.c0 + c) * R + r], acc[r]);
    }
}

// w[k * R + r] = D[perm[k]] * b[perm[k] + r * ldb]
template <class T>
__global__ void k_gather_scale(int64_t n, int32_t R, const int64_t* __restrict__ perm, const double* __restrict__ D,
                               const double* __restrict__ bd, int64_t ldb, double* __restrict__ wd)
{
#pragma clang fp contract(off)
    const T* b = reinterpret_cast<const T*>(bd);
    T* w = reinterpret_cast<T*>(wd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n * R; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = idx / R, r = idx % R, src = perm[k];
        w[idx] = Sc<T>::scale(b[src + r * ldb], D[src]);
    }
}

// b[perm[k] + r * ldb] = D[perm[k]] * w[k * R + r]
template <class T>
__global__ void k_scatter_scale(int64_t n, int32_t R, const int64_t* __restrict__ perm, const double* __restrict__ D,
                                const double* __restrict__ wd, double* __restrict__ bd, int64_t ldb)
{
#pragma clang fp contract(off)
    const T* w = reinterpret_cast<const T*>(wd);
    T* b = reinterpret_cast<T*>(bd);
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < n * R; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = idx / R, r = idx % R, dst = perm[k];
        b[dst + r * ldb] = Sc<T>::scale(w[idx], D[dst]);
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
int solve_rhs_per_pass(int vs) { return vs == 1 ? RM<double>::v : RM<Zd>::v; }

void lfwd_gather(int vs, int32_t nrow, const int32_t* rows, const int64_t* incp, const int64_t* incs, const double* contrib,
                 double* x, int32_t R, double bytes, hipStream_t s)
{
    if (nrow <= 0) return;
    ProfScope ps("lfwd_gather", bytes, s);
    const unsigned g = (unsigned)((nrow + TB / 64 - 1) / (TB / 64));
    if (vs == 1) hipLaunchKernelGGL(k_lfwd_gather<double>, dim3(g), dim3(TB), 0, s, nrow, rows, incp, incs, contrib, x, R);
    else hipLaunchKernelGGL(k_lfwd_gather<Zd>, dim3(g), dim3(TB), 0, s, nrow, rows, incp, incs, contrib, x, R);
}
void lfwd_pull(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* x, double* y, int32_t R,
               double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("lfwd_pull", bytes, s);
    if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, false, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
    else hipLaunchKernelGGL((k_tri_prod<Zd, false, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
}
void lfwd_push(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int64_t* slot_off, const double* y,
               double* contrib, int32_t R, double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("lfwd_push", bytes, s);
    if (vs == 1) hipLaunchKernelGGL(k_lfwd_push<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, slot_off, y, contrib, R);
    else hipLaunchKernelGGL(k_lfwd_push<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, slot_off, y, contrib, R);
}
void ub_dots(int vs, int32_t ntile, const int4* tiles, const BlockDesc* blocks, const int32_t* pa, const int32_t* pj,
             const int64_t* pu, const double* Ux, const double* x, double* part, int32_t R, double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("ub_dots", bytes, s);
    if (vs == 1) hipLaunchKernelGGL(k_ub_dots<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, pa, pj, pu, Ux, x, part, R);
    else hipLaunchKernelGGL(k_ub_dots<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, pa, pj, pu, Ux, x, part, R);
}
void ub_red(int vs, int32_t ntile, const int4* tiles, const BlockDesc* blocks, const double* part, double* x, int32_t R,
            hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("ub_red", 0.0, s);
    if (vs == 1) hipLaunchKernelGGL(k_ub_red<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, part, x, R);
    else hipLaunchKernelGGL(k_ub_red<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, part, x, R);
}
void ub_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* y, double* x, int32_t R,
             double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("ub_prod", bytes, s);
    if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, true, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
    else hipLaunchKernelGGL((k_tri_prod<Zd, true, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
}
void utdots(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int64_t* Up, const int32_t* Ui,
            const double* Ux, const double* y, double* x, int32_t R, int conj, double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("utdots", bytes, s);
    if (vs == 1) hipLaunchKernelGGL(k_utdots<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, Up, Ui, Ux, y, x, R, conj);
    else hipLaunchKernelGGL(k_utdots<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, Up, Ui, Ux, y, x, R, conj);
}
void ltdots(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int32_t* pinv_rows, const double* x,
            double* y, int32_t R, int conj, double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("ltdots", bytes, s);
    if (vs == 1) hipLaunchKernelGGL(k_ltdots<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, pinv_rows, x, y, R, conj);
    else hipLaunchKernelGGL(k_ltdots<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, pinv_rows, x, y, R, conj);
}
void ut_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* x, double* y, int32_t R, int conj,
             double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("ut_prod", bytes, s);
    if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, false, true>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, conj);
    else hipLaunchKernelGGL((k_tri_prod<Zd, false, true>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, conj);
}
void lt_prod(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const double* y, double* x, int32_t R, int conj,
             double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("lt_prod", bytes, s);
    if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, true, true>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, conj);
    else hipLaunchKernelGGL((k_tri_prod<Zd, true, true>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, conj);
}
void gather_scale(int vs, int64_t n, int32_t R, const int64_t* perm, const double* D, const double* b, int64_t ldb, double* w,
                  hipStream_t s)
{
    if (n <= 0 || R <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_gather_scale<double>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, perm, D, b, ldb, w);
    else hipLaunchKernelGGL(k_gather_scale<Zd>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, perm, D, b, ldb, w);
}
void scatter_scale(int vs, int64_t n, int32_t R, const int64_t* perm, const double* D, const double* w, double* b, int64_t ldb,
                   hipStream_t s)
{
    if (n <= 0 || R <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_scatter_scale<double>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, perm, D, w, b, ldb);
    else hipLaunchKernelGGL(k_scatter_scale<Zd>, dim3(gridn(n * R)), dim3(TB), 0, s, n, R, perm, D, w, b, ldb);
}
}  // namespace k
}  // namespace fpa
