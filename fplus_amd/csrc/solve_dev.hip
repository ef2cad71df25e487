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
#include <algorithm>
#include <atomic>

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
__device__ __forceinline__ void gather_row(int64_t i, const int64_t* __restrict__ incp, const int64_t* __restrict__ incs,
                                  const T* __restrict__ contrib, T* __restrict__ x, int32_t R)
{
    constexpr int RQ = RM<T>::v / 4;
    const int lane = threadIdx.x & 63, g = lane >> 2, q = lane & 3;
    const int64_t q0 = incp[i], q1 = incp[i + 1];
    T acc[RQ];
#pragma unroll
    for (int u = 0; u < RQ; u++) acc[u] = Sc<T>::zero();
    // the lane's slots g, g + 16, .. in batches of 4: the slot numbers, then the values, loaded before they are added
    for (int64_t p0 = q0 + g; p0 < q1; p0 += 64) {
        int64_t sl[4];
#pragma unroll
        for (int j = 0; j < 4; j++) sl[j] = incs[min(p0 + 16 * j, q1 - 1)] * R;
        T v[4][RQ];
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int u = 0; u < RQ; u++) v[j][u] = (q + 4 * u < R) ? contrib[sl[j] + q + 4 * u] : Sc<T>::zero();
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (p0 + 16 * j < q1) {
#pragma unroll
                for (int u = 0; u < RQ; u++)
                    if (q + 4 * u < R) acc[u] = Sc<T>::add(acc[u], v[j][u]);
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
template <class T>
__global__ __launch_bounds__(TB) void k_lfwd_gather(int32_t nrow, const int32_t* __restrict__ rows, const int64_t* __restrict__ incp,
                                                     const int64_t* __restrict__ incs, const double* __restrict__ contribd,
                                                     double* __restrict__ xd, int32_t R)
{
    const int32_t it = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    if (it >= nrow) return;
    gather_row<T>(rows[it], incp, incs, reinterpret_cast<const T*>(contribd), reinterpret_cast<T*>(xd), R);
}

// The blocks' diagonal triangles in the sweeps, as products with the inverses kept from the factorization:
// out(block rows) = Op in(block rows), Op = L11^-1 (lower, unit diagonal) or U11^-1 (upper) for the sweeps with A, and
// their transposes (conjugated for the sweeps with A^H: conj) for those with A^T; the inverses are w x w column-major
// (ld = w), so Op(row, c) = M[row + c w], or M[c + row w] transposed.  UPPER: Op is upper triangular.  One workgroup
// per 64 rows of a block: lane = row; the four waves take a quarter each of the columns the rows reach (lower: up to
// the chunk's last row, upper: from its first row), with the block's input for those columns and all right-hand sides
// in LDS; the quarters' sums are added in wave order -- a fixed order that does not depend on the number of right-hand
// sides.  The output goes to another vector: the block's input is still read by its other row chunks.
template <class T, bool UPPER, bool TRANS, int RC = RM<T>::v>
__device__ __forceinline__ void tri_prod_tile(const int2 tl, const BlockDesc* __restrict__ blocks, const double* __restrict__ ind,
                                     double* __restrict__ outd, int32_t R, int conj, T* xs)
{
    const BlockDesc B = blocks[tl.x];   // tl: block, first row
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
        // the column loads in batches of PF, every one of a batch issued before any is used (latency, not bandwidth,
        // bounds a tile): clamped into the range, the entries past it not applied
        constexpr int PF = RC == 1 ? 32 : 16;
        for (int c0 = lo; c0 < hi; c0 += PF) {
            T v[PF];
#pragma unroll
            for (int j2 = 0; j2 < PF; j2++) {
                const int c = min(c0 + j2, hi - 1);
                v[j2] = TRANS ? cj(gload(M, c + (int64_t)row * w), conj) : gload(M, row + (int64_t)c * w);
            }
#pragma unroll
            for (int j2 = 0; j2 < PF; j2++)
                if (c0 + j2 < hi) {
#pragma unroll
                    for (int r = 0; r < RC; r++)
                        if (r < R) acc[r] = Sc<T>::fma_(acc[r], v[j2], xs[(c0 + j2 - ca) * RC + r]);
                }
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
template <class T, bool UPPER, bool TRANS, int RC = RM<T>::v>
__global__ __launch_bounds__(TB) void k_tri_prod(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                  const double* __restrict__ ind, double* __restrict__ outd, int32_t R,
                                                  int conj)
{
    __shared__ T xs[TB * RC];
    tri_prod_tile<T, UPPER, TRANS, RC>(tiles[blockIdx.x], blocks, ind, outd, R, conj, xs);
}

// L y = b, part 2: the rows below a block, 64 rows per workgroup (lane = row), the four waves a quarter of the block's
// columns each (the block's y for all columns and right-hand sides in LDS), the quarters' sums added in wave order;
// slot_off[b] is the block's first slot
template <class T, int RC = RM<T>::v>
__device__ __forceinline__ void push_tile(const int2 tl, const BlockDesc* __restrict__ blocks, const int64_t* __restrict__ slot_off,
                                 const double* __restrict__ yd, double* __restrict__ contribd, int32_t R, T* ys)
{
    const BlockDesc B = blocks[tl.x];   // tl: block, first row below the block (0-based)
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
        // loads in batches, as in tri_prod_tile
        constexpr int PF = RC == 1 ? 32 : 16;
        for (int c0 = qa; c0 < qb; c0 += PF) {
            T v[PF];
#pragma unroll
            for (int j2 = 0; j2 < PF; j2++) v[j2] = gload(V, (int64_t)min(c0 + j2, qb - 1) * B.nrows + row);
#pragma unroll
            for (int j2 = 0; j2 < PF; j2++)
                if (c0 + j2 < qb) {
#pragma unroll
                    for (int r = 0; r < RC; r++)
                        if (r < R) acc[r] = Sc<T>::fma_(acc[r], v[j2], ys[(c0 + j2) * RC + r]);
                }
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
template <class T, int RC = RM<T>::v>
__global__ __launch_bounds__(TB) void k_lfwd_push(const int2* __restrict__ tiles, const BlockDesc* __restrict__ blocks,
                                                   const int64_t* __restrict__ slot_off, const double* __restrict__ yd,
                                                   double* __restrict__ contribd, int32_t R)
{
    __shared__ T ys[TB * RC];
    push_tile<T, RC>(tiles[blockIdx.x], blocks, slot_off, yd, contribd, R, ys);
}

// U x = y, part 1: the products of a block's pivot rows with their U rows beyond the block, one share of the block's
// pieces per workgroup.  A piece is a later column j's run over the block's rows [a, w), the pieces in column order: thread t
// owns block row t and adds U(t, j) x_j for every piece with a <= t, so consecutive rows read consecutive U entries; the
// pieces go in batches of UB_PC, whose x_j for all right-hand sides sit in LDS.  A share's sums go to its own slot of
// `part`; ub_red subtracts a block's slots from x in slot order (a fixed order, the same for any number of right-hand
// sides).
// (the sums of pieces [p0, p0 + np) of block b go to dst[t * stride + r] for block row t)
template <class T>
__device__ __forceinline__ void dots_tile(int32_t b, int32_t p0t, int32_t np, const BlockDesc* __restrict__ blocks,
                                 const int32_t* __restrict__ pa, const int32_t* __restrict__ pj, const int64_t* __restrict__ pu,
                                 const double* __restrict__ Uxd, const double* __restrict__ xd, T* __restrict__ dst,
                                 int64_t stride, int32_t R, T* sx, int32_t* sa, int64_t* su)
{
    constexpr int RC = RM<T>::v, PC = k::UB_PC;
    const int4 tl = make_int4(b, p0t, np, 0);
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
            // the batch's U entries loaded before any is used (latency bounds the tile), the products in piece order
            T u[PC];
#pragma unroll
            for (int q = 0; q < PC; q++) u[q] = (q < npc && sa[q] <= t) ? Ux[su[q] + t] : Sc<T>::zero();
#pragma unroll
            for (int q = 0; q < PC; q++)
                if (q < npc) {
#pragma unroll
                    for (int r = 0; r < RC; r++)
                        if (r < R) acc[r] = Sc<T>::fma_(acc[r], u[q], sx[q * RC + r]);
                }
        }
    }
    if (t >= w) return;
    T* part = dst + (int64_t)t * stride;
#pragma unroll
    for (int r = 0; r < RC; r++)
        if (r < R) part[r] = acc[r];
}
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
    dots_tile<T>(tl.x, tl.y, tl.z, blocks, pa, pj, pu, Uxd, xd, reinterpret_cast<T*>(partd) + (int64_t)tl.w * TB * RC, RC, R, sx,
                 sa, su);
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
    // two chains, joined at the end: the order is fixed by the slot count alone; the slots' loads in batches of 16,
    // each batch issued before its sums
    constexpr int SB = 16;
    for (int c0 = 0; c0 < tl.w; c0 += SB) {
        T v[SB];
#pragma unroll
        for (int j = 0; j < SB; j++) v[j] = part[(int64_t)min(c0 + j, tl.w - 1) * TB * RC];
#pragma unroll
        for (int j = 0; j < SB; j += 2) {
            if (c0 + j < tl.w) s0 = Sc<T>::add(s0, v[j]);
            if (c0 + j + 1 < tl.w) s1 = Sc<T>::add(s1, v[j + 1]);
        }
    }
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

// ---- the sweeps with A as one persistent launch each (dataflow)
//
// Every level's work becomes items in one list, in level order: L y = b has per block gather tiles G (64 of its pivot
// rows subtract their slots), triangle tiles T (64 rows: y = L11^-1 x) and push tiles P (64 rows below); U x = y has per
// block dot-product groups D (a share of its pieces into the group's own partial rows), one reduction Rd (y on its rows
// minus the groups' sums, in group order) and triangle tiles X (64 rows: x = U11^-1 y).  A workgroup takes the next item
// by an atomic ticket and first waits until what the item reads is complete: per block counters of finished items,
// raised after a release fence.  G waits for the push tiles of every block writing its pivot rows, T for its block's G, P for every
// T of its block; D waits for the X tiles of every block its pieces' columns belong to, Rd for the block's D groups, X
// for Rd.  An item waits only for items earlier in the list, and every item earlier in the list was taken by a
// workgroup already running, so the waits end whatever the grid's size and residency.  A wait gives up after
// DF_SPIN_LIMIT polls, raises the abort flag and the workgroup returns; the other waits see the flag and return too.
// Counters (int32): [0] the L ticket, [1] the U ticket, [2] the abort flag, [3] spare, then per block G, T, P, D, Rd, X.
// Every sum runs in a fixed order (a row's slots in slot order, a block's groups in group order): the same bits on every
// run and for any number of right-hand sides.
constexpr long DF_SPIN_LIMIT = 1L << 22;
constexpr int DF_TYPE_SHIFT = 29;

__device__ inline bool df_wait(const int32_t* c, int32_t target, int32_t* abortf)
{
    long spins = 0;
    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
        if ((++spins & 63) == 0 && __hip_atomic_load(abortf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        if (spins > DF_SPIN_LIMIT) {
            __hip_atomic_store(abortf, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return true;
}
__device__ inline int32_t df_tiles64(int32_t n) { return (n + 63) / 64; }

// takes the next item and waits for its inputs (thread 0), then publishes it to the workgroup; -1: no more work
template <class WAIT>
__device__ inline int32_t df_next(int32_t* ticket, int32_t nitems, int32_t* abortf, const WAIT& wait_for, int32_t* sh)
{
    __syncthreads();   // the previous item is done with the LDS and sh
    if (threadIdx.x == 0) {
        const int32_t it = __hip_atomic_fetch_add(ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int32_t out = -1;
        if (it < nitems && wait_for(it)) {
            __threadfence();   // acquire: what the finished items wrote
            out = it;
        }
        sh[0] = out;
    }
    __syncthreads();
    return sh[0];
}
// the item is done: its writes released, then the block's counter raised
__device__ inline void df_done(int32_t* c)
{
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        __hip_atomic_fetch_add(c, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <class T>
__global__ __launch_bounds__(TB) void k_lsweep_df(int32_t nitems, const int4* __restrict__ items, const BlockDesc* __restrict__ blocks,
                                                   const int32_t* __restrict__ dep_off, const int32_t* __restrict__ dep,
                                                   const int64_t* __restrict__ incp, const int64_t* __restrict__ incs,
                                                   const int64_t* __restrict__ slot_off, double* __restrict__ wd,
                                                   double* __restrict__ yd, double* __restrict__ contribd, int32_t* cnt,
                                                   int32_t nb, int32_t R)
{
    __shared__ T lds[TB * RM<T>::v];
    __shared__ int32_t sh[1];
    int32_t* abortf = cnt + 2;
    int32_t *cG = cnt + 4, *cT = cG + nb, *cP = cT + nb;
    auto wait_for = [&](int32_t it) -> bool {
        const int4 x = items[it];
        const int type = x.x >> DF_TYPE_SHIFT, b = x.x & ((1 << DF_TYPE_SHIFT) - 1);
        if (type == 0) {
            for (int32_t k2 = dep_off[b]; k2 < dep_off[b + 1]; k2++) {
                const int32_t bp = dep[k2];
                if (!df_wait(cP + bp, df_tiles64(blocks[bp].nrows - blocks[bp].w), abortf)) return false;
            }
            return true;
        }
        if (type == 1) return df_wait(cG + b, df_tiles64(blocks[b].w), abortf);
        return df_wait(cT + b, df_tiles64(blocks[b].w), abortf);
    };
    for (;;) {
        const int32_t it = df_next(cnt, nitems, abortf, wait_for, sh);
        if (it < 0) return;
        const int4 x = items[it];
        const int type = x.x >> DF_TYPE_SHIFT, b = x.x & ((1 << DF_TYPE_SHIFT) - 1);
        if (type == 0) {
            // G: rows [x.y, x.y + 64) of the block, a thread per (row, right-hand sides t % 4, t % 4 + 4, ..): the row's
            // slots summed in slot order
            const BlockDesc B = blocks[b];
            const int row = x.y + (int)(threadIdx.x >> 2), q = threadIdx.x & 3;
            if (row < B.w && q < R) {
                const T* contrib = reinterpret_cast<const T*>(contribd);
                T* wv = reinterpret_cast<T*>(wd);
                const int64_t i = (int64_t)B.c0 + row, q0 = incp[i], q1 = incp[i + 1];
                constexpr int RQ = RM<T>::v / 4;
                T acc[RQ];
#pragma unroll
                for (int u = 0; u < RQ; u++) acc[u] = Sc<T>::zero();
                for (int64_t p = q0; p < q1; p++) {
                    const int64_t sl = incs[p] * R;
#pragma unroll
                    for (int u = 0; u < RQ; u++)
                        if (q + 4 * u < R) acc[u] = Sc<T>::add(acc[u], contrib[sl + q + 4 * u]);
                }
#pragma unroll
                for (int u = 0; u < RQ; u++)
                    if (q + 4 * u < R) wv[i * R + q + 4 * u] = Sc<T>::sub(wv[i * R + q + 4 * u], acc[u]);
            }
            df_done(cG + b);
        } else if (type == 1) {
            tri_prod_tile<T, false, false>(make_int2(b, x.y), blocks, wd, yd, R, 0, lds);
            df_done(cT + b);
        } else {
            push_tile<T>(make_int2(b, x.y), blocks, slot_off, yd, contribd, R, lds);
            df_done(cP + b);
        }
    }
}

// part: per block its groups' partial sums, group g's row t right-hand side r at (ubase[b] + g w + t) R + r
template <class T>
__global__ __launch_bounds__(TB) void k_usweep_df(int32_t nitems, const int4* __restrict__ items, const BlockDesc* __restrict__ blocks,
                                                   const int32_t* __restrict__ dep_off, const int32_t* __restrict__ dep,
                                                   const int32_t* __restrict__ pa, const int32_t* __restrict__ pj,
                                                   const int64_t* __restrict__ pu, const double* __restrict__ Uxd,
                                                   const int64_t* __restrict__ ubase, double* __restrict__ xd,
                                                   double* __restrict__ yd, double* __restrict__ partd, int32_t* cnt,
                                                   int32_t nb, int32_t R)
{
    constexpr int RC = RM<T>::v, PC = k::UB_PC;
    __shared__ T lds[TB * RC];
    __shared__ int32_t sa[PC];
    __shared__ int64_t su[PC];
    __shared__ int32_t sh[1];
    int32_t* abortf = cnt + 2;
    int32_t *cD = cnt + 4 + 3 * nb, *cR = cD + nb, *cX = cR + nb;
    T* part = reinterpret_cast<T*>(partd);
    auto wait_for = [&](int32_t it) -> bool {
        const int4 x = items[it];
        const int type = x.x >> DF_TYPE_SHIFT, b = x.x & ((1 << DF_TYPE_SHIFT) - 1);
        if (type == 0) {
            for (int32_t k2 = dep_off[b]; k2 < dep_off[b + 1]; k2++) {
                const int32_t bj = dep[k2];
                if (!df_wait(cX + bj, df_tiles64(blocks[bj].w), abortf)) return false;
            }
            return true;
        }
        if (type == 1) return df_wait(cD + b, x.y, abortf);
        return df_wait(cR + b, x.z, abortf);
    };
    for (;;) {
        const int32_t it = df_next(cnt + 1, nitems, abortf, wait_for, sh);
        if (it < 0) return;
        const int4 x = items[it];
        const int type = x.x >> DF_TYPE_SHIFT, b = x.x & ((1 << DF_TYPE_SHIFT) - 1);
        if (type == 0) {
            // D: pieces [x.y, x.y + x.z) into group x.w
            const int32_t w = blocks[b].w;
            dots_tile<T>(b, x.y, x.z, blocks, pa, pj, pu, Uxd, xd, part + (ubase[b] + (int64_t)x.w * w) * R, R, R, lds, sa, su);
            df_done(cD + b);
        } else if (type == 1) {
            // Rd: y on the block's rows minus its x.y groups, in group order
            const BlockDesc B = blocks[b];
            T* y = reinterpret_cast<T*>(yd);
            const T* pb = part + ubase[b] * R;
            for (int idx = threadIdx.x; idx < B.w * R; idx += TB) {
                const int t = idx / R, r = idx - t * R;
                T acc = Sc<T>::zero();
                for (int g = 0; g < x.y; g++) acc = Sc<T>::add(acc, pb[((int64_t)g * B.w + t) * R + r]);
                T* yp = y + ((int64_t)B.c0 + t) * R + r;
                *yp = Sc<T>::sub(*yp, acc);
            }
            df_done(cR + b);
        } else {
            tri_prod_tile<T, true, false>(make_int2(b, x.y), blocks, yd, xd, R, 0, lds);
            df_done(cX + b);
        }
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
    if (R == 1) {   // one right-hand side: one accumulator per thread
        if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, false, false, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
        else hipLaunchKernelGGL((k_tri_prod<Zd, false, false, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
    } else if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, false, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
    else hipLaunchKernelGGL((k_tri_prod<Zd, false, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, x, y, R, 0);
}
void lfwd_push(int vs, int32_t ntile, const int2* tiles, const BlockDesc* blocks, const int64_t* slot_off, const double* y,
               double* contrib, int32_t R, double bytes, hipStream_t s)
{
    if (ntile <= 0) return;
    ProfScope ps("lfwd_push", bytes, s);
    if (R == 1) {
        if (vs == 1) hipLaunchKernelGGL((k_lfwd_push<double, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, slot_off, y, contrib, R);
        else hipLaunchKernelGGL((k_lfwd_push<Zd, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, slot_off, y, contrib, R);
    } else if (vs == 1) hipLaunchKernelGGL(k_lfwd_push<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks, slot_off, y, contrib, R);
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
    if (R == 1) {
        if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, true, false, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
        else hipLaunchKernelGGL((k_tri_prod<Zd, true, false, 1>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
    } else if (vs == 1) hipLaunchKernelGGL((k_tri_prod<double, true, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
    else hipLaunchKernelGGL((k_tri_prod<Zd, true, false>), dim3(ntile), dim3(TB), 0, s, tiles, blocks, y, x, R, 0);
}
int df_grid(int vs, int upper)
{
    // the workgroups the device holds at once (the persistent sweeps' grid), per kernel
    static std::atomic<int> cached[4] = {-1, -1, -1, -1};
    const int i = (vs == 1 ? 0 : 2) + (upper ? 1 : 0);
    int g = cached[i].load(std::memory_order_relaxed);
    if (g < 0) {
        int per = 0;
        const void* f = vs == 1 ? (upper ? (const void*)k_usweep_df<double> : (const void*)k_lsweep_df<double>)
                                : (upper ? (const void*)k_usweep_df<Zd> : (const void*)k_lsweep_df<Zd>);
        hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, f, TB, 0), "occupancy of the persistent sweeps");
        g = std::max(1, per) * std::max(1, devctx().ncu);
        cached[i].store(g, std::memory_order_relaxed);
    }
    return g;
}
void lsweep_df(int vs, int32_t nitems, const int4* items, const BlockDesc* blocks, const int32_t* dep_off, const int32_t* dep,
               const int64_t* incp, const int64_t* incs, const int64_t* slot_off, double* w, double* y, double* contrib,
               int32_t* cnt, int32_t nb, int32_t R, double bytes, hipStream_t s)
{
    if (nitems <= 0) return;
    ProfScope ps("lsweep_df", bytes, s);
    const unsigned g = (unsigned)std::min(nitems, df_grid(vs, 0));
    if (vs == 1)
        hipLaunchKernelGGL(k_lsweep_df<double>, dim3(g), dim3(TB), 0, s, nitems, items, blocks, dep_off, dep, incp, incs, slot_off, w,
                           y, contrib, cnt, nb, R);
    else
        hipLaunchKernelGGL(k_lsweep_df<Zd>, dim3(g), dim3(TB), 0, s, nitems, items, blocks, dep_off, dep, incp, incs, slot_off, w, y,
                           contrib, cnt, nb, R);
}
void usweep_df(int vs, int32_t nitems, const int4* items, const BlockDesc* blocks, const int32_t* dep_off, const int32_t* dep,
               const int32_t* pa, const int32_t* pj, const int64_t* pu, const double* Ux, const int64_t* ubase, double* x,
               double* y, double* part, int32_t* cnt, int32_t nb, int32_t R, double bytes, hipStream_t s)
{
    if (nitems <= 0) return;
    ProfScope ps("usweep_df", bytes, s);
    const unsigned g = (unsigned)std::min(nitems, df_grid(vs, 1));
    if (vs == 1)
        hipLaunchKernelGGL(k_usweep_df<double>, dim3(g), dim3(TB), 0, s, nitems, items, blocks, dep_off, dep, pa, pj, pu, Ux, ubase, x,
                           y, part, cnt, nb, R);
    else
        hipLaunchKernelGGL(k_usweep_df<Zd>, dim3(g), dim3(TB), 0, s, nitems, items, blocks, dep_off, dep, pa, pj, pu, Ux, ubase, x, y,
                           part, cnt, nb, R);
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
