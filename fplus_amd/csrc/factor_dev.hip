// factor_dev.hip -- the numeric kernels of the supernodal factorization (gfx950).
//
// A panel of P columns lives in a compact workspace W (column-major, ld = ldw): its rows are the rows the panel's
// columns touch -- first the rows already pivotal (the U segments), then the rows not yet pivotal -- found by the host
// search; rowmap[row] gives a row's position.  The kernels here
//   * scatter the panel's columns of A into W (scatter_cols);
//   * apply the earlier supernodes to the panel, block by block: the unit-lower solve of the block's pivot rows the
//     panel's columns reach (the TRSM) and the rectangle below subtracted through rowmap (the GEMM, with the
//     relative-index scatter) as one product on the FP64 matrix cores with the block's inverse form (blk_upd) -- one
//     launch per dependency level, every block of a level in it -- then the solved segments back into W (x_back);
//   * make each new block's inverse form (blk_inverse: its triangle's inverse, then the rows below times it);
//   * factor the rows not yet pivotal by dense partial pivoting (getf2 per sub-panel of nb columns, then the
//     trailing TRSM and GEMM), LAPACK-style with whole-row swaps, so the pivot rows end at the top in column order.
// Entries of W outside a column's sparse pattern are exact zeros and stay so, which makes the dense panel LU produce
// the sparse factorization's values and pivots (a row outside the pattern has modulus 0 and cannot win).
#include <atomic>

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

constexpr int RT = FPA_ROW_TILE; // rows per GEMM tile
// panel columns per tile of the panel-LU GEMM: 64 real, 32 complex (LDS within 64 KB)
template <class T> struct CTile { static constexpr int v = 64; };
template <> struct CTile<Zd> { static constexpr int v = 32; };
constexpr int PNB = FPA_PANEL_NB;  // sub-panel width of the dense panel LU
__device__ inline double shfl_t(double v, int src) { return __shfl(v, src); }
__device__ inline Zd shfl_t(Zd v, int src) { return zmake(__shfl(v.x, src), __shfl(v.y, src)); }

template <class T> __device__ inline T shfl_val(T v, int src);
template <> __device__ inline double shfl_val<double>(double v, int src) { return __shfl(v, src); }
template <> __device__ inline Zd shfl_val<Zd>(Zd v, int src) { return zmake(__shfl(v.x, src), __shfl(v.y, src)); }

// The updates use each block in an inverse form, computed once when the block is made:
//     IV = [ L11^-1 ; L21 L11^-1 ]   (nrows x w, column-major like the block's values),
// L11 the block's unit lower triangle (its pivot rows) and L21 the rows below.  The inverse of a trailing part
// L11[k0:w, k0:w] is the trailing part of L11^-1, and (L21 L11^-1)[:, k0:w] = L21[:, k0:w] L11^-1[k0:w, k0:w], so a task
// whose columns reach the segment [k0, w) is one product IV[k0:nrows, k0:w] S with S the segment's rows of W: its top
// m = w - k0 rows are the solved segment X (the TRSM), the rest the update of the rows below (the GEMM).

// column c of L11^-1 (one workgroup per (block, column); thread t owns row t, a block being at most 256 columns): the
// unit-lower solve of e_c in chunks of 16 rows -- a chunk's triangle inside its wavefront (the row just finished
// broadcast by a lane shuffle), one barrier, then the rows below subtract the chunk's 16 columns -- starting at the
// chunk that holds row c (the rows above it stay zero); the next chunk's factor values are read during the current one
template <class T>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_blk_linv(const int2* __restrict__ jobs, const BlockDesc* __restrict__ blocks)
{
    constexpr int CH = 16;
    __shared__ T xs[TB];
    const int2 jb = jobs[blockIdx.x];
    const BlockDesc B = blocks[jb.x];
    const int32_t w = B.w, c = jb.y, t = threadIdx.x;
    const gptr<const T> L = as_global(reinterpret_cast<const T*>(B.vals));   // global: generic loads would hold up the LDS waits
    const bool own = t < w;
    T x = t == c ? Sc<T>::one() : Sc<T>::zero();
    auto fetch = [&](int ch, T* lv) {
        const int r0 = ch * CH;
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int k = r0 + u;
            lv[u] = (own && k < t) ? gload(L, (int64_t)k * B.nrows + t) : Sc<T>::zero();
        }
    };
    T lcur[CH], lnxt[CH];
    const int ch0 = c / CH, nch = (w + CH - 1) / CH;
    fetch(ch0, lcur);
    for (int ch = ch0; ch < nch; ch++) {
        const int r0 = ch * CH, r1 = min(w, r0 + CH);
        if (ch + 1 < nch) fetch(ch + 1, lnxt);
        if (t >= r0 && t < r1) {
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const int k = r0 + u;
                if (k >= r1) break;
                const T xk = shfl_val<T>(x, k & 63);
                if (t > k) x = Sc<T>::fms(x, lcur[u], xk);
            }
            xs[t] = x;
        }
        __syncthreads();
        if (own && t >= r1) {
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const int k = r0 + u;
                if (k >= r1) break;
                x = Sc<T>::fms(x, lcur[u], xs[k]);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) lcur[u] = lnxt[u];
    }
    if (own) {
        gstore(as_global(reinterpret_cast<T*>(B.ivals)), t + (int64_t)c * B.nrows, x);
        gstore(as_global(reinterpret_cast<T*>(B.linv)), t + (int64_t)c * w, x);
    }
}

// column c of U11^-1 (one workgroup per (block, column); thread t owns row t): the upper solve of e_c, rows from c
// down, in chunks of 16 rows -- a chunk's triangle inside its wavefront (each row divided by its diagonal, then
// broadcast by a lane shuffle), one barrier, then the rows above subtract the chunk's 16 columns; rows below c stay
// zero.  The next chunk's factor values are read during the current one.
template <class T>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_blk_uinv(const int2* __restrict__ jobs, const BlockDesc* __restrict__ blocks)
{
    constexpr int CH = 16;
    __shared__ T xs[TB];
    const int2 jb = jobs[blockIdx.x];
    const BlockDesc B = blocks[jb.x];
    const int32_t w = B.w, c = jb.y, t = threadIdx.x;
    const gptr<const T> U = as_global(reinterpret_cast<const T*>(B.vals));   // global, as in k_blk_linv
    const bool own = t <= c;
    T x = t == c ? Sc<T>::one() : Sc<T>::zero();
    const T d = own ? gload(U, (int64_t)t * B.nrows + t) : Sc<T>::one();
    // U(t, k) for the chunk's columns k, t < k <= c
    auto fetch = [&](int ch, T* lv) {
        const int r0 = ch * CH;
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int k = r0 + u;
            lv[u] = (own && k > t && k <= c) ? gload(U, (int64_t)k * B.nrows + t) : Sc<T>::zero();
        }
    };
    T lcur[CH], lnxt[CH];
    const int ch0 = c / CH;
    fetch(ch0, lcur);
    for (int ch = ch0; ch >= 0; ch--) {
        const int r0 = ch * CH, r1 = min(c + 1, r0 + CH);
        if (ch > 0) fetch(ch - 1, lnxt);
        if (t >= r0 && t < r1) {
#pragma unroll
            for (int u = CH - 1; u >= 0; u--) {
                const int k = r0 + u;
                if (k >= r1) continue;
                if (t == k) x = Sc<T>::div(x, d);
                const T xk = shfl_val<T>(x, k & 63);
                if (t < k) x = Sc<T>::fms(x, lcur[u], xk);
            }
            xs[t] = x;
        }
        __syncthreads();
        if (t < r0) {
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const int k = r0 + u;
                if (k >= r1) break;
                x = Sc<T>::fms(x, lcur[u], xs[k]);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) lcur[u] = lnxt[u];
    }
    if (t < w) gstore(as_global(reinterpret_cast<T*>(B.uinv)), t + (int64_t)c * w, x);
}

// Matrix cores: D(16x16) += A(16x4) B(4x16) per wavefront; lane l supplies A[l % 16][l / 16] and B[l / 16][l % 16].
// FP64 (v_mfma_f64_16x16x4f64): lane l holds D[l / 16 + 4 v][l % 16], v < 4; FP32 (v_mfma_f32_16x16x4f32): lane l holds
// D[4 (l / 16) + v][l % 16] (both checked on gfx950)
typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));
__device__ inline dx4 mfma16(double a, double b, dx4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline fx4 mfma16(float a, float b, fx4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

struct Zf {
    float x, y;
};

// the arithmetic of a product tile (PR): 0 FP64 operands and accumulation; 1 FP32 (the operands rounded to FP32 when
// they are staged in LDS, FP32 accumulation, the result widened to FP64 for the update of W); 2 BF16 hi/lo split products
// (staged in FP32 like 1; each FP32 operand x split into hi = bf16(x) and lo = bf16(x - hi), the product taken as
// Ahi Bhi + Ahi Blo + Alo Bhi on the BF16 matrix cores with FP32 accumulation: about 16 bits of each operand)
constexpr int PR64 = 0, PR32 = 1, PRBF = 2;
template <class T, int PR> struct Ar {
    using S = T;
    using acc = dx4;
    __device__ static S cvt(T v) { return v; }
    __device__ static S zero() { return Sc<T>::zero(); }
    // the column of its 16 x 16 part that lane `lane`'s accumulator element v holds (its row is lane % 16)
    __device__ static int col(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Ar<double, PR32> {
    using S = float;
    using acc = fx4;
    __device__ static S cvt(double v) { return (float)v; }
    __device__ static S zero() { return 0.0f; }
    __device__ static int col(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct Ar<Zd, PR32> {
    using S = Zf;
    using acc = fx4;
    __device__ static S cvt(Zd v) { Zf r; r.x = (float)v.x; r.y = (float)v.y; return r; }
    __device__ static S zero() { Zf r; r.x = 0.0f; r.y = 0.0f; return r; }
    __device__ static int col(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <> struct Ar<double, PRBF> : Ar<double, PR32> {};
template <> struct Ar<Zd, PRBF> : Ar<Zd, PR32> {};

// BF16 round to nearest even (finite inputs), and the three-product split multiply on 16 x 16 x 16 BF16 matrix cores:
// lane l supplies A[l % 16][4 (l / 16) + t] and B[4 (l / 16) + t][l % 16], t < 4; the result lands as in the FP32 form
typedef short s4v __attribute__((ext_vector_type(4)));
__device__ inline short f2bf(float x)
{
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ inline float bf2f(short h) { return __uint_as_float(((uint32_t)(uint16_t)h) << 16); }
__device__ inline fx4 mfma_bf3(const float (&a)[4], const float (&b)[4], fx4 c)
{
    s4v ah, al, bh, bl;
#pragma unroll
    for (int t = 0; t < 4; t++) {
        ah[t] = f2bf(a[t]);
        al[t] = f2bf(a[t] - bf2f(ah[t]));
        bh[t] = f2bf(b[t]);
        bl[t] = f2bf(b[t] - bf2f(bh[t]));
    }
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, c, 0, 0, 0);
    return c;
}

constexpr int TN = 64;           // columns per product tile
constexpr int BLD = TN + 1;      // LDS row length of B: one element of padding moves each k row to other banks
template <class T, int TM> struct TileCfg {
    // loads two chunks ahead, except for complex 64-row tiles: their accumulators leave no room for a second register set
    static constexpr bool DEEP = !(Sc<T>::vs == 2 && TM == 64);
    // complex products with three real ones (Gauss): T1 = sum br (ar + ai), T2 = sum ar (bi - br), T3 = sum ai (br + bi),
    // re = T1 - T3, im = T1 + T2 -- a fourth fewer matrix-core cycles; the error stays within the same normwise bound.
    // Not for complex 64-row tiles: a third accumulator set does not fit their registers.
    static constexpr bool G3 = Sc<T>::vs == 2 && TM == 32;
    static constexpr int KC = 8;   // k per staged chunk (one barrier per chunk; 16 was measured slower)
    static constexpr int LPA = KC * TM / TB, LPB = KC * TN / TB;   // elements of A and of B each thread stages
    static constexpr int WR = TM / 32, WC = 4 / WR, SJ = TN / WC / 16;   // waves: WR x WC, each 32 x (16 SJ)
    static constexpr int STG = 2 * KC * TM + 2 * KC * BLD;   // LDS scalars: two chunks of A[k][row] and B[k][col]
};

// C (TM x 64) = A (TM x [kb, ke)) B ([kb, ke) x 64) on the matrix cores (four real products per complex one, or three):
// k in chunks of 8 staged in LDS (two buffers), with the global loads two chunks ahead of the products where registers
// allow (TileCfg::DEEP): chunk i + 2 is read into registers while chunk i multiplies and chunk i + 1 waits in the other
// register set, so a load has two chunks' products to land; else one chunk ahead.  la(r, k) and lb(k, c) give the
// operands (zero outside the tile) in FP64; PR 1 and 2 round them to FP32 when they are staged.  The products are formed
// transposed (B^T A^T), so each lane ends with rows of C along lanes: out(r, c, v) is called for every element, and a
// column's 16 consecutive rows come from 16 lanes -- the write-out needs no pass through LDS.  pre() runs once the
// first chunk is staged, before the barrier that publishes it: LDS state the output needs is stored there, from loads
// that were in flight with the first chunks'.  stg: LDS of TileCfg<T, TM>::STG scalars T.
template <int N> struct IC { static constexpr int value = N; };

template <class T, int TM, int PR, class LA, class LB, class OUT, class PRE>
__device__ inline void tile_mfma(int32_t kb, int32_t ke, int32_t mrows, int32_t ncols, const LA& la, const LB& lb, const OUT& out,
                                 const PRE& pre, T* stg)
{
    using Cf = TileCfg<T, TM>;
    using AR = Ar<T, PR>;
    using S = typename AR::S;
    using ACC = typename AR::acc;
    constexpr bool CPLX = Sc<T>::vs == 2;
    constexpr int KC = Cf::KC;
    S* const As = reinterpret_cast<S*>(stg);
    S* const Bs = As + 2 * KC * TM;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = (wv / Cf::WC) * 32, wc = (wv % Cf::WC) * (16 * Cf::SJ);   // the wave's part of the tile
    // a wave whose part of the tile holds no used row or column (of mrows x ncols) skips its products: they are zero
    // and never written, and its matrix-core cycles go to other waves (finer skips, per 16 rows or columns inside a
    // wave's part, cost the complex tiles their registers)
    const bool on = mrows > wr && ncols > wc;
    ACC accr[2][Cf::SJ], acci[2][Cf::SJ], acc3[2][Cf::SJ];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < Cf::SJ; j++) {
            accr[i][j] = ACC{0, 0, 0, 0};
            acci[i][j] = ACC{0, 0, 0, 0};
            acc3[i][j] = ACC{0, 0, 0, 0};
        }
    // thread element e = tid + q * TB; A: row e % TM, k e / TM (rows contiguous); B: k e % KC, column e / KC
    T ra0[Cf::LPA], rb0[Cf::LPB], ra1[Cf::LPA], rb1[Cf::LPB];
    // Every load of a chunk is issued, also past the end (k clamped to ke - 1 < ke, so the address stays inside the
    // operand), and the elements past ke are zeroed when the chunk is staged, where the data are waited for anyway:
    // loads under branches, or values used right after their load, make the compiler wait for every load in flight --
    // the prefetched chunks' too -- at each LDS store.  la and lb read rows past the tile's and columns past its width
    // from valid places too; the products of those rows and columns are not written out.
    auto fetch = [&](int32_t kc, T (&ra)[Cf::LPA], T (&rb)[Cf::LPB]) {
#pragma unroll
        for (int q = 0; q < Cf::LPA; q++) {
            const int e = tid + q * TB;
            ra[q] = la(e % TM, min(kc + e / TM, ke - 1));
        }
#pragma unroll
        for (int q = 0; q < Cf::LPB; q++) {
            const int e = tid + q * TB;
            rb[q] = lb(min(kc + e % KC, ke - 1), e / KC, q);
        }
    };
    auto stage = [&](int buf, int32_t kc, const T (&ra)[Cf::LPA], const T (&rb)[Cf::LPB]) {
#pragma unroll
        for (int q = 0; q < Cf::LPA; q++) {
            const int e = tid + q * TB;
            As[(buf * KC + e / TM) * TM + e % TM] = kc + e / TM < ke ? AR::cvt(ra[q]) : AR::zero();
        }
#pragma unroll
        for (int q = 0; q < Cf::LPB; q++) {
            const int e = tid + q * TB;
            Bs[(buf * KC + e % KC) * BLD + e / KC] = kc + e % KC < ke ? AR::cvt(rb[q]) : AR::zero();
        }
    };
    auto mma_n = [&](auto ni_c, auto nj_c, int buf) {
        constexpr int NI = decltype(ni_c)::value, NJ = decltype(nj_c)::value;
        if constexpr (PR == PRBF) {
            // the chunk's 8 k in one 16-deep BF16 product: lane groups 0 and 1 supply k 0-3 and 4-7, groups 2 and 3 zeros
            static_assert(KC == 8, "the BF16 form takes chunks of 8");
            const int kq = lane >> 4, kb4 = 4 * (kq & 1);
            const float on4 = kq < 2 ? 1.0f : 0.0f;
            S a[2][4], b[Cf::SJ][4];
#pragma unroll
            for (int t = 0; t < 4; t++) {
#pragma unroll
                for (int i = 0; i < NI; i++) a[i][t] = As[(buf * KC + kb4 + t) * TM + wr + i * 16 + (lane & 15)];
#pragma unroll
                for (int j = 0; j < NJ; j++) b[j][t] = Bs[(buf * KC + kb4 + t) * BLD + wc + j * 16 + (lane & 15)];
            }
#pragma unroll
            for (int i = 0; i < NI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    float p[4], q[4];
                    if constexpr (Cf::G3) {
#pragma unroll
                        for (int t = 0; t < 4; t++) { p[t] = on4 * b[j][t].x; q[t] = on4 * (a[i][t].x + a[i][t].y); }
                        accr[i][j] = mfma_bf3(p, q, accr[i][j]);
#pragma unroll
                        for (int t = 0; t < 4; t++) { p[t] = on4 * (b[j][t].y - b[j][t].x); q[t] = on4 * a[i][t].x; }
                        acci[i][j] = mfma_bf3(p, q, acci[i][j]);
#pragma unroll
                        for (int t = 0; t < 4; t++) { p[t] = on4 * (b[j][t].x + b[j][t].y); q[t] = on4 * a[i][t].y; }
                        acc3[i][j] = mfma_bf3(p, q, acc3[i][j]);
                    } else if constexpr (CPLX) {
                        float py[4], pn[4];
#pragma unroll
                        for (int t = 0; t < 4; t++) {
                            p[t] = on4 * b[j][t].x;
                            py[t] = on4 * b[j][t].y;
                            pn[t] = -py[t];
                            q[t] = on4 * a[i][t].x;
                        }
                        accr[i][j] = mfma_bf3(p, q, accr[i][j]);    // re += b.x a.x
                        acci[i][j] = mfma_bf3(py, q, acci[i][j]);   // im += b.y a.x
#pragma unroll
                        for (int t = 0; t < 4; t++) q[t] = on4 * a[i][t].y;
                        accr[i][j] = mfma_bf3(pn, q, accr[i][j]);   // re -= b.y a.y
                        acci[i][j] = mfma_bf3(p, q, acci[i][j]);    // im += b.x a.y
                    } else {
#pragma unroll
                        for (int t = 0; t < 4; t++) { p[t] = on4 * b[j][t]; q[t] = on4 * a[i][t]; }
                        accr[i][j] = mfma_bf3(p, q, accr[i][j]);
                    }
                }
            return;
        }
#pragma unroll
        for (int ks = 0; ks < KC; ks += 4) {
            const int kk = ks + (lane >> 4);
            S a[2], b[Cf::SJ];   // a: A[row lane % 16][k lane / 16]; b: B[k lane / 16][column lane % 16]
#pragma unroll
            for (int i = 0; i < NI; i++) a[i] = As[(buf * KC + kk) * TM + wr + i * 16 + (lane & 15)];
#pragma unroll
            for (int j = 0; j < NJ; j++) b[j] = Bs[(buf * KC + kk) * BLD + wc + j * 16 + (lane & 15)];
#pragma unroll
            for (int i = 0; i < NI; i++)
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    if constexpr (Cf::G3) {
                        accr[i][j] = mfma16(b[j].x, a[i].x + a[i].y, accr[i][j]);
                        acci[i][j] = mfma16(b[j].y - b[j].x, a[i].x, acci[i][j]);
                        acc3[i][j] = mfma16(b[j].x + b[j].y, a[i].y, acc3[i][j]);
                    } else if constexpr (CPLX) {
                        accr[i][j] = mfma16(b[j].x, a[i].x, accr[i][j]);
                        accr[i][j] = mfma16(-b[j].y, a[i].y, accr[i][j]);
                        acci[i][j] = mfma16(b[j].x, a[i].y, acci[i][j]);
                        acci[i][j] = mfma16(b[j].y, a[i].x, acci[i][j]);
                    } else {
                        accr[i][j] = mfma16(b[j], a[i], accr[i][j]);
                    }
                }
        }
    };
    auto mma = [&](int buf) {
        if (on) mma_n(IC<2>(), IC<Cf::SJ>(), buf);
    };
    fetch(kb, ra0, rb0);
    if constexpr (!Cf::DEEP) {
        stage(0, kb, ra0, rb0);
        pre();
        __syncthreads();
        int buf = 0;
        for (int32_t kc = kb; kc < ke; kc += KC) {
            const bool more = kc + KC < ke;
            fetch(kc + KC, ra0, rb0);
            mma(buf);
            if (more) stage(buf ^ 1, kc + KC, ra0, rb0);
            __syncthreads();
            buf ^= 1;
        }
    } else {
    fetch(kb + KC, ra1, rb1);
    stage(0, kb, ra0, rb0);
    pre();
    __syncthreads();
    // chunk i = kb + i KC; the loop runs while both of its halves stage a next chunk, and the last one or two chunks
    // are multiplied after it: a loop with no exit in its body lets the compiler keep the second register set's loads
    // in flight across the loop head (an exit in the body merges with the back edge and forces a full wait there).
    // The trip count depends on kb and ke only: every wave takes the same path.
    const int32_t nch = (ke - kb + KC - 1) / KC;
    int32_t kc = kb;
    for (int32_t i = 0; i + 2 < nch; i += 2, kc += 2 * KC) {
        // chunk kc in LDS buffer 0, chunk kc + KC in registers 1
        fetch(kc + 2 * KC, ra0, rb0);
        mma(0);
        stage(1, kc + KC, ra1, rb1);
        __syncthreads();
        // chunk kc + KC in LDS buffer 1, chunk kc + 2 KC in registers 0
        fetch(kc + 3 * KC, ra1, rb1);
        mma(1);
        stage(0, kc + 2 * KC, ra0, rb0);
        __syncthreads();
    }
    // one or two chunks left: kc in LDS buffer 0, kc + KC (if any) in registers 1
    mma(0);
    if (kc + KC < ke) {
        stage(1, kc + KC, ra1, rb1);
        __syncthreads();
        mma(1);
    }
    }
    // lane l holds C[row l % 16][column AR::col(l, v)] of each 16 x 16 part
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < Cf::SJ; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                T val;
                if constexpr (Cf::G3)
                    val = zmake((double)accr[i][j][v] - (double)acc3[i][j][v], (double)accr[i][j][v] + (double)acci[i][j][v]);
                else if constexpr (CPLX) val = zmake(accr[i][j][v], acci[i][j][v]);
                else val = accr[i][j][v];
                out(wr + i * 16 + (lane & 15), wc + j * 16 + AR::col(lane, v), val);
            }
}

// the rows below a new block's triangle in inverse form, one 64 x 64 tile: IV[w + r, c] = sum_k L21[r, k] L11^-1[k, c]
// for k from the tile's first column on (L11^-1 is lower triangular)
template <class T>
__global__ __launch_bounds__(TB) void k_blk_ivgemm(const GemmTile* __restrict__ tiles, const BlockDesc* __restrict__ blocks)
{
    constexpr int TM = 64;
    __shared__ T stg[TileCfg<T, TM>::STG];
    const GemmTile tl = tiles[blockIdx.x];
    const BlockDesc B = blocks[tl.task];
    const int32_t r0 = B.w + tl.row0;
    const int32_t nr = min(TM, B.nrows - r0), nc = min(TN, B.w - tl.col0);
    const gptr<const T> L = as_global(reinterpret_cast<const T*>(B.vals));
    const gptr<const T> IVr = as_global(reinterpret_cast<const T*>(B.ivals));
    // rows past nr and columns past nc read the tile's last ones (nr, nc >= 1): their products are not written out
    auto la = [&](int r, int32_t k) { return gload(L, (int64_t)k * B.nrows + r0 + min(r, nr - 1)); };
    auto lb = [&](int32_t k, int c, int) { return gload(IVr, k + (int64_t)(tl.col0 + min(c, nc - 1)) * B.nrows); };
    const gptr<T> out_base = as_global(reinterpret_cast<T*>(B.ivals) + r0 + (int64_t)tl.col0 * B.nrows);
    auto out = [&](int r, int c, T v) {
        if (r < nr && c < nc) gstore(out_base, r + (int64_t)c * B.nrows, v);
    };
    tile_mfma<T, TM, PR64>(tl.col0, B.w, nr, nc, la, lb, out, [] {}, stg);
}

// one tile of a task's update, rows [row0, row0 + TM) of the block from the segment start k0 on: the product
// IV[rows, k0:w] S with S the segment's rows of W (consecutive in W from spos0).  Rows of the segment give the solved
// segment X, stored compact in the task's slice (W's segment rows are still being read by the task's other tiles);
// rows below are subtracted from W through rowmap, the relative-index scatter -- with atomics where another task of
// the launch may write the same rows (tl.shared), else with plain loads and stores.

template <class T, int TM, int PR, bool LOG = false>
__device__ inline void upd_tile(const UpdTile& tl, const int32_t* __restrict__ cols, const int32_t* __restrict__ rowmap,
                                double* __restrict__ W, int64_t ldw, double* __restrict__ Xb, T* stg, int32_t* rpos, int32_t* cidx,
                                typename Ar<T, PR32>::S* p32 = nullptr, unsigned long long* blk_err = nullptr)
{
    constexpr bool CPLX = Sc<T>::vs == 2;
    const int32_t m = tl.w - tl.k0, nr = min(TM, tl.nrows - tl.row0), nc = tl.nc;
    const int tid = threadIdx.x;
    // the columns are needed for the first chunk, the rows' W positions only for the output: their rowmap loads are
    // issued here and land with the first chunk's
    int32_t row = -1;
    // (a global load: a generic one would wait with the descriptor's scalar loads and hold back the column load)
    if (tid < TM && tid < nr && tl.row0 + tid >= tl.w) row = as_global(tl.rows)[tl.row0 + tid];
    if (tid < TN) cidx[tid] = tid < nc ? cols[tl.cfirst + tid] : 0;
    __syncthreads();
    const int32_t rp = row >= 0 ? rowmap[row] : 0;
    const gptr<const T> IV = as_global(reinterpret_cast<const T*>(tl.iv) + tl.row0);
    const T* S = reinterpret_cast<const T*>(W) + (tl.spos0 - tl.k0);
    // rows past nr read the tile's last row (nr >= 1), columns past nc W's column 0 (cidx): their products are not
    // written out.  A thread stages the same B columns in every chunk (column (tid + q TB) / KC of its q-th element):
    // their W offsets are taken once
    using Cf = TileCfg<T, TM>;
    int64_t cofs[Cf::LPB];
#pragma unroll
    for (int q = 0; q < Cf::LPB; q++) cofs[q] = (int64_t)cidx[(tid + q * TB) / Cf::KC] * ldw;
    auto la = [&](int r, int32_t k) { return gload(IV, (int64_t)k * tl.nrows + min(r, nr - 1)); };
    auto lb = [&](int32_t k, int, int q) { return S[k + cofs[q]]; };
    const gptr<T> X = as_global(reinterpret_cast<T*>(Xb) + tl.xo + (tl.row0 - tl.k0));
    T* Wt = reinterpret_cast<T*>(W);
    const int32_t rw = tl.w - tl.row0;   // tile rows from here on are below the block's triangle
    const bool shared = (tl.flags & FPA_TILE_SHARED) != 0;
    auto out = [&](int r, int c, T v) {
        if (r >= nr || c >= nc) return;
        if (r < rw) {
            gstore(X, r + (int64_t)c * m, v);
            return;
        }
        T* p = Wt + rpos[r] + (int64_t)cidx[c] * ldw;
        if (shared) {
            if constexpr (CPLX) {
                atomicAdd(&p->x, -v.x);
                atomicAdd(&p->y, -v.y);
            } else {
                atomicAdd(p, -v);
            }
        } else {
            *p = Sc<T>::sub(*p, v);
        }
    };
    // IV's rows inside the triangle are rows of L11^-1, zero right of the diagonal: a tile whose rows all lie above w
    // needs k only up to its last row
    const int32_t ke = min(tl.w, tl.row0 + TM);
    if constexpr (!LOG) {
        tile_mfma<T, TM, PR>(tl.k0, ke, nr, nc, la, lb, out, [&] { if (tid < TM) rpos[tid] = rp; }, stg);
    } else {
        // the logging form, for training the precision selector: the FP32 product first (kept in LDS, not applied),
        // then the FP64 one (applied); the largest difference and the largest FP64 value go to the block's entries of
        // blk_err (non-negative doubles as bits, max-accumulated)
        using AR = Ar<T, PR32>;
        auto out32 = [&](int r, int c, T v) {
            if (r < nr && c < nc) p32[r * TN + c] = AR::cvt(v);
        };
        tile_mfma<T, TM, PR32>(tl.k0, ke, nr, nc, la, lb, out32, [&] { if (tid < TM) rpos[tid] = rp; }, stg);
        __syncthreads();
        double dmax = 0.0, vmax = 0.0;
        auto out64 = [&](int r, int c, T v) {
            if (r >= nr || c >= nc) return;
            const typename AR::S q = p32[r * TN + c];
            T qd;
            if constexpr (CPLX) qd = zmake(q.x, q.y);
            else qd = (double)q;
            dmax = fmax(dmax, Sc<T>::absval(Sc<T>::sub(v, qd)));
            vmax = fmax(vmax, Sc<T>::absval(v));
            out(r, c, v);
        };
        tile_mfma<T, TM, PR64>(tl.k0, ke, nr, nc, la, lb, out64, [] {}, stg);
        for (int off = 32; off > 0; off >>= 1) {
            dmax = fmax(dmax, __shfl_xor(dmax, off));
            vmax = fmax(vmax, __shfl_xor(vmax, off));
        }
        if ((tid & 63) == 0) {
            atomic_max_nonneg(blk_err + 2 * (int64_t)tl.block, dmax);
            atomic_max_nonneg(blk_err + 2 * (int64_t)tl.block + 1, vmax);
        }
    }
}

// one workgroup per tile of a dependency level (every task of the level in one launch).  MODE 0: FP64 products; 1: FP32
// products; 2: each tile takes the FP32 or the FP64 product by its block's mark (uniform over the workgroup); 3: FP64
// products applied, the FP32 ones computed beside them and their differences logged per block; 4: BF16 hi/lo split
// products
template <class T, int TM, int MODE>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(4))) void k_blk_upd(
    const UpdTile* __restrict__ tiles, const int32_t* __restrict__ cols, const int32_t* __restrict__ rowmap,
    double* __restrict__ W, int64_t ldw, double* __restrict__ X, const uint8_t* __restrict__ blk_f32,
    unsigned long long* __restrict__ blk_err)
{
    __shared__ T stg[TileCfg<T, TM>::STG];
    __shared__ int32_t rpos[TM];
    __shared__ int32_t cidx[TN];
    const UpdTile tl = tiles[blockIdx.x];
    if constexpr (MODE == 3) {
        __shared__ typename Ar<T, PR32>::S p32[TM * TN];
        upd_tile<T, TM, PR64, true>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx, p32, blk_err);
    } else if constexpr (MODE == 1) {
        upd_tile<T, TM, PR32>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx);
    } else if constexpr (MODE == 2) {
        if (blk_f32[tl.block]) upd_tile<T, TM, PR32>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx);
        else upd_tile<T, TM, PR64>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx);
    } else if constexpr (MODE == 4) {
        upd_tile<T, TM, PRBF>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx);
    } else {
        upd_tile<T, TM, PR64>(tl, cols, rowmap, W, ldw, X, stg, rpos, cidx);
    }
}

// the solved segments back into W (one workgroup per task), for the U values copied out after the panel
template <class T>
__global__ __launch_bounds__(TB) void k_x_back(const UpdTask* __restrict__ tasks, const int32_t* __restrict__ cols,
                                                const BlockDesc* __restrict__ blocks, double* __restrict__ W, int64_t ldw,
                                                const double* __restrict__ xbuf)
{
    const UpdTask tk = tasks[blockIdx.x];
    const int32_t m = blocks[tk.block].w - tk.k0;   // at most 256 = TB rows: one per thread
    const T* X = reinterpret_cast<const T*>(xbuf) + tk.xoff;
    T* S = reinterpret_cast<T*>(W) + tk.spos0;
    const int t = threadIdx.x;
    if (t >= m) return;
#pragma unroll 4
    for (int32_t c = 0; c < tk.ncol; c++) S[t + (int64_t)cols[tk.col_off + c] * ldw] = X[t + (int64_t)c * m];
}

// ---- dense partial-pivot LU of the panel's lower section

constexpr int LU_TB = 1024;

// one pivot candidate merged into another: the larger modulus, the lower row on equal moduli (rows < 0: no candidate);
// tie: another candidate has the kept one's modulus exactly (the panel LU counts such columns: there the reference,
// which takes the first candidate in its column's own search order, can take another row)
__device__ inline void cand_merge(double& bv, int32_t& br, int& bt, double ov, int32_t orr, int ot)
{
    if (orr < 0) return;
    if (br < 0 || ov > bv) { bv = ov; br = orr; bt = ot; return; }
    if (ov == bv) {
        bt = 1;
        if (orr < br) br = orr;
    }
}

template <class T>
__global__ __launch_bounds__(LU_TB) void k_getf2(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a,
                                                  int32_t nb, const double* __restrict__ colmax, double pivot_rel,
                                                  int32_t* __restrict__ piv, PanelStatus* __restrict__ st,
                                                  PanelOut* __restrict__ hout)
{
    __shared__ double bestv[LU_TB / 64];
    __shared__ int32_t bestr[LU_TB / 64], bestt[LU_TB / 64];
    __shared__ int32_t pr;
    T* Wl = reinterpret_cast<T*>(Wd);
    if (st->err) return;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int32_t jend = min(a + nb, P);
    int32_t nties = 0;   // (thread 0) the sub-panel's columns with an exact tie for the pivot
    for (int32_t j = a; j < jend; j++) {
        // argmax of the modulus over rows j .. nr-1, the lowest row on ties
        double bv = -1.0;
        int32_t br = -1;
        int bt = 0;
        for (int32_t r = j + threadIdx.x; r < nr; r += LU_TB) {
            const double v = Sc<T>::absval(Wl[r + j * ld]);
            if (v > bv) { bv = v; br = r; bt = 0; }
            else if (v == bv) bt = 1;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int32_t orr = __shfl_xor(br, off);
            const int ot = __shfl_xor(bt, off);
            cand_merge(bv, br, bt, ov, orr, ot);
        }
        if (lane == 0) { bestv[wid] = bv; bestr[wid] = br; bestt[wid] = bt; }
        __syncthreads();
        if (threadIdx.x == 0) {
            double v = -1.0;
            int32_t r = -1;
            int t = 0;
            for (int w = 0; w < LU_TB / 64; w++) cand_merge(v, r, t, bestv[w], bestr[w], bestt[w]);
            if (t && v > 0.0) nties++;
            const double thr = pivot_rel * colmax[j];
            if (r < 0 || v <= 0.0 || v < thr) {
                st->err = j + 1;
                st->best = v > 0.0 ? v : 0.0;
                st->thr = thr;
                hout->st = *st;
                pr = -1;
            } else {
                pr = r;
                piv[j] = r;
                hout->piv[j] = r;
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
    if (threadIdx.x == 0) {   // the sub-panel is factored: its status for the host
        hout->st.ties = nties;
        hout->st.err = 0;
    }
}

// The same sub-panel LU on many workgroups at once (a grid the device holds at once: every workgroup is resident).  Workgroup g
// holds the sub-panel's nb columns of the positions a + 256 g + t in registers, one row per thread; per column, every
// workgroup posts its best candidate (modulus, row, the row's values) and the owner of position j posts row j, then
// all meet at a barrier (a counter that only grows: this launch's arrivals start at `base`), pick the same pivot (largest modulus, lowest row on ties) and update
// their rows.  Slots alternate between two buffers by column parity: a workgroup can be at most one barrier ahead.
template <class T> struct Cand {
    double val;
    int32_t row, pad;
    T v[PNB];
};
constexpr long COOP_SPIN_LIMIT = 1L << 24;   // a barrier not met after this many polls ends the kernel with an error

template <class T>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_getf2_coop(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a,
                                                    int32_t nb, const double* __restrict__ colmax, double pivot_rel,
                                                    int32_t* __restrict__ piv, PanelStatus* __restrict__ st,
                                                    Cand<T>* __restrict__ cand, T* __restrict__ rowj,
                                                    int32_t* __restrict__ counter, int32_t base, PanelOut* __restrict__ hout)
{
    __shared__ double rv[TB / 64];
    __shared__ int32_t rr[TB / 64], rtie[TB / 64];
    __shared__ int32_t sh_tie;
    int32_t nties = 0;   // (workgroup 0, thread 0) the sub-panel's columns with an exact tie for the pivot
    __shared__ T pvals[PNB];
    __shared__ T Lsub[PNB][PNB + 1];
    __shared__ int32_t sh_row, sh_gi, sh_abort;
    __shared__ double sh_val;
    T* Wl = reinterpret_cast<T*>(Wd);
    const int G = gridDim.x, g = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    if (st->err) return;
    const int32_t r = a + g * TB + tid;
    const bool have = r < nr;
    T x[PNB];
#pragma unroll
    for (int c = 0; c < PNB; c++) x[c] = (have && c < nb) ? Wl[r + (int64_t)(a + c) * ld] : Sc<T>::zero();
    for (int jj = 0; jj < nb; jj++) {
        const int32_t j = a + jj;
        const int par = jj & 1;
        double bv = -1.0;
        int32_t br = -1;
        int bt = 0;
        if (have && r >= j) {
#pragma unroll
            for (int c = 0; c < PNB; c++)
                if (c == jj) bv = Sc<T>::absval(x[c]);
            br = r;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off);
            const int32_t orr = __shfl_xor(br, off);
            const int ot = __shfl_xor(bt, off);
            cand_merge(bv, br, bt, ov, orr, ot);
        }
        if (lane == 0) { rv[wid] = bv; rr[wid] = br; rtie[wid] = bt; }
        __syncthreads();
        if (tid == 0) {
            double v = -1.0;
            int32_t row = -1;
            int t = 0;
            for (int w = 0; w < TB / 64; w++) cand_merge(v, row, t, rv[w], rr[w], rtie[w]);
            cand[par * G + g].val = v;
            cand[par * G + g].row = row;
            cand[par * G + g].pad = t;
            sh_row = row;
        }
        __syncthreads();
        if (have && r == sh_row) {
#pragma unroll
            for (int c = 0; c < PNB; c++) cand[par * G + g].v[c] = x[c];
        }
        if (have && r == j) {
#pragma unroll
            for (int c = 0; c < PNB; c++) rowj[par * PNB + c] = x[c];
        }
        __syncthreads();
        // one thread per workgroup: the release (the workgroup's stores above are complete at the barrier), the arrival,
        // the wait, the acquire
        if (tid == 0) {
            __threadfence();
            __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            const int32_t target = base + (jj + 1) * G;
            long spins = 0;
            int ab = 0;
            while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > COOP_SPIN_LIMIT) { ab = 1; break; }
                __builtin_amdgcn_s_sleep(1);
            }
            __threadfence();
            sh_abort = ab;
        }
        __syncthreads();
        if (sh_abort) {
            if (tid == 0) {
                atomicExch(&st->err, -1);
                hout->st.err = -1;
                // counter[1]: a flag that is never cleared while the factorization runs -- the host reads it at the end,
                // so a barrier given up anywhere fails the factorization even when the other workgroups went on
                atomicExch(counter + 1, 1);
            }
            return;
        }
        if (wid == 0) {
            double v = -1.0;
            int32_t row = -1, gi = -1;
            int t = 0;
            for (int q = lane; q < G; q += 64) {
                const int32_t r0 = row;
                cand_merge(v, row, t, cand[par * G + q].val, cand[par * G + q].row, cand[par * G + q].pad);
                if (row != r0) gi = q;
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(v, off);
                const int32_t orr = __shfl_xor(row, off), og = __shfl_xor(gi, off);
                const int ot = __shfl_xor(t, off);
                const int32_t r0 = row;
                cand_merge(v, row, t, ov, orr, ot);
                if (row != r0) gi = og;
            }
            if (lane == 0) { sh_val = v; sh_row = row; sh_gi = gi; sh_tie = t; }
        }
        __syncthreads();
        const double thr = pivot_rel * colmax[j];
        const int32_t pr = sh_row;
        if (pr < 0 || sh_val <= 0.0 || sh_val < thr) {
            if (g == 0 && tid == 0) {
                st->err = j + 1;
                st->best = sh_val > 0.0 ? sh_val : 0.0;
                st->thr = thr;
                hout->st.best = st->best;
                hout->st.thr = thr;
                hout->st.err = j + 1;
            }
            return;
        }
        if (tid < PNB) pvals[tid] = cand[par * G + sh_gi].v[tid];
        __syncthreads();
        if (g == 0 && tid == 0) {
            piv[j] = pr;
            hout->piv[j] = pr;
            nties += sh_tie;
        }
        if (pr != j) {
            if (have && r == j) {
#pragma unroll
                for (int c = 0; c < PNB; c++) x[c] = pvals[c];
            } else if (have && r == pr) {
#pragma unroll
                for (int c = 0; c < PNB; c++) x[c] = rowj[par * PNB + c];
            }
            if (g == 0)
                for (int32_t c = tid; c < P; c += TB)
                    if (c < a || c >= a + nb) {
                        const T t0 = Wl[j + (int64_t)c * ld];
                        Wl[j + (int64_t)c * ld] = Wl[pr + (int64_t)c * ld];
                        Wl[pr + (int64_t)c * ld] = t0;
                    }
        }
        if (have && r > j) {
            T l = Sc<T>::zero();
#pragma unroll
            for (int c = 0; c < PNB; c++)
                if (c == jj) { l = Sc<T>::div(x[c], pvals[c]); x[c] = l; }
#pragma unroll
            for (int c = 0; c < PNB; c++)
                if (c > jj && c < nb) x[c] = Sc<T>::fms(x[c], l, pvals[c]);
        }
    }
    // the sub-panel is factored: its status for the host (unless a workgroup gave up at a barrier meanwhile)
    if (g == 0 && tid == 0 && __hip_atomic_load(&st->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        hout->st.ties = nties;
        hout->st.err = 0;
    }
    if (have) {
#pragma unroll
        for (int c = 0; c < PNB; c++)
            if (c < nb) Wl[r + (int64_t)(a + c) * ld] = x[c];
    }
    // the sub-panel's pivot rows in the columns right of it: the unit-lower solve with the sub-panel's pivot block, by
    // workgroup 0, which holds those rows (positions a .. a + nb - 1) and alone has swapped the columns outside the
    // sub-panel (the other workgroups never touch them)
    if (g == 0) {
        if (tid < PNB) {
#pragma unroll
            for (int c = 0; c < PNB; c++) Lsub[tid][c] = (have && tid < nb) ? x[c] : Sc<T>::zero();
        }
        __syncthreads();
        // L11 y = b, sixteen lanes per column, lane i holding y_i and taking y_k from lane k of its group; the passes of
        // TB / PNB columns are interleaved (every pass's column loaded first, each elimination step applied to all of
        // them), so the workgroup's steps overlap instead of running one pass after another
        constexpr int NPASS = TB / (TB / PNB);   // the panel is at most TB columns wide
        const int i = tid % PNB, lane = tid & 63;
        const int32_t cg = a + nb + tid / PNB;
        T y[NPASS];
#pragma unroll
        for (int q = 0; q < NPASS; q++) {
            const int32_t c = cg + q * (TB / PNB);
            y[q] = (c < P && i < nb) ? Wl[(a + i) + (int64_t)c * ld] : Sc<T>::zero();
        }
        for (int kk = 0; kk < nb - 1; kk++) {
            const T l = Lsub[i][kk];
            const bool below = i > kk;
#pragma unroll
            for (int q = 0; q < NPASS; q++) {
                const T yk = shfl_t(y[q], (lane & ~(PNB - 1)) | kk);
                if (below) y[q] = Sc<T>::fms(y[q], l, yk);
            }
        }
#pragma unroll
        for (int q = 0; q < NPASS; q++) {
            const int32_t c = cg + q * (TB / PNB);
            if (c < P && i < nb) Wl[(a + i) + (int64_t)c * ld] = y[q];
        }
    }
}

// rows [a, a + nb) of the columns right of the sub-panel: the unit-lower solve with the sub-panel's pivot block (after
// the one-workgroup panel LU; the cooperative one does it itself)
template <class T>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_trsm_sub(double* __restrict__ Wd, int64_t ld, int32_t P, int32_t a, int32_t nb,
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
    // sixteen lanes per column (TB / PNB columns per workgroup), lane i holding x_i and taking x_k from lane k
    const int i = threadIdx.x % PNB, lane = threadIdx.x & 63;
    const int32_t c = a + nb + blockIdx.x * (TB / PNB) + threadIdx.x / PNB;
    const bool mine = c < P && i < nb;
    T x = mine ? Wl[(a + i) + (int64_t)c * ld] : Sc<T>::zero();
    for (int kk = 0; kk < nb - 1; kk++) {
        const T xk = shfl_t(x, (lane & ~(PNB - 1)) | kk);
        if (i > kk) x = Sc<T>::fms(x, Ls[i][kk], xk);
    }
    if (mine) Wl[(a + i) + (int64_t)c * ld] = x;
}

// W[a+nb:nr, a+nb:P] -= W[a+nb:nr, a:a+nb] * W[a:a+nb, a+nb:P]
template <class T>
__global__ __launch_bounds__(TB) void k_gemm_sub(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a,
                                                  int32_t nb, const PanelStatus* __restrict__ st)
{
    constexpr int CT = CTile<T>::v, TC = CT / 16;
    __shared__ T As[RT][PNB + 1];
    __shared__ T Bs[PNB][CT + 1];
    if (st && st->err) return;   // (st null: a refactorization, which records small pivots itself)
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
    // the thread's part of the target, read with the operands (not after the product)
    const int tr = (threadIdx.x % 16) * 4, tc = (threadIdx.x / 16) * TC;
    T cv[4][TC];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < TC; j++)
            cv[i][j] = (tr + i < mr && tc + j < mc) ? Wl[(r0 + tr + i) + (int64_t)(c0 + tc + j) * ld] : Sc<T>::zero();
    __syncthreads();
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
            if (tr + i < mr && tc + j < mc) Wl[(r0 + tr + i) + (int64_t)(c0 + tc + j) * ld] = Sc<T>::sub(cv[i][j], acc[i][j]);
}

// ---- the panel LU with the pivots known (a refactorization in the pivot order of an earlier factorization)
//
// The lower section's rows are already in pivot order, so sub-panel [a, a + nb) is an LU without row exchanges: every
// workgroup factors the sub-panel's nb x nb diagonal block in LDS (the same operations in every workgroup, so the same
// values), then its own rows below (L21 = A21 U11^-1, one row per thread, the sub-panel's nb values in registers);
// the workgroup that reads the diagonal block last writes it back factored, and the workgroups share the solve of the
// block's rows in the columns right of the sub-panel (U12 = L11^-1 A12, sixteen lanes per column).  The operations per entry are those of the pivoting kernels (divide by
// the pivot, then subtract the multiple of the pivot row column by column).  A pivot whose modulus is below
// pivot_rel * colmax of its column: mode 0 records the first such column in st (the refactorization fails and the
// caller factors with pivoting); mode 1 replaces the pivot by pivot_rel * colmax with the pivot's phase (static
// pivoting; an exact zero becomes that real value) and counts it.
template <class T>
__global__ __launch_bounds__(TB) void k_lu_nopiv(double* __restrict__ Wd, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb,
                                                  const double* __restrict__ colmax, double pivot_rel, int mode, int64_t j0,
                                                  RefactorStatus* __restrict__ st)
{
    __shared__ T D[PNB][PNB + 1];
    __shared__ double cmx[PNB];
    T* Wl = reinterpret_cast<T*>(Wd);
    const int tid = threadIdx.x, g = blockIdx.x;
    const int dr = tid % PNB, dc = tid / PNB;   // TB = PNB * PNB: one element of the diagonal block per thread
    __shared__ int sh_last;
    D[dr][dc] = (dr < nb && dc < nb) ? Wl[(a + dr) + (int64_t)(a + dc) * ld] : Sc<T>::zero();
    if (tid < PNB) cmx[tid] = tid < nb ? colmax[a + tid] : 0.0;
    __syncthreads();
    // the factored pivot block goes back to W from the workgroup that read it last: its arrival ticket, taken once the
    // block is in LDS, is the launch's last one only when every workgroup has read the block (a workgroup may start
    // only after another one has finished, when the device is busy with other streams' work)
    if (tid == 0) {
        __threadfence();
        sh_last = atomicAdd(&st->ticket, 1u) == gridDim.x - 1;
    }
    // workgroup 0's statistics, kept in thread 0 and posted once after the loop (a barrier waits for a thread's
    // memory operations in flight: an atomic per column would put its round trip on every step)
    int32_t n_small = 0, first_small = -1;
    double min_ratio = INFINITY;
    for (int jj = 0; jj < nb; jj++) {
        // every thread takes the pivot (replaced by the same rule in all of them) and its operands before any write
        T pv = D[jj][jj];
        const double m = Sc<T>::absval(pv), thr = pivot_rel * cmx[jj];
        const bool small = !(m >= thr) || m == 0.0;
        if (small && mode == 1) pv = m > 0.0 ? Sc<T>::scale(pv, thr / m) : Sc<T>::scale(Sc<T>::one(), thr > 0.0 ? thr : 1.0);
        const bool below = dr > jj && dr < nb, right = dc > jj && dc < nb;
        const T drj = D[dr][jj], djc = D[jj][dc];
        __syncthreads();
        if (tid == 0) {
            if (small) {
                n_small++;
                if (first_small < 0) first_small = jj;
            }
            if (m > 0.0 && cmx[jj] > 0.0) min_ratio = fmin(min_ratio, m / cmx[jj]);
            D[jj][jj] = pv;
        }
        if (below) {
            const T l = Sc<T>::div(drj, pv);
            if (dc == jj) D[dr][jj] = l;
            else if (right) D[dr][dc] = Sc<T>::fms(D[dr][dc], l, djc);
        }
        __syncthreads();
    }
    if (g == 0 && tid == 0) {
        if (n_small > 0) {
            if (mode == 1) atomicAdd(&st->nreplaced, n_small);
            else atomicCAS(&st->err_col, 0, (int32_t)(j0 + a + first_small + 1));
        }
        if (min_ratio < INFINITY) atomicMin(&st->min_ratio_bits, (unsigned long long)__double_as_longlong(min_ratio));
    }
    // the pivots' reciprocals: the rows below multiply by them (one division per pivot, not one per row)
    __shared__ T rcp[PNB];
    if (tid < PNB) rcp[tid] = tid < nb ? Sc<T>::div(Sc<T>::one(), D[tid][tid]) : Sc<T>::zero();
    __syncthreads();
    // the rows below the sub-panel: workgroup g takes rows a + nb + g TB + tid
    const int32_t row = a + nb + g * TB + tid;
    if (row < nr) {
        T x[PNB];
#pragma unroll
        for (int c = 0; c < PNB; c++) x[c] = c < nb ? Wl[row + (int64_t)(a + c) * ld] : Sc<T>::zero();
#pragma unroll
        for (int jj = 0; jj < PNB; jj++) {
            if (jj < nb) {   // (no early exit: x stays in registers, indexed by constants only)
                const T l = Sc<T>::mul(x[jj], rcp[jj]);
                x[jj] = l;
#pragma unroll
                for (int c = jj + 1; c < PNB; c++)
                    if (c < nb) x[c] = Sc<T>::fms(x[c], l, D[jj][c]);
            }
        }
#pragma unroll
        for (int c = 0; c < PNB; c++)
            if (c < nb) Wl[row + (int64_t)(a + c) * ld] = x[c];
    }
    if (sh_last) {
        const int r = tid % PNB, c = tid / PNB;
        if (r < nb && c < nb) Wl[(a + r) + (int64_t)(a + c) * ld] = D[r][c];
        if (tid == 0) atomicExch(&st->ticket, 0u);   // every arrival of this launch is in: ready for the next one
    }
    // the sub-panel's pivot rows in the columns right of it, L11 y = b: sixteen lanes per column, lane i holding y_i and
    // taking y_k from lane k of its group (a column per thread would keep the whole triangle of L11 in registers); the
    // passes of TB / PNB columns go round the workgroups, each of which holds the factored pivot block
    const int i = tid % PNB, lane = tid & 63;
    for (int32_t c0 = a + nb + g * (TB / PNB); c0 < P; c0 += (int32_t)gridDim.x * (TB / PNB)) {
        const int32_t c = c0 + tid / PNB;
        const bool mine = c < P && i < nb;
        T y = mine ? Wl[(a + i) + (int64_t)c * ld] : Sc<T>::zero();
        for (int kk = 0; kk < nb - 1; kk++) {
            const T yk = shfl_t(y, (lane & ~(PNB - 1)) | kk);
            if (i > kk) y = Sc<T>::fms(y, D[i][kk], yk);
        }
        if (mine) Wl[(a + i) + (int64_t)c * ld] = y;
    }
}

// one wavefront per run
template <class T>
__global__ void k_u_runs(int32_t nruns, const URun* __restrict__ runs, const double* __restrict__ W, double* __restrict__ Ux,
                         int32_t* __restrict__ Ui)
{
    const int lane = threadIdx.x & 63;
    const int32_t r = blockIdx.x * (TB / 64) + (threadIdx.x >> 6);
    if (r >= nruns) return;
    const URun u = runs[r];
    for (int32_t t = lane; t < u.len; t += 64) {
        Ui[u.uoff + t] = u.row0 + t;
        st_val<T>(Ux, u.uoff + t, ld_val<T>(W, u.woff + t));
    }
}

template <class T>
__global__ void k_extract_cols(const ExtractCol* __restrict__ d, const int32_t* __restrict__ pos,
                               const int32_t* __restrict__ lrows, const double* __restrict__ Wl, int64_t ldw)
{
    const ExtractCol e = d[blockIdx.x];
    const T* src = reinterpret_cast<const T*>(Wl) + (int64_t)e.ci * ldw;
    T* dst = reinterpret_cast<T*>(e.dst);
    for (int32_t r = threadIdx.x; r < e.nrows; r += TB) {
        const int32_t q = pos[e.pos_off + r];
        dst[r] = src[q];
        if (e.rdst) e.rdst[r] = lrows[q];
    }
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
void extract_cols(int vs, int32_t ncols, const ExtractCol* d, const int32_t* pos, const int32_t* lrows, const double* Wl,
                  int64_t ldw, hipStream_t s)
{
    if (ncols <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_extract_cols<double>, dim3(ncols), dim3(TB), 0, s, d, pos, lrows, Wl, ldw);
    else hipLaunchKernelGGL(k_extract_cols<Zd>, dim3(ncols), dim3(TB), 0, s, d, pos, lrows, Wl, ldw);
}
void u_runs(int vs, int32_t nruns, const URun* runs, const double* W, double* Ux, int32_t* Ui, hipStream_t s)
{
    if (nruns <= 0) return;
    const unsigned g = (unsigned)((nruns + TB / 64 - 1) / (TB / 64));
    if (vs == 1) hipLaunchKernelGGL(k_u_runs<double>, dim3(g), dim3(TB), 0, s, nruns, runs, W, Ux, Ui);
    else hipLaunchKernelGGL(k_u_runs<Zd>, dim3(g), dim3(TB), 0, s, nruns, runs, W, Ux, Ui);
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
void blk_inverse(int vs, int32_t njobs, const int2* jobs, int32_t ntile, const GemmTile* tiles, const BlockDesc* blocks,
                 double flops, hipStream_t s)
{
    if (njobs <= 0) return;
    ProfScope ps("blk_inverse", flops, s);
    if (vs == 1) hipLaunchKernelGGL(k_blk_linv<double>, dim3(njobs), dim3(TB), 0, s, jobs, blocks);
    else hipLaunchKernelGGL(k_blk_linv<Zd>, dim3(njobs), dim3(TB), 0, s, jobs, blocks);
    if (ntile <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_blk_ivgemm<double>, dim3(ntile), dim3(TB), 0, s, tiles, blocks);
    else hipLaunchKernelGGL(k_blk_ivgemm<Zd>, dim3(ntile), dim3(TB), 0, s, tiles, blocks);
}
void blk_uinv(int vs, int64_t njobs, const int2* jobs, const BlockDesc* blocks, double flops, hipStream_t s)
{
    if (njobs <= 0) return;
    if (njobs > INT32_MAX) throw std::invalid_argument("blk_uinv: too many jobs for one launch");
    ProfScope ps("blk_uinv", flops, s);
    if (vs == 1) hipLaunchKernelGGL(k_blk_uinv<double>, dim3((unsigned)njobs), dim3(TB), 0, s, jobs, blocks);
    else hipLaunchKernelGGL(k_blk_uinv<Zd>, dim3((unsigned)njobs), dim3(TB), 0, s, jobs, blocks);
}
template <class T, int TM>
static void blk_upd_t(int mode, int32_t ntile, const UpdTile* tiles, const int32_t* cols, const int32_t* rowmap, double* W,
                      int64_t ldw, double* X, const uint8_t* bf, unsigned long long* be, hipStream_t s)
{
    if (mode == 1) hipLaunchKernelGGL((k_blk_upd<T, TM, 1>), dim3(ntile), dim3(TB), 0, s, tiles, cols, rowmap, W, ldw, X, bf, be);
    else if (mode == 2) hipLaunchKernelGGL((k_blk_upd<T, TM, 2>), dim3(ntile), dim3(TB), 0, s, tiles, cols, rowmap, W, ldw, X, bf, be);
    else if (mode == 3) hipLaunchKernelGGL((k_blk_upd<T, TM, 3>), dim3(ntile), dim3(TB), 0, s, tiles, cols, rowmap, W, ldw, X, bf, be);
    else if (mode == 4) hipLaunchKernelGGL((k_blk_upd<T, TM, 4>), dim3(ntile), dim3(TB), 0, s, tiles, cols, rowmap, W, ldw, X, bf, be);
    else hipLaunchKernelGGL((k_blk_upd<T, TM, 0>), dim3(ntile), dim3(TB), 0, s, tiles, cols, rowmap, W, ldw, X, bf, be);
}
void blk_upd(int vs, int32_t ntile, const UpdTile* tiles, const int32_t* cols, const int32_t* rowmap, double* W, int64_t ldw,
             int tile_rows, double flops, double* X, int prec_mode, const uint8_t* blk_f32, unsigned long long* blk_err,
             hipStream_t s)
{
    if (ntile <= 0) return;
    if (prec_mode == 2 && !blk_f32) throw std::invalid_argument("blk_upd: per-block precision without the block marks");
    if (prec_mode == 3 && !blk_err) throw std::invalid_argument("blk_upd: logging without the per-block error array");
    if (prec_mode < 0 || prec_mode > 4) throw std::invalid_argument("blk_upd: precision mode 0 to 4");
    static const char* names[5] = {"blk_upd", "blk_upd_f32", "blk_upd_mix", "blk_upd_log", "blk_upd_bf16"};
    ProfScope ps(names[prec_mode], flops, s);
    if (tile_rows == 32) {
        if (vs == 1) blk_upd_t<double, 32>(prec_mode, ntile, tiles, cols, rowmap, W, ldw, X, blk_f32, blk_err, s);
        else blk_upd_t<Zd, 32>(prec_mode, ntile, tiles, cols, rowmap, W, ldw, X, blk_f32, blk_err, s);
    } else {
        if (vs == 1) blk_upd_t<double, 64>(prec_mode, ntile, tiles, cols, rowmap, W, ldw, X, blk_f32, blk_err, s);
        else blk_upd_t<Zd, 64>(prec_mode, ntile, tiles, cols, rowmap, W, ldw, X, blk_f32, blk_err, s);
    }
}
void x_back(int vs, int32_t ntask, const UpdTask* tasks, const int32_t* cols, const BlockDesc* blocks, double* W, int64_t ldw,
            const double* xbuf, hipStream_t s)
{
    if (ntask <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_x_back<double>, dim3(ntask), dim3(TB), 0, s, tasks, cols, blocks, W, ldw, xbuf);
    else hipLaunchKernelGGL(k_x_back<Zd>, dim3(ntask), dim3(TB), 0, s, tasks, cols, blocks, W, ldw, xbuf);
}
void getf2(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const double* colmax,
           double pivot_rel, int32_t* piv, PanelStatus* st, PanelOut* hout, void* coop, int32_t* counter, int64_t* counted,
           bool allow_coop, hipStream_t s)
{
    ProfScope ps("getf2", (vs == 1 ? 2.0 : 8.0) * (double)nr * nb * nb / 2, s);
    const int32_t G = (int32_t)((nr - a + TB - 1) / TB);
    if (P > TB) throw std::invalid_argument("panel LU: the panel is wider than a workgroup");
    // the workgroups the device holds at once (occupancy times compute units), per scalar type
    // (found by the first call of each type; calls on other threads may find it too, with the same result)
    static std::atomic<int> resident[2] = {-1, -1};
    int cap = resident[vs == 1 ? 0 : 1].load(std::memory_order_relaxed);
    if (cap < 0) {
        int per = 0;
        hip_check(vs == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_getf2_coop<double>, TB, 0)
                          : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)k_getf2_coop<Zd>, TB, 0),
                  "occupancy of the panel LU");
        cap = per * devctx().ncu;
        resident[vs == 1 ? 0 : 1].store(cap, std::memory_order_relaxed);
    }
    if (allow_coop && G >= 1 && G <= std::min(cap, getf2_coop_max_wg()) && nb <= PNB) {
        // the barrier counter keeps growing across launches (nb G arrivals each); it is cleared before it could overflow
        if (*counted + (int64_t)nb * G > ((int64_t)1 << 30)) {
            k::fill_i32(counter, 1, 0, s);
            *counted = 0;
        }
        int32_t arrived = (int32_t)*counted;
        *counted += (int64_t)nb * G;
        void* args[] = {&Wl, &ld, &nr, &P, &a, &nb, (void*)&colmax, &pivot_rel, &piv, &st, &coop, (void*)&coop, &counter, &arrived,
                        &hout};
        // cand (2 G slots) first in coop, then rowj (2 PNB scalars)
        char* base = static_cast<char*>(coop);
        void* rowj = base + (size_t)2 * G * (vs == 1 ? sizeof(Cand<double>) : sizeof(Cand<Zd>));
        args[11] = &rowj;
        // an ordinary launch: a grid the device holds at once, after the stream's previous kernel has ended, has all its
        // workgroups resident together (the cooperative launch API guarantees the same, but brackets every launch with
        // a queue switch that costs about 20 us); should a workgroup not be scheduled, the bounded wait at the barrier
        // ends the kernel with an error status instead of waiting forever
        hipError_t e = vs == 1 ? hipLaunchKernel((const void*)k_getf2_coop<double>, dim3(G), dim3(TB), args, 0, s)
                               : hipLaunchKernel((const void*)k_getf2_coop<Zd>, dim3(G), dim3(TB), args, 0, s);
        hip_check(e, "panel LU launch");
        return;
    }
    if (vs == 1) hipLaunchKernelGGL(k_getf2<double>, dim3(1), dim3(LU_TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, piv, st, hout);
    else hipLaunchKernelGGL(k_getf2<Zd>, dim3(1), dim3(LU_TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, piv, st, hout);
    trsm_sub(vs, Wl, ld, P, a, nb, st, s);
}
constexpr int COOP_MAX_WG = 512;   // the most workgroups of the cooperative panel LU (its candidate slots are sized for it)
// the cap in use (a tuning knob: 0 forces the one-workgroup kernel)
static std::atomic<int> coop_cap{COOP_MAX_WG};
size_t getf2_coop_bytes(int vs) { return (size_t)2 * COOP_MAX_WG * (vs == 1 ? sizeof(Cand<double>) : sizeof(Cand<Zd>)) + 2 * PNB * 16 + 256; }
int getf2_coop_max_wg() { return coop_cap.load(std::memory_order_relaxed); }
void set_getf2_coop_max_wg(int cap) { coop_cap.store(std::max(0, std::min(cap, COOP_MAX_WG)), std::memory_order_relaxed); }
void lu_nopiv(int vs, double* Wl, int64_t ld, int32_t nr, int32_t P, int32_t a, int32_t nb, const double* colmax,
              double pivot_rel, int mode, int64_t j0, RefactorStatus* st, hipStream_t s)
{
    if (nb > PNB || P > TB || a + nb > nr) throw std::invalid_argument("lu_nopiv: sub-panel larger than the kernel takes");
    ProfScope ps("lu_nopiv", (vs == 1 ? 2.0 : 8.0) * (double)nr * nb * nb / 2, s);
    const unsigned g = (unsigned)std::max<int64_t>(1, ((int64_t)nr - a - nb + TB - 1) / TB);
    if (vs == 1) hipLaunchKernelGGL(k_lu_nopiv<double>, dim3(g), dim3(TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, mode, j0, st);
    else hipLaunchKernelGGL(k_lu_nopiv<Zd>, dim3(g), dim3(TB), 0, s, Wl, ld, nr, P, a, nb, colmax, pivot_rel, mode, j0, st);
}
void trsm_sub(int vs, double* Wl, int64_t ld, int32_t P, int32_t a, int32_t nb, const PanelStatus* st, hipStream_t s)
{
    const int32_t nc = P - a - nb;
    if (nc <= 0) return;
    ProfScope ps("trsm_sub", 0.0, s);
    const unsigned g = (unsigned)((nc + TB / PNB - 1) / (TB / PNB));
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
