// rowlu_dev.hip -- the row kernel's elimination on the device: one workgroup of 1024 threads runs every step of one
// system (reference src/factor_row.c:1330-1560, restated for the GPU).
//
// A step is: the pivot search (a block-wide arg-max whose ties go to the first entry in the reference's scan order:
// rows by position, then columns), the row and column exchanges, the pivot row's drops (incomplete LU), and the
// elimination of every row with an entry in the pivot column.  Active rows hold only columns >= k at step k, so the
// pivot column's entry of a row, when present, is its first: the candidates are the rows whose leading column is k.
// The pivot row (without its pivot) is staged in LDS once per step; each wavefront then merges one candidate row with
// it: every entry's place in the union comes from a binary search into the other row and a ballot prefix of the
// matches, the union goes to the wavefront's scratch with dropped entries marked, and a second pass compacts it into
// the row's block (or a fresh one at the pool's top when it has outgrown its room).  Values are formed with the
// reference's operations in its order (m = a_ik / a_kk; a_ij + (-m) a_kj; (-m) a_kj), compiled without contraction, so
// real factors are bit-identical to the reference's; complex division uses Smith's method.
#include "dev_common.h"
#include "prof.h"
#include "rowlu_dev.h"

namespace fpa {
namespace {

constexpr int RL_THREADS = 1024;
constexpr int RL_WAVES = RL_THREADS / 64;
constexpr int RL_BMAX = 2048;   // pivot-row entries staged in LDS (longer rows are read from the pool)
constexpr int RL_ASTAGE = 512;  // a merged row's columns staged in its wavefront's LDS (longer rows are searched in the pool)

__device__ inline uint64_t lanes_below() { return (1ull << (threadIdx.x & 63)) - 1ull; }

__device__ inline double shx(double v, int m) { return __shfl_xor(v, m); }
__device__ inline int64_t shx(int64_t v, int m) { return (int64_t)__shfl_xor((long long)v, m); }
__device__ inline Zd shx(Zd v, int m) { return zmake(__shfl_xor(v.x, m), __shfl_xor(v.y, m)); }

// a larger magnitude wins; equal magnitudes go to the smaller key (the earlier entry in the reference's scan)
__device__ inline bool better(double a, int64_t ka, double b, int64_t kb) { return a > b || (a == b && ka < kb); }

template <class T> struct Best {
    double a;
    int64_t k;
    T v;
};

template <class T> __device__ inline Best<T> none() { return Best<T>{-1.0, INT64_MAX, Sc<T>::zero()}; }

template <class T> __device__ inline void wave_best(Best<T>& b)
{
    for (int m = 32; m > 0; m >>= 1) {
        const double oa = shx(b.a, m);
        const int64_t ok = shx(b.k, m);
        const T ov = shx(b.v, m);
        if (better(oa, ok, b.a, b.k)) { b.a = oa; b.k = ok; b.v = ov; }
    }
}

// the block's best candidate, the same in every thread
template <class T> __device__ Best<T> block_best(Best<T> b, double* s_a, int64_t* s_k, T* s_v)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_best(b);
    if (lane == 0) { s_a[wave] = b.a; s_k[wave] = b.k; s_v[wave] = b.v; }
    __syncthreads();
    if (wave == 0) {
        Best<T> c = none<T>();
        if (lane < RL_WAVES) { c.a = s_a[lane]; c.k = s_k[lane]; c.v = s_v[lane]; }
        wave_best(c);
        if (lane == 0) { s_a[RL_WAVES] = c.a; s_k[RL_WAVES] = c.k; s_v[RL_WAVES] = c.v; }
    }
    __syncthreads();
    const Best<T> r{s_a[RL_WAVES], s_k[RL_WAVES], s_v[RL_WAVES]};
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double* s_a)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m));
    if (lane == 0) s_a[wave] = v;
    __syncthreads();
    double r = 0.0;
    for (int w = 0; w < RL_WAVES; w++) r = fmax(r, s_a[w]);
    __syncthreads();
    return r;
}

// first index of c[0 .. cnt) holding a column >= x
__device__ inline int32_t lbound(const int32_t* c, int32_t cnt, int32_t x)
{
    int32_t lo = 0, hi = cnt;
    while (lo < hi) {
        const int32_t mid = (lo + hi) >> 1;
        if (c[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline int32_t find_col(const int32_t* c, int32_t cnt, int32_t x)
{
    const int32_t t = lbound(c, cnt, x);
    return t < cnt && c[t] == x ? t : -1;
}

// a value formed by the elimination stays when nonzero and, for the incomplete factorization, above drop_tol
__device__ inline bool keep_formed(double av, double dt) { return av > 0.0 && (dt == 0.0 || av > dt); }

// the largest |entry| of a row and the first column holding it (-1 and RL_NONE for an empty row), one thread
template <class T> __device__ void row_max(const int32_t* c, const T* v, int32_t nz, double& m, int32_t& mc)
{
    m = -1.0;
    mc = RL_NONE;
    for (int32_t t = 0; t < nz; t++) {
        const double av = Sc<T>::absval(v[t]);
        if (av > m) { m = av; mc = c[t]; }
    }
}

// the same over a row by one wavefront (the same in every lane)
template <class T> __device__ void wave_row_max(const int32_t* c, const T* v, int32_t nz, double& m, int32_t& mc)
{
    Best<T> b = none<T>();
    for (int32_t t = threadIdx.x & 63; t < nz; t += 64) {
        const double av = Sc<T>::absval(v[t]);
        if (better(av, c[t], b.a, b.k)) b = Best<T>{av, c[t], Sc<T>::zero()};
    }
    wave_best(b);
    m = b.a;
    mc = b.k == INT64_MAX ? RL_NONE : (int32_t)b.k;
}

// columns j1 < j2 exchange their labels in the active row at position i (src/factor_row.c:215-238)
template <class T, int S> __device__ void col_swap_row(const RowLuArgs& a, int32_t i, int32_t j1, int32_t j2)
{
    const int32_t r = a.slot[i];
    const int32_t nz = a.nnz[r];
    if (nz == 0) return;
    int32_t* c = a.cols + a.off[r];
    T* v = reinterpret_cast<T*>(a.vals) + a.off[r];
    const int32_t t1 = find_col(c, nz, j1), t2 = find_col(c, nz, j2);
    if (t1 < 0 && t2 < 0) return;
    if (t1 >= 0 && t2 >= 0) {
        const T x = v[t1];
        v[t1] = v[t2];
        v[t2] = x;
        if (S == 0) row_max(c, v, nz, a.rmax[i], a.rmaxc[i]);
        return;
    }
    int32_t t = t1 >= 0 ? t1 : t2;
    const int32_t col = t1 >= 0 ? j2 : j1;
    const T x = v[t];
    while (t + 1 < nz && c[t + 1] < col) { c[t] = c[t + 1]; v[t] = v[t + 1]; t++; }
    while (t > 0 && c[t - 1] > col) { c[t] = c[t - 1]; v[t] = v[t - 1]; t--; }
    c[t] = col;
    v[t] = x;
    a.lead[i] = c[0];
    if (S == 0) row_max(c, v, nz, a.rmax[i], a.rmaxc[i]);
}

// an incomplete-LU breakdown: the entry (k, k) of the row at position k set to rep, inserted when absent (one thread)
template <class T> __device__ bool ilu_replace(const RowLuArgs& a, int32_t k, double rep, unsigned long long* pool_top)
{
    T* vals = reinterpret_cast<T*>(a.vals);
    int32_t* cols = a.cols;
    const int32_t r = a.slot[k];
    const T rv = Sc<T>::scale(Sc<T>::one(), rep);
    int64_t o = a.off[r];
    const int32_t nz = a.nnz[r];
    if (nz > 0 && cols[o] == k) { vals[o] = rv; return true; }
    if (o > a.base[r]) {
        o--;
    } else if (o + nz < a.cap_end[r]) {
        for (int32_t t = nz; t > 0; t--) { cols[o + t] = cols[o + t - 1]; vals[o + t] = vals[o + t - 1]; }
    } else {
        const int64_t size = (int64_t)nz + nz / 2 + 8;
        const int64_t s = (int64_t)atomicAdd(pool_top, (unsigned long long)size);
        if (s + size > a.pool_cap) return false;
        for (int32_t t = 0; t < nz; t++) { cols[s + 1 + t] = cols[o + t]; vals[s + 1 + t] = vals[o + t]; }
        a.base[r] = s;
        a.cap_end[r] = s + size;
        o = s;
    }
    cols[o] = k;
    vals[o] = rv;
    a.off[r] = o;
    a.nnz[r] = nz + 1;
    a.lead[k] = k;
    return true;
}

// one wavefront eliminates column k from the row at position ip with the pivot row (piv, then Bc / Bv: its nB later
// entries) (src/factor_row.c:1507-1548 and the merge of :930-1002).  sA: the wavefront's LDS for the row's columns,
// skeep: its two LDS words of kept union positions; pool_top / log_top: the workgroup's LDS counters of the pool and
// the log.  Rows of at most 64 entries past the pivot column merged with a pivot row of at most 64 are merged in
// registers (the union's kept positions as a 128-bit mask, each entry's place its popcount); longer ones through the
// wavefront's scratch in the pool.
template <class T, int S>
__device__ void eliminate_row(const RowLuArgs& a, int32_t ip, int32_t k, T piv, const int32_t* Bc, const T* Bv, int32_t nB,
                              int32_t* scols, T* svals, int32_t* sA, unsigned long long* skeep, unsigned long long* pool_top,
                              unsigned long long* log_top, unsigned long long* s_work, int32_t* s_status)
{
    const int lane = threadIdx.x & 63;
    const int32_t r = a.slot[ip];
    int32_t* cols = a.cols;
    T* vals = reinterpret_cast<T*>(a.vals);
    const double dt = a.drop_tol;
    const int64_t o = a.off[r];
    const int32_t nz = a.nnz[r];
    const int32_t nA = nz - 1;
    const int64_t oA = o + 1;
    // the row's first chunk past column k, loaded with its entry in column k
    const bool v0 = lane < nA;
    const int32_t c0 = v0 ? cols[oA + lane] : RL_NONE;
    const T a0 = v0 ? vals[oA + lane] : Sc<T>::zero();
    const T m = Sc<T>::div(vals[o], piv);
    if (dt > 0.0 && Sc<T>::absval(m) <= dt) {   // the multiplier is dropped with its entry
        if (lane == 0) {
            a.off[r] = o + 1;
            a.nnz[r] = nz - 1;
            a.lead[ip] = nz > 1 ? c0 : RL_NONE;
        }
        if (S == 0) {
            double mx;
            int32_t mc;
            wave_row_max(cols + oA, vals + oA, nA, mx, mc);
            if (lane == 0) { a.rmax[ip] = mx; a.rmaxc[ip] = mc; }
        }
        return;
    }
    if (lane == 0) {
        const unsigned long long li = atomicAdd(log_top, 1ull);
        if ((int64_t)li < a.log_cap) {
            a.log_row[li] = r;
            a.log_step[li] = k;
            reinterpret_cast<T*>(a.log_val)[li] = m;
        } else {
            *s_status = RL_RESTART;
        }
    }
    const T alpha = Sc<T>::scale(m, -1.0);
    const uint64_t below = lanes_below();
    // the row's new place: its own block when the kept entries fit, else a fresh block at the pool's top
    auto place = [&](int32_t kept) -> int64_t {
        int64_t d = -1;
        if (lane == 0) {
            if (a.base[r] + kept <= a.cap_end[r]) {
                d = a.base[r];
            } else {
                const int64_t size = (int64_t)kept + kept / 2 + 8;
                const int64_t s = (int64_t)atomicAdd(pool_top, (unsigned long long)size);
                if (s + size <= a.pool_cap) {
                    a.base[r] = s;
                    a.cap_end[r] = s + size;
                    d = s;
                } else {
                    *s_status = RL_RESTART;
                }
            }
        }
        return (int64_t)__shfl((long long)d, 0);
    };
    if (nA <= 64 && nB <= 64) {
        // row entry (lane < nA): its place in the union and its value
        int32_t lbA = 0;
        bool mA = false;
        if (v0) { lbA = lbound(Bc, nB, c0); mA = lbA < nB && Bc[lbA] == c0; }
        const uint64_t mmA = __ballot(mA);
        const int32_t uA = lane + lbA - __popcll(mmA & below);
        T wA = a0;
        bool kA = v0;
        if (mA) { wA = Sc<T>::add(a0, Sc<T>::mul(alpha, Bv[lbA])); kA = keep_formed(Sc<T>::absval(wA), dt); }
        // pivot-row entry (lane < nB) the row lacks: fill
        if (v0) sA[lane] = c0;
        if (lane == 0) { skeep[0] = 0ull; skeep[1] = 0ull; }
        __threadfence_block();
        const bool vB = lane < nB;
        const int32_t cB = vB ? Bc[lane] : RL_NONE;
        int32_t laB = 0;
        bool mB = false;
        if (vB) { laB = lbound(sA, nA, cB); mB = laB < nA && sA[laB] == cB; }
        const uint64_t mmB = __ballot(mB);
        const int32_t uB = lane + laB - __popcll(mmB & below);
        T wB = Sc<T>::zero();
        bool kB = false;
        if (vB && !mB) { wB = Sc<T>::mul(alpha, Bv[lane]); kB = keep_formed(Sc<T>::absval(wB), dt); }
        if (kA) atomicOr(&skeep[uA >> 6], 1ull << (uA & 63));
        if (kB) atomicOr(&skeep[uB >> 6], 1ull << (uB & 63));
        __threadfence_block();
        const uint64_t w0 = skeep[0], w1 = skeep[1];
        const int32_t kept = __popcll(w0) + __popcll(w1);
        auto rank = [&](int32_t u) -> int32_t {
            return u < 64 ? __popcll(w0 & ((1ull << u) - 1ull)) : __popcll(w0) + __popcll(w1 & ((1ull << (u - 64)) - 1ull));
        };
        const int64_t d = place(kept);
        if (d < 0) return;
        if (lane == 0) atomicAdd(s_work, (unsigned long long)(nA + nB + kept + 1));
        if (kA) { const int64_t dd = d + rank(uA); cols[dd] = c0; vals[dd] = wA; }
        if (kB) { const int64_t dd = d + rank(uB); cols[dd] = cB; vals[dd] = wB; }
        int32_t first = RL_NONE;
        if (kept) {
            const int32_t u0 = w0 ? __ffsll((unsigned long long)w0) - 1 : 64 + __ffsll((unsigned long long)w1) - 1;
            const uint64_t fa = __ballot(kA && uA == u0), fb = __ballot(kB && uB == u0);
            first = fa ? __shfl(c0, __ffsll((unsigned long long)fa) - 1) : __shfl(cB, __ffsll((unsigned long long)fb) - 1);
        }
        if (S == 0) {   // the row's largest kept entry, first column among equals
            Best<T> b = none<T>();
            if (kA) b = Best<T>{Sc<T>::absval(wA), c0, Sc<T>::zero()};
            if (kB) {
                const double av = Sc<T>::absval(wB);
                if (better(av, cB, b.a, b.k)) b = Best<T>{av, cB, Sc<T>::zero()};
            }
            wave_best(b);
            if (lane == 0) { a.rmax[ip] = b.a; a.rmaxc[ip] = b.k == INT64_MAX ? RL_NONE : (int32_t)b.k; }
        }
        if (lane == 0) {
            a.off[r] = d;
            a.nnz[r] = kept;
            a.lead[ip] = first;
        }
        return;
    }
    const int32_t ucap = a.n;
    const bool astaged = nA <= RL_ASTAGE;
    // row r's entries past column k: their places in the union
    int32_t carry = 0, kept = 0;
    bool bad = false;
    for (int32_t b0 = 0; b0 < nA; b0 += 64) {
        const int32_t i = b0 + lane;
        const bool valid = i < nA;
        const int32_t c = b0 == 0 ? c0 : (valid ? cols[oA + i] : RL_NONE);
        if (valid && astaged) sA[i] = c;
        const int32_t lb = valid ? lbound(Bc, nB, c) : 0;
        const bool match = valid && lb < nB && Bc[lb] == c;
        const uint64_t mm = __ballot(match);
        const int32_t mb = carry + __popcll(mm & below);
        carry += __popcll(mm);
        bool keep = false;
        if (valid) {
            T w = b0 == 0 ? a0 : vals[oA + i];
            keep = true;
            if (match) {
                w = Sc<T>::add(w, Sc<T>::mul(alpha, Bv[lb]));
                keep = keep_formed(Sc<T>::absval(w), dt);
            }
            const int32_t u = i + lb - mb;
            if (u >= 0 && u < ucap) { scols[u] = keep ? c : -1; svals[u] = w; } else bad = true;
        }
        kept += __popcll(__ballot(keep));
    }
    const int32_t matches = carry;
    __threadfence_block();
    // the pivot row's entries that row r lacks: fill
    const int32_t* Ac = astaged ? sA : cols + oA;
    carry = 0;
    for (int32_t b0 = 0; b0 < nB; b0 += 64) {
        const int32_t j = b0 + lane;
        const bool valid = j < nB;
        const int32_t c = valid ? Bc[j] : RL_NONE;
        const int32_t la = valid ? lbound(Ac, nA, c) : 0;
        const bool match = valid && la < nA && Ac[la] == c;
        const uint64_t mm = __ballot(match);
        const int32_t mb = carry + __popcll(mm & below);
        carry += __popcll(mm);
        bool keep = false;
        if (valid && !match) {
            const T w = Sc<T>::mul(alpha, Bv[j]);
            keep = keep_formed(Sc<T>::absval(w), dt);
            const int32_t u = j + la - mb;
            if (u >= 0 && u < ucap) { scols[u] = keep ? c : -1; svals[u] = w; } else bad = true;
        }
        kept += __popcll(__ballot(keep));
    }
    if (__ballot(bad)) {
        if (lane == 0) *s_status = RL_INTERNAL;
        return;
    }
    const int32_t U = nA + nB - matches;
    __threadfence_block();
    const int64_t d = place(kept);
    if (d < 0) return;
    if (lane == 0) atomicAdd(s_work, (unsigned long long)(nA + nB + kept + 1));
    int32_t w = 0, first = RL_NONE;
    Best<T> rb = none<T>();
    for (int32_t b0 = 0; b0 < U; b0 += 64) {
        const int32_t u = b0 + lane;
        const bool valid = u < U;
        const int32_t c = valid ? scols[u] : -1;
        const T v = valid ? svals[u] : Sc<T>::zero();
        const bool keep = c >= 0;
        const uint64_t mm = __ballot(keep);
        if (keep) {
            const int64_t dd = d + w + __popcll(mm & below);
            cols[dd] = c;
            vals[dd] = v;
            const double av = Sc<T>::absval(v);
            if (better(av, c, rb.a, rb.k)) rb = Best<T>{av, c, Sc<T>::zero()};
        }
        if (w == 0 && mm) first = __shfl(c, __ffsll((unsigned long long)mm) - 1);
        w += __popcll(mm);
    }
    if (S == 0) {
        wave_best(rb);
        if (lane == 0) { a.rmax[ip] = rb.a; a.rmaxc[ip] = rb.k == INT64_MAX ? RL_NONE : (int32_t)rb.k; }
    }
    if (lane == 0) {
        a.off[r] = d;
        a.nnz[r] = w;
        a.lead[ip] = w ? first : RL_NONE;
    }
}

// the elimination of one system by the whole workgroup; S: the pivot strategy (0 full, 1 partial, 2 rook)
template <class T, int S> __device__ __forceinline__ void row_lu_body(const RowLuArgs& a)
{
    __shared__ int32_t sBc[RL_BMAX];
    __shared__ T sBv[RL_BMAX];
    __shared__ double s_a[RL_WAVES + 1];
    __shared__ int64_t s_k[RL_WAVES + 1];
    __shared__ T s_v[RL_WAVES + 1];
    __shared__ int32_t s_ncand, s_status, s_rk, s_npl, s_plcol;
    __shared__ int32_t sA[RL_WAVES][RL_ASTAGE];
    __shared__ unsigned long long s_keep[RL_WAVES][2];
    __shared__ unsigned long long s_pool_top, s_log_top, s_cand_nnz, s_work, s_log0;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int32_t n = a.n;
    T* vals = reinterpret_cast<T*>(a.vals);
    int32_t* cols = a.cols;
    int32_t* scols = a.scratch_col + (int64_t)wave * n;
    T* svals = reinterpret_cast<T*>(a.scratch_val) + (int64_t)wave * n;
    const double dt = a.drop_tol;
    const uint64_t below = lanes_below();

    double in_max, max0 = 0.0, thr = 0.0, best = 0.0;
    int64_t nrep = 0, kstop = n;
    int32_t phase = 0;
    if (a.k0 == 0 && !a.resume) {
        double mx = 0.0;   // max |a_ij| of the input
        for (int32_t r = tid; r < n; r += RL_THREADS) {
            const int64_t o = a.off[r];
            const int32_t nz = a.nnz[r];
            for (int32_t t = 0; t < nz; t++) mx = fmax(mx, Sc<T>::absval(vals[o + t]));
        }
        in_max = block_max(mx, s_a);
        if (S == 0)
            for (int32_t r = tid; r < n; r += RL_THREADS) row_max(cols + a.off[r], vals + a.off[r], a.nnz[r], a.rmax[r], a.rmaxc[r]);
    } else {
        in_max = a.status->in_max;
        max0 = a.status->max0;
        thr = a.status->thr;
        nrep = a.status->nrep;
    }
    if (tid == 0) {
        s_status = RL_OK;
        s_pool_top = *a.pool_top;
        s_log_top = *a.log_top;
        s_log0 = s_log_top;
        s_work = 0;
    }

    for (int32_t k = a.k0; k < n; k++) {
        const bool elim_only = k == a.k0 && a.resume;
        if (tid == 0) { s_ncand = 0; s_cand_nnz = 0; s_npl = 0; s_plcol = -1; }
        __syncthreads();
        if (!elim_only) {
            // ---- 1. pivot search (src/factor_row.c:273-342, :423-479, :558-658)
            int64_t pi = -1, pj = -1;
            bool found = false;
            T pv = Sc<T>::zero();
            if (S == 1) {                                   // partial: the rows led by column k
                Best<T> b = none<T>();
                unsigned long long cnz = 0;
                for (int32_t i0 = (k & ~3) + 4 * tid; i0 < n; i0 += 4 * RL_THREADS) {
                    int32_t l[4];
                    if (i0 + 3 < n) {
                        const int4 v4 = *reinterpret_cast<const int4*>(a.lead + i0);
                        l[0] = v4.x; l[1] = v4.y; l[2] = v4.z; l[3] = v4.w;
                    } else {
#pragma unroll
                        for (int j = 0; j < 4; j++) l[j] = i0 + j < n ? a.lead[i0 + j] : RL_NONE;
                    }
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int32_t i = i0 + j;
                        if (i < k || l[j] != k) continue;
                        const int32_t r = a.slot[i];
                        const T v = vals[a.off[r]];
                        cnz += (unsigned long long)a.nnz[r];
                        const double av = Sc<T>::absval(v);
                        if (better(av, i, b.a, b.k)) b = Best<T>{av, i, v};
                        a.cand[atomicAdd(&s_ncand, 1)] = i;
                    }
                }
                if (cnz) atomicAdd(&s_cand_nnz, cnz);
                b = block_best(b, s_a, s_k, s_v);
                if (b.k != INT64_MAX) { pi = b.k; pv = b.v; }
                pj = k;
                found = pi >= 0 && b.a > 0.0;
            } else if (S == 0) {
                // full: every active entry, rows then columns -- each active row's largest entry and its first column
                // are kept up to date by the merges and exchanges, so the search reads one maximum per row
                Best<T> b = none<T>();
                for (int32_t i = k + tid; i < n; i += RL_THREADS) {
                    const double av = a.rmax[i];
                    if (av < 0.0) continue;
                    const int64_t key = (int64_t)i * n + a.rmaxc[i];
                    if (better(av, key, b.a, b.k)) b = Best<T>{av, key, Sc<T>::zero()};
                }
                b = block_best(b, s_a, s_k, s_v);
                if (b.k != INT64_MAX) {
                    pi = b.k / n;
                    pj = b.k % n;
                    const int32_t r = a.slot[pi];
                    const int64_t o = a.off[r];
                    if (tid == 0) {
                        const int32_t t = find_col(cols + o, a.nnz[r], (int32_t)pj);
                        if (t < 0) s_status = RL_INTERNAL;
                        s_v[RL_WAVES] = t < 0 ? Sc<T>::zero() : vals[o + t];
                    }
                    __syncthreads();
                    pv = s_v[RL_WAVES];
                    __syncthreads();
                }
                found = pi >= 0 && b.a > 0.0;
            } else {                                                  // rook: column and row scans in turn, 8 rounds
                int64_t ci = k, ri = -1;
                double babs = -1.0;
                T bval = Sc<T>::zero();
                bool done = false;
                for (int it = 0; it < 8; it++) {
                    if (it > 0) {                                      // the rows of this scan's column, listed
                        if (tid == 0) { s_npl = 0; s_plcol = (int32_t)ci; }
                        __syncthreads();
                    }
                    Best<T> b = none<T>();
                    for (int32_t i = k + tid; i < n; i += RL_THREADS) {
                        int64_t o;
                        int32_t t = 0;
                        if (ci == k) {
                            if (a.lead[i] != k) continue;
                            o = a.off[a.slot[i]];
                        } else {
                            const int32_t r = a.slot[i];
                            o = a.off[r];
                            t = find_col(cols + o, a.nnz[r], (int32_t)ci);
                            if (t < 0) continue;
                            a.cand[atomicAdd(&s_npl, 1)] = i;
                        }
                        const T v = vals[o + t];
                        const double av = Sc<T>::absval(v);
                        if (better(av, i, b.a, b.k)) b = Best<T>{av, i, v};
                    }
                    b = block_best(b, s_a, s_k, s_v);
                    if (b.k == INT64_MAX) { pi = -1; pj = -1; pv = Sc<T>::zero(); found = false; done = true; break; }
                    ri = b.k; bval = b.v; babs = b.a;
                    const int32_t r = a.slot[ri];
                    const int64_t o = a.off[r];
                    const int32_t nz = a.nnz[r];
                    Best<T> c = none<T>();
                    for (int32_t t = tid; t < nz; t += RL_THREADS) {
                        const int32_t cc = cols[o + t];
                        if (cc < k) continue;
                        const T v = vals[o + t];
                        const double av = Sc<T>::absval(v);
                        if (better(av, cc, c.a, c.k)) c = Best<T>{av, cc, v};
                    }
                    c = block_best(c, s_a, s_k, s_v);
                    if (c.k == INT64_MAX) { pi = -1; pj = -1; pv = Sc<T>::zero(); found = false; done = true; break; }
                    if (c.k == ci) { pi = ri; pj = ci; pv = bval; found = babs > 0.0; done = true; break; }
                    ci = c.k; bval = c.v; babs = c.a;
                }
                if (!done) { pi = ri; pj = ci; pv = bval; found = babs > 0.0; }
            }
            best = Sc<T>::absval(pv);
            if (k == 0) { max0 = fmax(best, in_max); thr = a.pivot_rel * max0; }
            // ---- 1b. an incomplete-LU breakdown: the pivot (k, k) replaced by drop_tol * max|a_ij|
            if ((!found || best < thr || best == 0.0) && dt > 0.0 && dt * max0 > 0.0) {
                const double rep = dt * max0;
                if (tid == 0 && !ilu_replace<T>(a, k, rep, &s_pool_top)) s_status = RL_RESTART;
                pi = k; pj = k; found = true; best = rep;
                nrep++;
                __syncthreads();
                if (s_status != RL_OK) { kstop = k; break; }
            }
            if (!found || best < thr) {
                if (tid == 0) s_status = RL_SINGULAR;
                kstop = k;
                break;
            }
            // ---- 2, 3. the exchanges
            if (tid == 0 && pi != k) {
                const int32_t t = a.slot[k]; a.slot[k] = a.slot[pi]; a.slot[pi] = t;
                const int32_t l = a.lead[k]; a.lead[k] = a.lead[pi]; a.lead[pi] = l;
                if (S == 0) {
                    const double m = a.rmax[k]; a.rmax[k] = a.rmax[pi]; a.rmax[pi] = m;
                    const int32_t mc = a.rmaxc[k]; a.rmaxc[k] = a.rmaxc[pi]; a.rmaxc[pi] = mc;
                }
                const int64_t x = a.p[k]; a.p[k] = a.p[pi]; a.p[pi] = x;
            }
            if (S != 1) __syncthreads();
            if (pj != k) {
                // the active rows only (finished rows hold input column ids)
                if (tid == 0) { const int64_t x = a.q[k]; a.q[k] = a.q[pj]; a.q[pj] = x; }
                if (S == 2 && s_plcol == (int32_t)pj) {
                    // rook, the last column scan was pj's: the rows holding column k (led by it), then those of the
                    // scan's list that hold pj but not k
                    for (int32_t i = k + tid; i < n; i += RL_THREADS)
                        if (a.lead[i] == k) col_swap_row<T, S>(a, i, k, (int32_t)pj);
                    __syncthreads();
                    const int32_t npl = s_npl;
                    for (int32_t c = tid; c < npl; c += RL_THREADS) {
                        int32_t i = a.cand[c];
                        i = i == (int32_t)pi ? k : (i == k ? (int32_t)pi : i);   // the positions after the row exchange
                        if (a.lead[i] != k) col_swap_row<T, S>(a, i, k, (int32_t)pj);
                    }
                } else {
                    for (int32_t i = k + tid; i < n; i += RL_THREADS) col_swap_row<T, S>(a, i, k, (int32_t)pj);
                }
                __syncthreads();
            }
        }
        if (S != 1 || elim_only) {                       // the rows led by column k after the exchanges
            unsigned long long cnz = 0;
            for (int32_t i = k + 1 + tid; i < n; i += RL_THREADS)
                if (a.lead[i] == k) {
                    a.cand[atomicAdd(&s_ncand, 1)] = i;
                    cnz += (unsigned long long)a.nnz[a.slot[i]];
                }
            if (cnz) atomicAdd(&s_cand_nnz, cnz);
        }
        if (tid == 0) {
            const int32_t r = a.slot[k];
            s_rk = r;
            if (a.nnz[r] == 0 || a.lead[k] != k) s_status = RL_INTERNAL;
        }
        __syncthreads();
        if (s_status != RL_OK) { kstop = k; break; }
        const int32_t rk = s_rk;
        // ---- 4b. the pivot row, final, loses its off-pivot entries of at most drop_tol
        if (dt > 0.0 && wave == 0 && !elim_only) {
            const int64_t o = a.off[rk];
            const int32_t nz = a.nnz[rk];
            int32_t w = 0;
            for (int32_t b0 = 0; b0 < nz; b0 += 64) {
                const int32_t t = b0 + lane;
                const bool valid = t < nz;
                const int32_t c = valid ? cols[o + t] : 0;
                const T v = valid ? vals[o + t] : Sc<T>::zero();
                const bool keep = valid && (c == k || Sc<T>::absval(v) > dt);
                const uint64_t mm = __ballot(keep);
                if (keep) {
                    const int32_t d = w + __popcll(mm & below);
                    cols[o + d] = c;
                    vals[o + d] = v;
                }
                w += __popcll(mm);
            }
            if (lane == 0) a.nnz[rk] = w;
        }
        __syncthreads();
        const int64_t ok = a.off[rk];
        const int32_t nB = a.nnz[rk] - 1;
        const T piv = vals[ok];
        const int32_t ncand = s_ncand;
        // the step's worst-case growth (every candidate row moving to a fresh block) must fit the pool, its
        // multipliers the log; otherwise the run stops here, before the elimination, for the host to grow them
        if (tid == 0) {
            const unsigned long long m = s_cand_nnz + (unsigned long long)ncand * (unsigned long long)nB;
            const unsigned long long need = m + m / 2 + 8ull * (unsigned long long)ncand;
            if (s_pool_top + need > (unsigned long long)a.pool_cap) s_status = RL_POOL;
            else if (s_log_top + (unsigned long long)ncand > (unsigned long long)a.log_cap) s_status = RL_LOG;
        }
        const bool staged = nB <= RL_BMAX;
        if (staged)
            for (int32_t t = tid; t < nB; t += RL_THREADS) { sBc[t] = cols[ok + 1 + t]; sBv[t] = vals[ok + 1 + t]; }
        __syncthreads();
        if (s_status != RL_OK) { kstop = k; phase = 1; break; }
        // ---- 5. elimination: one wavefront per row led by column k
        const int32_t* Bc = staged ? sBc : cols + ok + 1;
        const T* Bv = staged ? sBv : vals + ok + 1;
        for (int32_t c = wave; c < ncand; c += RL_WAVES) {
            const int32_t i = a.cand[c];
            if (i == k || a.lead[i] != k) continue;
            eliminate_row<T, S>(a, i, k, piv, Bc, Bv, nB, scols, svals, sA[wave], s_keep[wave], &s_pool_top, &s_log_top, &s_work,
                             &s_status);
        }
        if (S != 1)   // the finished row keeps input column ids from here on; later exchanges skip it
            for (int32_t t = tid; t <= nB; t += RL_THREADS) cols[ok + t] = (int32_t)a.q[cols[ok + t]];
        __syncthreads();
        if (s_status != RL_OK) { kstop = k; break; }
    }
    if (tid == 0) {
        *a.pool_top = s_pool_top;
        *a.log_top = s_log_top;
        RowLuStatus* st = a.status;
        st->code = s_status;
        st->phase = phase;
        st->step = kstop;
        st->nrep = nrep;
        st->best = best;
        st->thr = thr;
        st->max0 = max0;
        st->in_max = in_max;
        st->pool_used = (int64_t)s_pool_top;
        st->log_used = (int64_t)s_log_top;
        st->merges = (int64_t)(s_log_top - s_log0);
        st->merge_bytes = (double)s_work * (double)(sizeof(int32_t) + sizeof(T));
    }
}

template <class T, int S> __global__ __launch_bounds__(RL_THREADS) void k_row_lu(RowLuArgs a) { row_lu_body<T, S>(a); }

// independent systems, one workgroup each
template <class T, int S> __global__ __launch_bounds__(RL_THREADS) void k_row_lu_batch(const RowLuArgs* args)
{
    const RowLuArgs a = args[blockIdx.x];
    row_lu_body<T, S>(a);
}

}  // namespace

void row_lu_batch(int vs, int strategy, const RowLuArgs* d_args, int nsys, hipStream_t s)
{
    if (nsys <= 0) return;
    ProfScope ps("row_lu_batch", 0.0, s);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(nsys), dim3(RL_THREADS), 0, s, d_args); };
    if (vs == 2) {
        if (strategy == 0) go(k_row_lu_batch<Zd, 0>);
        else if (strategy == 1) go(k_row_lu_batch<Zd, 1>);
        else go(k_row_lu_batch<Zd, 2>);
    } else {
        if (strategy == 0) go(k_row_lu_batch<double, 0>);
        else if (strategy == 1) go(k_row_lu_batch<double, 1>);
        else go(k_row_lu_batch<double, 2>);
    }
}

void row_lu(int vs, const RowLuArgs& a, hipStream_t s)
{
    ProfScope ps("row_lu", 0.0, s);
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(1), dim3(RL_THREADS), 0, s, a); };
    if (vs == 2) {
        if (a.strategy == 0) go(k_row_lu<Zd, 0>);
        else if (a.strategy == 1) go(k_row_lu<Zd, 1>);
        else go(k_row_lu<Zd, 2>);
    } else {
        if (a.strategy == 0) go(k_row_lu<double, 0>);
        else if (a.strategy == 1) go(k_row_lu<double, 1>);
        else go(k_row_lu<double, 2>);
    }
}

}  // namespace fpa
