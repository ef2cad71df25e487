// ruiz.hip -- Ruiz L-inf equilibration on the device, fills and value gathers (gfx950).
//
// One iteration of the reference (src/equil.c:60-183): row maxima of the squared moduli re*re + im*im, a zero row
// reported, the residual max 0.5 |log m|, the row scale 1 / sqrt(sqrt(m)) folded into Dr; the rows scaled and the
// column maxima of the scaled entries taken; a zero column reported, the column scale folded into Dc; the columns
// scaled.  Maxima are exact in any order (atomic max on the bit patterns of non-negative doubles) and every scaling is
// elementwise, so with contraction off the results are bit-identical to the serial loops: the entry-parallel kernels
// below need no ordering.  The passes are HBM-bound streams over the nnz entries.
#include "dev_common.h"
#include "device_api.h"
#include "prof.h"

#include <climits>

namespace fpa {
namespace {

constexpr int TB = 256;

inline unsigned grid_for(int64_t n, int per = TB)
{
    int64_t g = (n + per - 1) / per;
    if (g < 1) g = 1;
    if (g > 65536) g = 65536;
    return (unsigned)g;
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n, unsigned long long v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_f64(double* p, int64_t n, double v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

template <int VS>
__global__ void k_gather_vals(int64_t cnt, const int64_t* __restrict__ map, const double* __restrict__ src, double* __restrict__ dst)
{
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < cnt; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = map[o];
        if (VS == 1) dst[o] = src[q];
        else reinterpret_cast<double2*>(dst)[o] = reinterpret_cast<const double2*>(src)[q];
    }
}

template <int VS>
__device__ inline double sqmod(const double* x, int64_t q)
{
#pragma clang fp contract(off)
    if (VS == 1) { const double v = x[q]; return v * v; }
    const double2 v = reinterpret_cast<const double2*>(x)[q];
    return v.x * v.x + v.y * v.y;
}

template <int VS>
__global__ void k_ruiz_rowmax(int64_t nnz, const int32_t* __restrict__ erow, const double* __restrict__ x,
                              unsigned long long* __restrict__ rmax)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x)
        atomic_max_nonneg(&rmax[erow[q]], sqmod<VS>(x, q));
}

// per line: zero check, residual, scale = 1 / sqrt(sqrt(m)), D *= scale
__global__ void k_ruiz_line_finish(int64_t n, const unsigned long long* __restrict__ m, double* __restrict__ D,
                                   double* __restrict__ scale, k::RuizStatus* st, int is_col)
{
#pragma clang fp contract(off)
    __shared__ double red[TB];
    double r = 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double v = __longlong_as_double((long long)m[i]);
        if (v == 0.0) {
            atomicMin((unsigned long long*)(is_col ? &st->bad_col : &st->bad_row), (unsigned long long)i);
            scale[i] = 1.0;
            continue;
        }
        const double lv = 0.5 * fabs(log(v));
        if (lv > r) r = lv;
        const double s = 1.0 / sqrt(sqrt(v));
        D[i] *= s;
        scale[i] = s;
    }
    red[threadIdx.x] = r;
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomic_max_nonneg(&st->res_bits, red[0]);
}

template <int VS>
__global__ void k_ruiz_scale_colmax(int64_t nnz, const int32_t* __restrict__ erow, const int32_t* __restrict__ ecol,
                                    double* __restrict__ x, const double* __restrict__ rscale, unsigned long long* __restrict__ cmax)
{
#pragma clang fp contract(off)
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const double s = rscale[erow[q]];
        double a;
        if (VS == 1) { const double v = x[q] * s; x[q] = v; a = v * v; }
        else {
            double2 v = reinterpret_cast<double2*>(x)[q];
            v.x = v.x * s; v.y = v.y * s;
            reinterpret_cast<double2*>(x)[q] = v;
            a = v.x * v.x + v.y * v.y;
        }
        atomic_max_nonneg(&cmax[ecol[q]], a);
    }
}

template <int VS>
__global__ void k_ruiz_scale_cols(int64_t nnz, const int32_t* __restrict__ ecol, double* __restrict__ x, const double* __restrict__ cs)
{
#pragma clang fp contract(off)
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const double s = cs[ecol[q]];
        if (VS == 1) x[q] = x[q] * s;
        else { double2 v = reinterpret_cast<double2*>(x)[q]; v.x = v.x * s; v.y = v.y * s; reinterpret_cast<double2*>(x)[q] = v; }
    }
}

__global__ void k_scale_all(int64_t m, double* x, double sc)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) x[q] = x[q] * sc;
}
__global__ void k_scale_vec2(int64_t n, double* a, double* b, double h)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        a[q] = a[q] * h;
        b[q] = b[q] * h;
    }
}

__global__ void k_absminmax(int64_t m, const double* __restrict__ x, unsigned long long* amax, unsigned long long* amin)
{
    double hi = 0.0, lo = INFINITY;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x) {
        const double a = fabs(x[q]);
        if (a > hi) hi = a;
        if (a > 0.0 && a < lo) lo = a;
    }
    atomic_max_nonneg(amax, hi);
    atomicMin(amin, (unsigned long long)__double_as_longlong(lo));
}

template <int VS>
__global__ void k_line_absmax(int64_t nnz, const int32_t* __restrict__ erow, const int32_t* __restrict__ ecol,
                              const double* __restrict__ x, unsigned long long* lmax, int64_t n)
{
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nnz; q += (int64_t)gridDim.x * blockDim.x) {
        const double a = VS == 1 ? fabs(x[q]) : fmax(fabs(x[2 * q]), fabs(x[2 * q + 1]));
        atomic_max_nonneg(&lmax[erow[q]], a);
        atomic_max_nonneg(&lmax[n + ecol[q]], a);
    }
}

__global__ void k_min_positive(int64_t m, const unsigned long long* __restrict__ v, unsigned long long* out)
{
    unsigned long long lo = ULLONG_MAX;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < m; q += (int64_t)gridDim.x * blockDim.x)
        if (v[q] != 0 && v[q] < lo) lo = v[q];
    atomicMin(out, lo);
}

}  // namespace

namespace k {
void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(TB), 0, s, p, n, v);
}
void fill_f64(double* p, int64_t n, double v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(TB), 0, s, p, n, v);
}
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(TB), 0, s, p, n, v);
}
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n)), dim3(TB), 0, s, p, n, v);
}
void gather_vals(int vs, int64_t cnt, const int64_t* map, const double* src, double* dst, hipStream_t s)
{
    if (cnt <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_gather_vals<1>, dim3(grid_for(cnt)), dim3(TB), 0, s, cnt, map, src, dst);
    else hipLaunchKernelGGL(k_gather_vals<2>, dim3(grid_for(cnt)), dim3(TB), 0, s, cnt, map, src, dst);
}
void ruiz_rowmax(int vs, int64_t nnz, const int32_t* erow, const double* x, unsigned long long* rmax, hipStream_t s)
{
    if (nnz <= 0) return;
    ProfScope ps("ruiz_rowmax", (double)nnz * (8.0 * vs + 4.0), s);
    if (vs == 1) hipLaunchKernelGGL(k_ruiz_rowmax<1>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, x, rmax);
    else hipLaunchKernelGGL(k_ruiz_rowmax<2>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, x, rmax);
}
void ruiz_line_finish(int64_t n, const unsigned long long* m, double* D, double* scale, RuizStatus* st, int is_col, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_ruiz_line_finish, dim3(grid_for(n)), dim3(TB), 0, s, n, m, D, scale, st, is_col);
}
void ruiz_scale_colmax(int vs, int64_t nnz, const int32_t* erow, const int32_t* ecol, double* x, const double* rscale,
                       unsigned long long* cmax, hipStream_t s)
{
    if (nnz <= 0) return;
    ProfScope ps("ruiz_scale_colmax", (double)nnz * (16.0 * vs + 8.0), s);
    if (vs == 1) hipLaunchKernelGGL(k_ruiz_scale_colmax<1>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, ecol, x, rscale, cmax);
    else hipLaunchKernelGGL(k_ruiz_scale_colmax<2>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, ecol, x, rscale, cmax);
}
void ruiz_scale_cols(int vs, int64_t nnz, const int32_t* ecol, double* x, const double* cscale, hipStream_t s)
{
    if (nnz <= 0) return;
    ProfScope ps("ruiz_scale_cols", (double)nnz * (16.0 * vs + 4.0), s);
    if (vs == 1) hipLaunchKernelGGL(k_ruiz_scale_cols<1>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, ecol, x, cscale);
    else hipLaunchKernelGGL(k_ruiz_scale_cols<2>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, ecol, x, cscale);
}
void ruiz_scale_all(int64_t m, double* x, double sc, hipStream_t s)
{
    if (m > 0) hipLaunchKernelGGL(k_scale_all, dim3(grid_for(m)), dim3(TB), 0, s, m, x, sc);
}
void ruiz_scale_vec2(int64_t n, double* a, double* b, double h, hipStream_t s)
{
    if (n > 0) hipLaunchKernelGGL(k_scale_vec2, dim3(grid_for(n)), dim3(TB), 0, s, n, a, b, h);
}
void absminmax(int64_t m, const double* x, unsigned long long* amax, unsigned long long* amin, hipStream_t s)
{
    if (m > 0) hipLaunchKernelGGL(k_absminmax, dim3(grid_for(m)), dim3(TB), 0, s, m, x, amax, amin);
}
void line_absmax(int vs, int64_t nnz, const int32_t* erow, const int32_t* ecol, const double* x, unsigned long long* lmax,
                 int64_t n, hipStream_t s)
{
    if (nnz <= 0) return;
    if (vs == 1) hipLaunchKernelGGL(k_line_absmax<1>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, ecol, x, lmax, n);
    else hipLaunchKernelGGL(k_line_absmax<2>, dim3(grid_for(nnz)), dim3(TB), 0, s, nnz, erow, ecol, x, lmax, n);
}
void min_positive(int64_t m, const unsigned long long* v, unsigned long long* out, hipStream_t s)
{
    if (m > 0) hipLaunchKernelGGL(k_min_positive, dim3(grid_for(m)), dim3(TB), 0, s, m, v, out);
}
}  // namespace k

KernelProf& kprof()
{
    static KernelProf p;
    return p;
}

namespace {
thread_local hipStream_t t_stream = nullptr;   // the calling thread's own stream, once it asked for one
thread_local bool t_own = false;
}  // namespace

static DevCtx& devctx_process()
{
    // made once, by the first caller, under the language's guarantee for local statics (callers on other threads wait
    // for it); a failed set-up throws and leaves it to the next call
    static DevCtx ctx = [] {
        DevCtx c;
        int cnt = 0;
        hipError_t e = hipGetDeviceCount(&cnt);
        if (e != hipSuccess || cnt < 1) throw GpuError("no HIP device available (the FPLUS MI355X build runs on the GPU only)");
        hip_check(hipSetDevice(0), "hipSetDevice");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
        c.device = 0;
        c.ncu = prop.multiProcessorCount;
        hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
        return c;
    }();
    return ctx;
}

DevCtx& devctx()
{
    DevCtx& ctx = devctx_process();
    if (!t_own) return ctx;
    // a thread with its own stream: the process's context with that stream
    thread_local DevCtx tl;
    tl = ctx;
    tl.stream = t_stream;
    return tl;
}

void set_thread_stream(bool on)
{
    devctx_process();
    if (on && !t_stream)   // kept until the process ends (no HIP call at exit), at the default priority
        hip_check(hipStreamCreateWithFlags(&t_stream, hipStreamNonBlocking), "hipStreamCreate (thread)");
    t_own = on;
}

}  // namespace fpa
