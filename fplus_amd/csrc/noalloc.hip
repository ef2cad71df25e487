// noalloc.hip -- fplus_amd_solve_device_noalloc: the device solve with its scratch (the permuted right-hand sides, the
// sweeps' intermediate vector, contribution slots, partial sums and the persistent sweeps' counters) in a fixed array
// of device memory that the library holds per GPU, for callers that may not allocate (code under graph capture, serving loops, harnesses that forbid
// allocation).  No hipMalloc, hipFree or allocating library routine runs on this path.
#include "numeric_obj.h"
#include "solve_dev.h"

#include <map>
#include <mutex>

using namespace fpa;

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;   // 64 MiB per GPU, held until the process ends
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

std::mutex g_mu;                    // one call at a time per process: every call on a GPU shares that GPU's array
std::map<int, double*> g_ptr;       // the array's device address, per device
size_t g_limit = kScratchBytes;     // the bytes a call may use (lowered by fplus_amd_tune, for tests of the limit)

// the array's address on the current device, looked up once per device (the array's name is not a device address
// in host code)
double* scratch_ptr(int dev)
{
    auto it = g_ptr.find(dev);
    if (it != g_ptr.end()) return it->second;
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    g_ptr[dev] = static_cast<double*>(p);
    return static_cast<double*>(p);
}

inline size_t up256(size_t doubles) { return (doubles + 31) / 32 * 32; }   // 256-byte aligned regions

}  // namespace

namespace fpa {
size_t noalloc_scratch_bytes() { return kScratchBytes; }
bool set_noalloc_limit(size_t bytes)
{
    if (bytes > kScratchBytes) return false;
    std::lock_guard<std::mutex> g(g_mu);
    g_limit = bytes == 0 ? kScratchBytes : bytes;
    return true;
}
}  // namespace fpa

extern "C" int fplus_amd_solve_device_noalloc(const fplus_numeric* num, void* d_b, int64_t nrhs, int64_t ldb, int trans, char* err,
                                              size_t errlen)
{
    if (!num) { errf(err, errlen, "no factorization"); return FPLUS_ERR_BADARG; }
    if (num->row) return row_handle_error(err, errlen);
    const int64_t n = num->n;
    if (!num->valid) { errf(err, errlen, "the last refactorization of this factorization failed: factor again"); return FPLUS_ERR_BADARG; }
    if (trans < 0 || trans > 2 || nrhs < 0 || ldb < n || (nrhs > 0 && n > 0 && !d_b)) {
        errf(err, errlen, "fplus_amd_solve_device_noalloc: bad trans, nrhs, ldb or b");
        return FPLUS_ERR_BADARG;
    }
    if (nrhs == 0 || n == 0) return FPLUS_OK;
    const int vs = num->F.vs;
    const int tr = vs == 1 ? (trans ? 1 : 0) : trans;
    std::lock_guard<std::mutex> g(g_mu);
    // right-hand sides per pass: the most (up to the sweeps' limit) whose scratch fits
    int32_t R = (int32_t)std::min<int64_t>(k::solve_rhs_per_pass(vs), nrhs);
    SweepScratch z;
    size_t need = 0;
    for (;; R--) {
        z = sweep_scratch_size(num, R);
        const size_t n_w = (size_t)n * R * vs;
        need = (up256(n_w) + up256(z.n_y) + (tr == 0 ? up256(z.n_contrib) + up256(z.n_part) + up256((z.n_cnt + 1) / 2) : 0)) *
               sizeof(double);
        if (need <= g_limit || R == 1) break;
    }
    if (need > g_limit) {
        errf(err, errlen, "fplus_amd_solve_device_noalloc: this factorization's solve needs %zu bytes of scratch for one right-hand "
                          "side; the library's array holds %zu (use fplus_amd_solve_device)", need, g_limit);
        return FPLUS_ERR_BADARG;
    }
    try {
        DevCtx& ctx = devctx();
        hipStream_t s = ctx.stream;
        double* base = scratch_ptr(ctx.device);
        double* w = base;
        z.y = w + up256((size_t)n * R * vs);
        if (tr == 0) {
            z.contrib = z.y + up256(z.n_y);
            z.part = z.contrib + up256(z.n_contrib);
            if (z.n_cnt) z.cnt = reinterpret_cast<int32_t*>(z.part + up256(z.n_part));
        }
        std::lock_guard<std::mutex> gs(num->solve_mu);
        for (int64_t r0 = 0; r0 < nrhs; r0 += R) {
            const int32_t Rp = (int32_t)std::min<int64_t>(R, nrhs - r0);
            double* br = static_cast<double*>(d_b) + (size_t)(r0 * ldb) * vs;
            if (tr == 0) k::gather_scale(vs, n, Rp, num->d_p.p, num->d_Dr.p, br, ldb, w, s);
            else k::gather_scale(vs, n, Rp, num->d_q.p, num->d_Dc.p, br, ldb, w, s);
            device_sweeps(num, w, Rp, tr, z, s);
            if (tr == 0) k::scatter_scale(vs, n, Rp, num->d_q.p, num->d_Dc.p, w, br, ldb, s);
            else k::scatter_scale(vs, n, Rp, num->d_p.p, num->d_Dr.p, w, br, ldb, s);
        }
        dev_sync(s);
    } catch (const GpuError& e) {
        errf(err, errlen, "GPU: %s", e.what());
        return FPLUS_ERR_GPU;
    }
    return FPLUS_OK;
}
