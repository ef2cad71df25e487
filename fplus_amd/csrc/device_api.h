// device_api.h -- the host-side interface of the HIP kernels (launchers) and device-memory helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace fpa {

struct GpuError : std::runtime_error {
    explicit GpuError(const std::string& m) : std::runtime_error(m) {}
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw GpuError(std::string(what) + ": " + hipGetErrorString(e));
}

// the library's device context: device 0 of the process, one stream at default priority
struct DevCtx {
    int device = -1;
    hipStream_t stream = nullptr;
    int ncu = 0;
};
DevCtx& devctx();   // initialises on first use; throws GpuError when no device is usable
// the calling thread's work on a stream of its own (on) or on the process's stream (off); see fplus_amd_thread_stream
void set_thread_stream(bool on);
// host threads of each factorization's pool (tuning knob "host_threads"; 0: the default, 15)
int host_threads();
bool set_host_threads(int n);

// copies between device memory and pageable host memory, through page-locked bounce buffers (hostcopy.cpp): copy_h2d
// returns once h may be reused (the device copy is ordered on s), copy_d2h once h holds the data
void copy_h2d(void* d, const void* h, size_t bytes, hipStream_t s);
void copy_d2h(void* h, const void* d, size_t bytes, hipStream_t s);

template <class T> struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t cnt) { alloc(cnt); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this == &o) return *this;
        release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0;
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t cnt)
    {
        release();
        n = cnt;
        hip_check(hipMalloc((void**)&p, (cnt > 0 ? cnt : 1) * sizeof(T)), "hipMalloc");
    }
    // grow to at least cnt (contents not kept)
    void reserve(size_t cnt) { if (cnt > n || !p) alloc(cnt + cnt / 4 + 16); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    // from / into pageable host memory (through the bounce buffers): h is free again, resp. filled, on return
    void upload(const T* h, size_t cnt, hipStream_t s, size_t off = 0) { copy_h2d(p + off, h, cnt * sizeof(T), s); }
    void download(T* h, size_t cnt, hipStream_t s, size_t off = 0) const { copy_d2h(h, p + off, cnt * sizeof(T), s); }
};

inline void dev_sync(hipStream_t s) { hip_check(hipStreamSynchronize(s), "hipStreamSynchronize"); }

// page-locked host memory (hipHostMalloc): device-to-host copies into it run asynchronously on the stream
template <class T> struct HostBuf {
    T* p = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void alloc(size_t cnt)
    {
        release();
        n = cnt;
        hip_check(hipHostMalloc((void**)&p, (cnt > 0 ? cnt : 1) * sizeof(T), hipHostMallocDefault), "hipHostMalloc");
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr; n = 0;
    }
};

// page-locked staging for a stream's uploads, in two areas used in turn: what is staged in an area goes to the device
// by asynchronous copies (no staging inside the runtime, no wait for the copy), and an area is reused only after the
// event recorded behind its last copy has completed -- the host can go on with the next round of work while the copies
// and the kernels reading them run
class UploadStage {
public:
    UploadStage() = default;
    UploadStage(const UploadStage&) = delete;
    UploadStage& operator=(const UploadStage&) = delete;
    ~UploadStage()
    {
        if (last_) (void)hipStreamSynchronize(last_);   // no copy still reads an area
        for (int i = 0; i < 2; i++) {
            if (ev_[i]) (void)hipEventDestroy(ev_[i]);
            for (HostBuf<char>* b : area_[i].bufs) delete b;
        }
    }
    // start a round of uploads in the other area, once its previous round's copies are done
    void begin()
    {
        cur_ ^= 1;
        Area& a = area_[cur_];
        if (!ev_[cur_]) hip_check(hipEventCreateWithFlags(&ev_[cur_], hipEventDisableTiming), "hipEventCreateWithFlags");
        if (a.recorded) hip_check(hipEventSynchronize(ev_[cur_]), "hipEventSynchronize");
        a.recorded = false;
        a.chunk = 0;
        a.used = 0;
    }
    // room for cnt T's in the current area, 256-byte aligned, valid until the area's next round
    template <class T> T* take(size_t cnt)
    {
        Area& a = area_[cur_];
        const size_t need = ((cnt > 0 ? cnt : 1) * sizeof(T) + 255) & ~(size_t)255;
        while (a.chunk < a.bufs.size() && a.used + need > a.bufs[a.chunk]->n) { a.chunk++; a.used = 0; }
        if (a.chunk == a.bufs.size()) {
            const size_t last = a.bufs.empty() ? 0 : a.bufs.back()->n;
            HostBuf<char>* b = new HostBuf<char>();
            b->alloc(std::max(need, std::max((size_t)4 << 20, 2 * last)));
            a.bufs.push_back(b);
            a.used = 0;
        }
        T* p = reinterpret_cast<T*>(a.bufs[a.chunk]->p + a.used);
        a.used += need;
        return p;
    }
    // cnt T's staged at h (from take) to the device at d, on stream s
    template <class T> void send(T* d, const T* h, size_t cnt, hipStream_t s)
    {
        if (!cnt) return;
        hip_check(hipMemcpyAsync(d, h, cnt * sizeof(T), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
        last_ = s;
    }
    // stage a copy of cnt T's and send it
    template <class T> void upload(T* d, const T* src, size_t cnt, hipStream_t s)
    {
        if (!cnt) return;
        T* h = take<T>(cnt);
        std::memcpy(h, src, cnt * sizeof(T));
        send(d, h, cnt, s);
    }
    // the round's copies are all issued: mark their end on the stream
    void end(hipStream_t s)
    {
        hip_check(hipEventRecord(ev_[cur_], s), "hipEventRecord");
        area_[cur_].recorded = true;
    }

private:
    struct Area {
        std::vector<HostBuf<char>*> bufs;
        size_t chunk = 0, used = 0;
        bool recorded = false;
    };
    Area area_[2];
    hipEvent_t ev_[2] = {nullptr, nullptr};
    hipStream_t last_ = nullptr;
    int cur_ = 1;
};

// events that mark points of a stream the host waits for one by one
struct EventList {
    std::vector<hipEvent_t> ev;
    EventList() = default;
    EventList(const EventList&) = delete;
    EventList& operator=(const EventList&) = delete;
    ~EventList() { for (hipEvent_t e : ev) (void)hipEventDestroy(e); }
    void resize(size_t cnt)
    {
        while (ev.size() < cnt) {
            hipEvent_t e;
            hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreateWithFlags");
            ev.push_back(e);
        }
    }
};

// ---------------------------------------------------------------- kernels: fills, gathers, Ruiz (ruiz.hip)
namespace k {
void fill_u64(unsigned long long* p, int64_t n, unsigned long long v, hipStream_t s);
void fill_f64(double* p, int64_t n, double v, hipStream_t s);
void fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
void fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);
// dst[o] = src[map[o]] for entries of vs doubles
void gather_vals(int vs, int64_t cnt, const int64_t* map, const double* src, double* dst, hipStream_t s);
// Ruiz, one iteration in pieces (erow / ecol: row and column of every entry, int32)
struct RuizStatus {
    unsigned long long res_bits;   // max 0.5 |log m| over the lines seen (non-negative double)
    int64_t bad_row, bad_col;      // first all-zero row / column, INT64_MAX when none
};
void ruiz_rowmax(int vs, int64_t nnz, const int32_t* erow, const double* x, unsigned long long* rmax, hipStream_t s);
void ruiz_line_finish(int64_t n, const unsigned long long* m, double* D, double* scale, RuizStatus* st, int is_col,
                      hipStream_t s);
void ruiz_scale_colmax(int vs, int64_t nnz, const int32_t* erow, const int32_t* ecol, double* x, const double* rscale,
                       unsigned long long* cmax, hipStream_t s);
void ruiz_scale_cols(int vs, int64_t nnz, const int32_t* ecol, double* x, const double* cscale, hipStream_t s);
void ruiz_scale_all(int64_t m, double* x, double sc, hipStream_t s);
void ruiz_scale_vec2(int64_t n, double* a, double* b, double h, hipStream_t s);
// |x| extremes over all doubles (bits), line maxima by max(|re|, |im|)
void absminmax(int64_t m, const double* x, unsigned long long* amax, unsigned long long* amin, hipStream_t s);
void line_absmax(int vs, int64_t nnz, const int32_t* erow, const int32_t* ecol, const double* x,
                 unsigned long long* lmax /* 2n: rows then columns */, int64_t n, hipStream_t s);
void min_positive(int64_t m, const unsigned long long* v, unsigned long long* out, hipStream_t s);
}  // namespace k

}  // namespace fpa
