// hostcopy.cpp -- copies between device memory and pageable host memory (the caller's arrays, std::vector storage)
// through page-locked bounce buffers.
//
// No DMA reads or writes pageable memory: the library stages every such copy through page-locked memory it allocated
// itself (hipHostMalloc), so the runtime never has to lock the caller's pages for a transfer (locking existing host
// memory is refused on some systems) and a copy's completion is known on return.  The cost is one host memcpy per
// copy, in 16 MB chunks that overlap the device transfers of the neighbouring chunks.
#include "device_api.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace fpa {
namespace {

// two page-locked halves used in turn, each with the event behind its last copy; process-wide (a mutex serialises the
// callers) and kept until the process ends
struct Bounce {
    static constexpr size_t CH = (size_t)16 << 20;
    std::mutex mu;
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool pending[2] = {false, false};
    void init()
    {
        if (buf[1]) return;
        for (int i = 0; i < 2; i++) {
            if (!buf[i]) hip_check(hipHostMalloc((void**)&buf[i], CH, hipHostMallocDefault), "hipHostMalloc (bounce)");
            if (!ev[i]) hip_check(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreateWithFlags (bounce)");
        }
    }
    void wait(int i)
    {
        if (pending[i]) hip_check(hipEventSynchronize(ev[i]), "hipEventSynchronize (bounce)");
        pending[i] = false;
    }
};

Bounce& bounce()
{
    static Bounce* b = new Bounce();   // never destroyed: no HIP call at process exit
    return *b;
}

}  // namespace

void copy_h2d(void* d, const void* h, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    Bounce& B = bounce();
    std::lock_guard<std::mutex> g(B.mu);
    B.init();
    int i = 0;
    for (size_t off = 0; off < bytes; off += Bounce::CH, i ^= 1) {
        const size_t c = std::min(Bounce::CH, bytes - off);
        B.wait(i);
        std::memcpy(B.buf[i], static_cast<const char*>(h) + off, c);
        hip_check(hipMemcpyAsync(static_cast<char*>(d) + off, B.buf[i], c, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
        hip_check(hipEventRecord(B.ev[i], s), "hipEventRecord (bounce)");
        B.pending[i] = true;
    }
}

void copy_d2h(void* h, const void* d, size_t bytes, hipStream_t s)
{
    if (!bytes) return;
    Bounce& B = bounce();
    std::lock_guard<std::mutex> g(B.mu);
    B.init();
    B.wait(0);
    B.wait(1);
    const size_t nch = (bytes + Bounce::CH - 1) / Bounce::CH;
    auto issue = [&](size_t k) {
        const int i = (int)(k & 1);
        const size_t off = k * Bounce::CH, c = std::min(Bounce::CH, bytes - off);
        hip_check(hipMemcpyAsync(B.buf[i], static_cast<const char*>(d) + off, c, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
        hip_check(hipEventRecord(B.ev[i], s), "hipEventRecord (bounce)");
        B.pending[i] = true;
    };
    issue(0);
    for (size_t k = 0; k < nch; k++) {
        if (k + 1 < nch) issue(k + 1);   // the next chunk lands in the other half while this one is copied out
        const int i = (int)(k & 1);
        B.wait(i);
        const size_t off = k * Bounce::CH, c = std::min(Bounce::CH, bytes - off);
        std::memcpy(static_cast<char*>(h) + off, B.buf[i], c);
    }
}

}  // namespace fpa
