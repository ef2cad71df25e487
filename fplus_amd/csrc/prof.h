// prof.h -- optional per-kernel timing with HIP events on the library's stream (off by default): the launches of a
// named kernel are bracketed by events; collect() waits and sums the elapsed times.  For benchmarks.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace fpa {

struct KernelProf {
    bool on = false;
    struct Rec { std::string name; hipEvent_t a, b; double work; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    struct Sum { double ms = 0, work = 0; int64_t launches = 0; };
    std::map<std::string, Sum> sums;
    std::mutex mu;
    hipEvent_t get()
    {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e;
        (void)hipEventCreate(&e);
        return e;
    }
    void begin(const char* name, double work, hipStream_t s, hipEvent_t* a)
    {
        if (!on) return;
        *a = get();
        (void)hipEventRecord(*a, s);
        (void)name; (void)work;
    }
    void end(const char* name, double work, hipStream_t s, hipEvent_t a)
    {
        if (!on) return;
        hipEvent_t b = get();
        (void)hipEventRecord(b, s);
        pending.push_back({name, a, b, work});
        if (pending.size() > 4096) collect();
    }
    void collect()
    {
        for (Rec& r : pending) {
            (void)hipEventSynchronize(r.b);
            float ms = 0;
            (void)hipEventElapsedTime(&ms, r.a, r.b);
            Sum& s = sums[r.name];
            s.ms += ms; s.work += r.work; s.launches++;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
};
KernelProf& kprof();

// brackets one launch of the named kernel; work = its algorithmic flops (or bytes)
struct ProfScope {
    const char* name; double work; hipStream_t s; hipEvent_t a = nullptr;
    ProfScope(const char* n, double w, hipStream_t st) : name(n), work(w), s(st) { kprof().begin(name, work, s, &a); }
    ~ProfScope() { kprof().end(name, work, s, a); }
};

}  // namespace fpa
