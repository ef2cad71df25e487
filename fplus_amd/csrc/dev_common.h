// dev_common.h -- scalar types and small helpers shared by the HIP sources (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpa {

// complex double, interleaved (re, im): the layout of C99 double complex
struct Zd {
    double x, y;
};

__host__ __device__ inline Zd zmake(double a, double b) { Zd r; r.x = a; r.y = b; return r; }

// A pointer the compiler cannot place (read from a descriptor in memory) is generic, and a generic access counts
// against the LDS counter as well as the vector-memory one: every wait for an LDS read then also waits for it, which
// undoes a prefetch.  Device buffers reached that way are made global pointers with as_global.
template <class T> using gptr = T __attribute__((address_space(1)))*;
template <class T> __device__ inline gptr<T> as_global(T* p) { return (gptr<T>)p; }
// element i through a global pointer (a complex one as one 16-byte access)
typedef double gdv2 __attribute__((ext_vector_type(2)));
__device__ inline double gload(gptr<const double> p, int64_t i) { return p[i]; }
__device__ inline Zd gload(gptr<const Zd> p, int64_t i)
{
    const gdv2 v = ((gptr<const gdv2>)p)[i];
    return zmake(v.x, v.y);
}
__device__ inline void gstore(gptr<double> p, int64_t i, double v) { p[i] = v; }
__device__ inline void gstore(gptr<Zd> p, int64_t i, Zd v) { ((gptr<gdv2>)p)[i] = gdv2{v.x, v.y}; }

template <class T> struct Sc;
template <> struct Sc<double> {
    static constexpr int vs = 1;
    __device__ static inline double zero() { return 0.0; }
    __device__ static inline double one() { return 1.0; }
    __device__ static inline double mul(double a, double b) { return a * b; }
    __device__ static inline double fms(double acc, double a, double b) { return acc - a * b; }   // acc - a b
    __device__ static inline double fma_(double acc, double a, double b) { return acc + a * b; }  // acc + a b
    __device__ static inline double sub(double a, double b) { return a - b; }
    __device__ static inline double add(double a, double b) { return a + b; }
    __device__ static inline double div(double a, double b) { return a / b; }
    __device__ static inline double absval(double a) { return fabs(a); }
    __device__ static inline double conj(double a) { return a; }
    __device__ static inline double scale(double a, double s) { return a * s; }
    __device__ static inline bool iszero(double a) { return a == 0.0; }
};
template <> struct Sc<Zd> {
    static constexpr int vs = 2;
    __device__ static inline Zd zero() { return zmake(0.0, 0.0); }
    __device__ static inline Zd one() { return zmake(1.0, 0.0); }
    __device__ static inline Zd mul(Zd a, Zd b) { return zmake(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
    __device__ static inline Zd fms(Zd acc, Zd a, Zd b)
    {
        return zmake(acc.x - (a.x * b.x - a.y * b.y), acc.y - (a.x * b.y + a.y * b.x));
    }
    __device__ static inline Zd fma_(Zd acc, Zd a, Zd b)
    {
        return zmake(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
    }
    __device__ static inline Zd sub(Zd a, Zd b) { return zmake(a.x - b.x, a.y - b.y); }
    __device__ static inline Zd add(Zd a, Zd b) { return zmake(a.x + b.x, a.y + b.y); }
    // Smith's algorithm: no overflow of the intermediate c^2 + d^2
    __device__ static inline Zd div(Zd a, Zd b)
    {
        if (fabs(b.x) >= fabs(b.y)) {
            const double r = b.y / b.x, d = b.x + b.y * r;
            return zmake((a.x + a.y * r) / d, (a.y - a.x * r) / d);
        }
        const double r = b.x / b.y, d = b.y + b.x * r;
        return zmake((a.x * r + a.y) / d, (a.y * r - a.x) / d);
    }
    __device__ static inline double absval(Zd a) { return hypot(a.x, a.y); }
    __device__ static inline Zd conj(Zd a) { return zmake(a.x, -a.y); }
    __device__ static inline Zd scale(Zd a, double s) { return zmake(a.x * s, a.y * s); }
    __device__ static inline bool iszero(Zd a) { return a.x == 0.0 && a.y == 0.0; }
};

__device__ inline void atomic_sub(double* p, double v) { atomicAdd(p, -v); }
__device__ inline void atomic_sub(Zd* p, Zd v)
{
    atomicAdd(&p->x, -v.x);
    atomicAdd(&p->y, -v.y);
}

// non-negative doubles order like their bit patterns
__device__ inline void atomic_max_nonneg(unsigned long long* p, double v)
{
    atomicMax(p, (unsigned long long)__double_as_longlong(v));
}

}  // namespace fpa
