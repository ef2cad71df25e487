// host_common.h -- helpers shared by the host sources of libfplus_amd: error messages and the input checks
// (reference src/fplus_internal.h:112-119, :147-182).
#pragma once
#include "fplus.h"

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace fpa {

inline void errf(char* err, size_t errlen, const char* fmt, ...)
{
    if (!err || errlen == 0) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, errlen, fmt, ap);
    va_end(ap);
}

// ptr[0] == 0, ptr non-decreasing, indices in [0, n); canonical: strictly increasing within each line
inline int check_compressed(int64_t n, const int64_t* ptr, const int64_t* idx, bool canonical, char* err, size_t errlen)
{
    if (n < 0) { errf(err, errlen, "n must be >= 0 (got %ld)", (long)n); return FPLUS_ERR_BADARG; }
    if (ptr[0] != 0) { errf(err, errlen, "indptr[0] must be 0 (got %ld)", (long)ptr[0]); return FPLUS_ERR_BADARG; }
    for (int64_t j = 0; j < n; j++) {
        if (ptr[j + 1] < ptr[j]) { errf(err, errlen, "indptr decreases at %ld", (long)j); return FPLUS_ERR_BADARG; }
        for (int64_t q = ptr[j]; q < ptr[j + 1]; q++) {
            if (idx[q] < 0 || idx[q] >= n) {
                errf(err, errlen, "index %ld out of range [0, %ld) in line %ld", (long)idx[q], (long)n, (long)j);
                return FPLUS_ERR_BADARG;
            }
            if (canonical && q > ptr[j] && idx[q] <= idx[q - 1]) {
                errf(err, errlen, "indices not strictly increasing in line %ld (canonical form required: sorted, no duplicates)", (long)j);
                return FPLUS_ERR_BADARG;
            }
        }
    }
    return FPLUS_OK;
}

inline int check_finite(int64_t nnz, const void* x, int is_complex, char* err, size_t errlen)
{
    const double* d = static_cast<const double*>(x);
    const int64_t m = is_complex ? 2 * nnz : nnz;
    for (int64_t k = 0; k < m; k++)
        if (!std::isfinite(d[k])) {
            errf(err, errlen, "non-finite value in entry %ld", (long)(is_complex ? k / 2 : k));
            return FPLUS_ERR_BADARG;
        }
    return FPLUS_OK;
}

inline int check_perm(int64_t n, const int64_t* p, const char* name, char* err, size_t errlen)
{
    std::vector<unsigned char> seen(n > 0 ? n : 1, 0);
    for (int64_t i = 0; i < n; i++) {
        if (p[i] < 0 || p[i] >= n || seen[p[i]]) {
            errf(err, errlen, "%s is not a permutation of 0..n-1 (entry %ld)", name, (long)i);
            return FPLUS_ERR_BADARG;
        }
        seen[p[i]] = 1;
    }
    return FPLUS_OK;
}

}  // namespace fpa
