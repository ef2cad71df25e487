// symbolic.h -- the symbolic analysis object and the per-pattern device data cached on it.
#pragma once
#include "device_api.h"

#include <cstdint>
#include <memory>
#include <mutex>
#include <vector>

namespace fpa {
// What a pattern needs for every factorization with its orders: the permuted pattern (host and device), the source
// entry of every permuted entry (the value gather), and the row / column of every permuted entry (Ruiz).  Built on
// the first factorization, reused by the next ones (same-pattern refactorizations).
struct PatternCache {
    int64_t n = 0, nnz = 0;
    std::vector<int64_t> Bp, Bi;
    DevBuf<int64_t> d_src, d_Bp;
    DevBuf<int32_t> d_Bi, d_ecol;
    // the same pattern by rows (for the residuals of the refined solves), made on first use: row pointers, the column of
    // each entry, and its position in the CSC arrays
    std::mutex csr_mu;
    bool have_csr = false;
    DevBuf<int64_t> d_Rp, d_Rpos;
    DevBuf<int32_t> d_Rj;
};
}  // namespace fpa

struct fplus_symbolic {
    int64_t n = 0, nnz = 0;
    std::vector<int64_t> row_perm, col_perm;
    int64_t n_tau_rows = 0, n_tau_cols = 0;
    mutable std::mutex mu;
    mutable std::shared_ptr<fpa::PatternCache> cache;
};
