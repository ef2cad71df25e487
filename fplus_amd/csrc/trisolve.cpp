// trisolve.cpp -- fplus_tri_solve and the solve-plan entry points on host CSC factors (placeholder until the generic
// device sweeps land).
#include "host_common.h"

struct fplus_solve_plan {
    int64_t n = 0;
};

extern "C" {
int fplus_tri_solve(int64_t, const int64_t*, const int64_t*, const void*, const int64_t*, const int64_t*, const void*, void*, int, int)
{
    return FPLUS_ERR_INTERNAL;
}
int fplus_solve_plan_create(int64_t n, const int64_t*, const int64_t*, const int64_t*, const int64_t*, fplus_solve_plan** plan,
                            char* err, size_t errlen)
{
    *plan = nullptr;
    fpa::errf(err, errlen, "not implemented");
    (void)n;
    return FPLUS_ERR_INTERNAL;
}
int fplus_tri_solve_plan(const fplus_solve_plan*, int64_t, const int64_t*, const int64_t*, const void*, const int64_t*,
                         const int64_t*, const void*, void*, int64_t, int64_t, int, int)
{
    return FPLUS_ERR_INTERNAL;
}
int64_t fplus_solve_plan_nsupernodes(const fplus_solve_plan*) { return -1; }
void fplus_solve_plan_free(fplus_solve_plan* plan) { delete plan; }
}
