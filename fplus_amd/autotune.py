"""The update-tile autotuner: times the factorization of a given matrix under each tile policy of the block updates
(tuning knob "update_tile_policy", include/fplus_amd.h) and leaves the fastest in force for the process.

The tile height of each dependency level's products (32 or 64 rows of the blocks' inverse forms against 64 columns of
the panel) trades the operand re-reads of short tiles against the parallelism of a level: which is faster depends on
the matrix's level structure and on the scalar type, so it is measured rather than predicted.  Every policy gives the
same factorization up to the rounding of the update sums (the pivots, partition and patterns do not depend on it).
"""
import time

from . import Numeric, stats, tune

POLICIES = {0: "rule: 64-row tiles on real levels with >= 4 tiles per CU, else 32", 1: "all 32-row", 2: "all 64-row",
            3: "64-row on levels with >= 4 tiles per CU (real and complex)",
            4: "64-row on levels with >= 16 tiles per CU (real and complex)"}


def autotune(sym, Ap, Ai, Ax, policies=(0, 1, 2, 3, 4), repeats=1, options=None):
    """factor (sym, Ap, Ai, Ax) under each policy `repeats` times; the fastest (median factor time) is set with
    tune("update_tile_policy", p).  Returns {"best": p, "times": {p: seconds}, "update_s": {p: block-update seconds}}."""
    times, upd = {}, {}
    for p in policies:
        tune("update_tile_policy", p)
        ts = []
        for _ in range(repeats):
            t0 = time.perf_counter()
            num = Numeric(sym, Ap, Ai, Ax, options)
            ts.append(time.perf_counter() - t0)
            st = stats(num)
            del num
        ts.sort()
        times[p] = ts[len(ts) // 2]
        upd[p] = st["t_factor"]
    best = min(times, key=times.get)
    tune("update_tile_policy", best)
    return {"best": best, "times": times, "factor_s": upd}
