// rowlu_dev.h -- the row kernel on the device: right-looking elimination on sparse rows with full, partial or rook
// pivoting and the incomplete-LU drop rules (reference src/factor_row.c:1191-1560), one workgroup per system.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpa {

constexpr int32_t RL_NONE = 0x7fffffff;   // the leading column of an empty row

// what the kernel reports back, and what a run that resumes an earlier one reads
struct RowLuStatus {
    int32_t code;        // RL_OK, RL_SINGULAR, RL_POOL / RL_LOG (entry pool / multiplier log too small for the next
                         // step: the state is that of the step's start, or of its elimination when phase is 1),
                         // RL_RESTART (a pool ran out within a step: run again from the start), RL_INTERNAL
    int32_t phase;       // where a resumed run starts step `step`: 0 its pivot search, 1 its elimination
    int64_t step;        // the step that stopped the elimination
    int64_t nrep;        // pivots replaced at incomplete-LU breakdowns
    double best, thr;    // the failing step's best pivot magnitude; the singularity threshold
    double max0;         // max |a_ij| of the input (and the first pivot)
    double in_max;       // max |a_ij| of the input
    int64_t pool_used, log_used;
    int64_t merges;      // row merges of this run
    double merge_bytes;  // their algorithmic bytes: each merged row's and the pivot row's entries read, the kept entries
                         // and the multiplier written (a column index and a value each)
};
enum { RL_OK = 0, RL_SINGULAR = 1, RL_POOL = 2, RL_LOG = 3, RL_INTERNAL = 4, RL_RESTART = 5 };

// The device state of one system.  Rows are addressed by a physical id r (the input row r at the start); slot[i] is
// the physical row at position i, p[i] / q[i] the input row / column at position i.  Row r's entries, columns
// ascending, sit at cols/vals[off[r] .. off[r] + nnz[r]) inside its block [base[r], cap_end[r]) of the entry pool;
// lead[i] is the first column of the row at position i (RL_NONE when empty).  With column exchanges (full and rook
// pivoting) the rows finished at earlier steps hold input column ids (q of their step), so that an exchange relabels
// the active rows only; the host maps them to the final positions.  Blocks that outgrow their room move to fresh space at the pool's
// top (pool_top).  Each multiplier m(i, k) is logged as (physical row, step, value) at log_top.  scratch: n entries
// per wavefront (the union of two rows before its drops).  vals / log_val / scratch_val hold T (double or complex).
struct RowLuArgs {
    int32_t n;
    int32_t strategy;                     // 0 full, 1 partial, 2 rook
    double drop_tol, pivot_rel;
    int64_t *off, *base, *cap_end;
    int32_t *nnz, *lead, *slot;
    int64_t *p, *q;
    int32_t* cols;
    double* vals;
    int64_t pool_cap;
    unsigned long long* pool_top;
    int32_t *log_row, *log_step;
    double* log_val;
    int64_t log_cap;
    unsigned long long* log_top;
    int32_t* cand;                        // positions of the rows with an entry in the pivot column (n)
    int32_t* scratch_col;
    double* scratch_val;
    RowLuStatus* status;
    double* rmax;                         // full pivoting: the largest |entry| of the row at each position (-1 empty)
    int32_t* rmaxc;                       // and the first column holding it
    int32_t k0;                           // the first step of this run (0, or the step a resumed run continues)
    int32_t resume;                       // 1: step k0 starts at its elimination (status holds max0, thr, nrep)
};

// the elimination of one system from step k0 on, in one launch of one workgroup (vs 1 real, 2 complex); it stops at
// the start of a step (or of its elimination) whose worst-case growth the pool or the log cannot take
void row_lu(int vs, const RowLuArgs& a, hipStream_t s);
// nsys independent systems of one scalar type and strategy in one launch, one workgroup each (d_args: their
// arguments in device memory)
void row_lu_batch(int vs, int strategy, const RowLuArgs* d_args, int nsys, hipStream_t s);

}  // namespace fpa
