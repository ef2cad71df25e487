/* fplus_amd.h -- entry points of the MI355X build beyond the reference interface (include/fplus.h).
 *
 * None of these has a counterpart in the reference; each serves a caller of this build named in its comment.
 */
#ifndef FPLUS_AMD_H
#define FPLUS_AMD_H

#include "fplus.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A symbolic analysis whose orders are given (callers that hold their own orders, e.g. the identity orders of the
 * Python SparseLU with colamd=False): A[row_perm, :][:, col_perm] is what fplus_factor factors. */
int fplus_amd_symbolic_from_orders(int64_t n, const int64_t *Ap, const int64_t *Ai, const int64_t *row_perm,
                                   const int64_t *col_perm, fplus_symbolic **sym, char *err, size_t errlen);

/* Statistics of a factorization: nnz(L), nnz(U), supernodes, device blocks, panels, dependency levels, factor flops
 * c * sum_k l_k u_k (c = 2 real, 8 complex), update GEMM flops, and the seconds spent in Ruiz, in the factorization,
 * on the host side of it, and waiting for the device, the Ruiz iterations, and the block updates' algorithmic bytes
 * (each update reads its block's rows in inverse form and its segment once, writes the solved segment once, and reads
 * and writes its target rows once).  For benchmarks and logs. */
typedef struct {
    int64_t nnzL, nnzU, nsupernodes, nblocks, panels, levels;
    double flops, gemm_flops;
    double t_ruiz, t_factor, t_host, t_wait, t_total;
    int ruiz_iters;
    double gemm_bytes;
    int64_t shared_levels;   /* dependency levels whose update tasks can write the same rows */
    int64_t max_block_w;     /* the widest device block (columns) */
    int coop_retried;        /* 1: the cooperative panel LU gave up at a barrier and the factorization ran again without it */
    int precision;           /* the block updates of the last (re)factorization: 0 FP64, 1 FP32, 2 per supernode */
    int64_t refactor_replaced;    /* the last refactorization: pivots replaced (static pivoting) */
    double refactor_min_ratio;    /* and its smallest |pivot| / (largest entry of the pivot's column) */
    double replay_bytes;          /* device bytes kept for refactorizations (schedule and inverse forms) */
    int64_t f32_blocks;           /* blocks whose updates ran in FP32 in the last (re)factorization */
    double f32_flop_share;        /* their share of the update flops */
    int64_t pivot_ties;           /* columns of the pivoting factorization whose largest candidate modulus was held by two
                                     rows or more: the device takes the lowest workspace position, the reference the first
                                     candidate of its column's search order, so there the pivots can differ */
} fplus_amd_stats;
int fplus_amd_numeric_stats(const fplus_numeric *num, fplus_amd_stats *st);

/* The row kernel's last factorization on the calling thread (fplus_factor_row, _perm, _sym, or fplus_factor with rook
 * / full pivoting or drop_tol > 0): its device runs (the first, those that resumed after growing the entry pool or the
 * multiplier log, and those that started over), the seconds of the elimination on the device and of the whole call,
 * the row merges and their algorithmic bytes (each merged row's and the pivot row's entries read, the kept entries and
 * the multiplier written, a column index and a value each), and the pool entries and multipliers it used.  For
 * benchmarks and logs; FPLUS_ERR_BADARG when the thread has made none. */
typedef struct {
    int64_t runs, restarts, merges, pool_entries, multipliers;
    double t_device, t_total, merge_bytes;
} fplus_amd_row_info;
int fplus_amd_row_stats(fplus_amd_row_info *info);

/* Many independent systems through the row kernel at once (no counterpart in the reference, which factors one system
 * per call): system i is the canonical CSR matrix (n[i], indptr[i], indices[i], data[i]), all of one scalar type and
 * factored with the same pivot_strategy, drop_tol and pivot_rel as fplus_factor_row would; each system runs as one
 * workgroup of a shared launch.  out[i] and rc[i] receive what fplus_factor_row would return for system i (CSR factors,
 * or its error code; out[i] zeroed then).  The call returns FPLUS_OK when the batch ran, or an error for the whole
 * batch (bad arguments, device failure, out of memory).  fplus_amd_row_stats then reports the whole batch. */
int fplus_amd_factor_row_batch(int64_t nsys, const int64_t *n, const int64_t *const *indptr, const int64_t *const *indices,
                               const void *const *data, int is_complex, double drop_tol, int pivot_strategy, double pivot_rel,
                               fplus_factors *out, int *rc, char *err, size_t errlen);

/* Device-resident inputs, for callers that keep the values and right-hand sides in HBM (refactorization loops,
 * benchmarks): fplus_factor with Ax a device pointer (the analysed pattern Ap, Ai on the host), and fplus_solve on a
 * device array b (column-major, leading dimension ldb), in place. */
int fplus_amd_factor_device(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *d_Ax, int is_complex,
                            const fplus_options *opt, fplus_numeric **num, char *err, size_t errlen);
int fplus_amd_solve_device(const fplus_numeric *num, void *d_b, int64_t nrhs, int64_t ldb, int trans, char *err, size_t errlen);

/* fplus_amd_solve_device without allocating: no counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The same kernels and results
 * as fplus_amd_solve_device; their scratch (the permuted right-hand sides, the sweeps' intermediate vector, contribution
 * slots and partial sums) lives in a 64 MiB array of device memory that the library holds on each GPU until the process
 * ends, instead of buffers it allocates.  Right-hand sides go in passes sized to fit the array; a factorization whose
 * solve does not fit for one right-hand side (about n > 2 M unknowns) is rejected with FPLUS_ERR_BADARG.  Every call on
 * a GPU shares that GPU's array, which keeps the last call's contents (no kernel reads what the call did not write);
 * the library runs one such call at a time per process, and work of the caller's that overlaps a call on the same
 * array (e.g. from a graph replayed on another stream) races with it.  Waits for the library's stream. */
int fplus_amd_solve_device_noalloc(const fplus_numeric *num, void *d_b, int64_t nrhs, int64_t ldb, int trans, char *err,
                                   size_t errlen);

/* Options of this build's factorization beyond fplus_options (fplus_amd_default_options: all zero).
 *   precision      the block updates (the Schur-complement products of the supernodal LU): 0 FP64 (the drop-in's);
 *                  1 the FP32 matrix cores (operands rounded to FP32, FP32 accumulation, the result subtracted in FP64);
 *                  2 per block (a supernode's columns within a panel), by the learned precision selector: as each block
 *                  is made its features go through the selector's model on the device, which marks whether the block's
 *                  updates of later columns may run in FP32; 3 FP64 updates with the FP32 products computed beside them
 *                  and their differences logged per block (fplus_amd_selector_log: the selector's training data); 4 BF16
 *                  hi/lo split products (each FP32-rounded operand split into two BF16 parts, three BF16 matrix-core
 *                  products with FP32 accumulation: about 16 bits per operand).  Pivot
 *                  selection, the panels' dense LU, Ruiz and the solves stay FP64.  A factorization with FP32 updates
 *                  is meant for fplus_amd_solve_refined, which wins the accuracy back.
 *   refactor       1: keep what fplus_amd_refactor replays -- the device schedule of this factorization and the blocks'
 *                  inverse forms (fplus_amd_stats.replay_bytes).
 *   static_pivots  (used by fplus_amd_refactor) */
typedef struct fplus_amd_selector fplus_amd_selector;
typedef struct {
    int precision;
    int refactor;
    int static_pivots;
    int reserved;
    const fplus_amd_selector *selector;   /* precision 2 (here or in a refactorization): the model that marks the blocks */
} fplus_amd_options;
void fplus_amd_default_options(fplus_amd_options *o);

/* fplus_factor with the options above; Ax in device memory when ax_on_device is 1 (the pattern Ap, Ai on the host). */
int fplus_amd_factor_ex(const fplus_symbolic *sym, const int64_t *Ap, const int64_t *Ai, const void *Ax, int ax_on_device,
                        int is_complex, const fplus_options *opt, const fplus_amd_options *aopt, fplus_numeric **num,
                        char *err, size_t errlen);

/* A same-pattern refactorization (new values Ax of the pattern num was factored with, e.g. the next shift of a
 * shift-invert sweep), in place: Ruiz on the new values, then the recorded factorization replayed in its pivot order
 * -- no pivot search, no host work per panel.  precision: as fplus_amd_options.precision.  static_pivots 0: a pivot whose
 * modulus falls below pivot_rel times its column's largest entry fails the refactorization (FPLUS_ERR_SINGULAR; factor
 * again with pivoting); 1: it is replaced by that bound with its phase (static pivoting; fplus_amd_stats counts them) and
 * fplus_amd_solve_refined recovers the accuracy.  After a failure num is unusable until a refactorization succeeds.
 * Needs a factorization made with fplus_amd_options.refactor = 1. */
int fplus_amd_refactor(fplus_numeric *num, const void *Ax, int ax_on_device, int precision, int static_pivots, char *err,
                       size_t errlen);

/* The learned precision selector: an MLP 12 -> 64 -> 64 -> 2 in FP32 over per-block features (the block's width and
 * rows, the largest entries of its inverse form, of L21 and of L11^-1, the spread and size of its pivots against their
 * columns' largest entries, the spread of the equilibration scales over its rows and columns, its position), trained
 * on logs that factorizations with precision 3 produce (each block's updates computed both ways), with batched training
 * and inference kernels on the device.  Class 1 (the block's FP32 updates keep its products within the training
 * threshold) above `threshold` marks the block FP32.
 *   create: a model with seeded random weights (untrained: set weights or train it before use).
 *   train: nsamples feature rows (12 floats each) with labels (1: FP32 safe, 0: not), SGD with momentum over
 *          minibatches of 256 for `epochs` passes; the features' means and scales are taken from the samples.
 *   predict: P(class 1) per feature row.
 *   get_weights / set_weights: the SEL_NW = 5146 floats (W1, b1, W2, b2, W3, b3, feature means, scales).
 *   log: after a factorization (or refactorization) with precision 3, per block its features, the largest difference
 *        |P32 - P64| of its update products computed in FP32 and in FP64, and the largest |P64| (both in the
 *        equilibrated scale of the factorization, Dr A Dc; 0 for a block that updated nothing). */
int fplus_amd_selector_create(uint64_t seed, fplus_amd_selector **sel);
void fplus_amd_selector_free(fplus_amd_selector *sel);
int fplus_amd_selector_train(fplus_amd_selector *sel, const float *features, const float *labels, int64_t nsamples, int epochs,
                             float lr, uint64_t seed, float *final_loss);
int fplus_amd_selector_predict(const fplus_amd_selector *sel, const float *features, int64_t nsamples, float *prob);
int fplus_amd_selector_set_threshold(fplus_amd_selector *sel, float threshold);
int fplus_amd_selector_get_weights(const fplus_amd_selector *sel, float *w, int64_t cap);
int fplus_amd_selector_set_weights(fplus_amd_selector *sel, const float *w, int64_t n);
int fplus_amd_selector_log(const fplus_numeric *num, float *features, double *diff, double *pmax, int64_t cap, int64_t *nblocks);

/* Refined solves: A x = b by iterative refinement in FP64 around the factorization, each correction solved by GMRES
 * preconditioned with the factors (GMRES-IR), in the equilibrated system Dr A Dc of the factorization -- for
 * factorizations with FP32 updates, replaced pivots (static pivoting) or a refactorization in an earlier pivot order.  The
 * residuals use the values the factorization was given (kept on the device).  Every right-hand side is refined until its
 * equilibrated normwise backward error ||Dr (b - A x)|| / (||Dr A Dc||_1 ||x / Dc|| + ||Dr b||) is at most tol, or
 * max_outer steps were taken.  b: n x nrhs column-major (leading dimension ldb), in host or device memory, overwritten
 * with x.  Defaults (fplus_amd_default_refine_options): 10 steps, GMRES restart 8, tol 1e-14, inner_tol 1e-4. */
typedef struct {
    int max_outer;      /* refinement steps at most */
    int restart;        /* GMRES iterations per step at most (1 to 32) */
    double tol;         /* the target equilibrated backward error */
    double inner_tol;   /* a step's GMRES stops once the preconditioned residual fell by this factor */
    int verbose;
    int reserved;
} fplus_amd_refine_options;
typedef struct {
    int outer;          /* refinement steps taken, over all batches of right-hand sides (0: the first solve met tol) */
    int inner;          /* GMRES iterations, each one triangular solve and one product with A */
    int converged;      /* 1: every right-hand side met tol */
    int reserved;
    double bwd_first;   /* the largest equilibrated backward error after the first solve */
    double bwd_last;    /* and at the end */
} fplus_amd_refine_info;
void fplus_amd_default_refine_options(fplus_amd_refine_options *o);
int fplus_amd_solve_refined(const fplus_numeric *num, void *b, int b_on_device, int64_t nrhs, int64_t ldb,
                            const fplus_amd_refine_options *ro, fplus_amd_refine_info *info, char *err, size_t errlen);

/* The refactor + solve sequence with the precision fallback: refactorization at `precision` (static pivoting), refined
 * solve; when the refinement stalls (not every right-hand side meets tol within max_outer steps), the factorization is
 * made again with FP64 updates and b solved again from the start.  *fallback: 0 none, 1 the FP64 refactorization ran.
 * b is kept in a device copy for the second attempt. */
int fplus_amd_refactor_solve(fplus_numeric *num, const void *Ax, int ax_on_device, void *b, int b_on_device, int64_t nrhs,
                             int64_t ldb, int precision, const fplus_amd_refine_options *ro, fplus_amd_refine_info *info,
                             int *fallback, char *err, size_t errlen);

/* A same-pattern sweep (e.g. the shifts of a shift-invert pencil): nshift value sets in Ax (nnz x nshift, the values of
 * shift k from Ax + k nnz, the pattern num was factored with) each refactored and solved as fplus_amd_refactor_solve
 * does, for its own nrhs right-hand sides in b (shift k's from b + k ldb nrhs, column-major, leading dimension ldb),
 * overwritten with x.  Ax and b in device memory when on_device is 1.  info[k] and fallback[k] (arrays of nshift, or
 * NULL) per shift.  With the tuning knob "refactor_graph" on, the refactorizations replay through a HIP graph captured
 * on the first one.  Stops at the first failing shift (the message names it). */
int fplus_amd_refactor_sweep(fplus_numeric *num, const void *Ax, int64_t nshift, void *b, int64_t nrhs, int64_t ldb,
                             int on_device, int precision, const fplus_amd_refine_options *ro, fplus_amd_refine_info *info,
                             int *fallback, char *err, size_t errlen);

/* Device memory and the library's stream for those callers: allocate, free, copy (kind 0 host to device, 1 device to
 * host, 2 device to device), wait for the library's stream. */
int fplus_amd_dev_alloc(size_t bytes, void **ptr);
int fplus_amd_dev_free(void *ptr);
int fplus_amd_dev_copy(void *dst, const void *src, size_t bytes, int kind);
int fplus_amd_dev_sync(void);

/* Independent systems side by side: on = 1 sends the library's device work from the calling thread to a stream of
 * its own (made on the first call of the thread, kept until the process ends), so that factorizations and solves made
 * on several threads overlap on the device; on = 0 returns the thread to the process's stream.  Handles made on one
 * stream may be used from another thread afterwards (the calls that return results wait for their own stream). */
int fplus_amd_thread_stream(int on);

/* Tuning knobs (process-wide; tests and benchmarks).  "panel_lu_max_workgroups": the most workgroups of the
 * cooperative panel LU (at most 512, the default); 0 runs every sub-panel on the one-workgroup kernel.
 * "xcd_tile_order": 1 (default) orders each update level's tiles so that those reading one segment slice run on one
 * XCD, 0 keeps them in task order (for measurements; applies to factorizations made afterwards).  "refactor_graph": 1
 * replays refactorizations through a HIP graph captured once per factorization and set of arguments, 0 (default)
 * launches them one by one.  "refactor_pipeline": 1 (default) replays a refactorization as a pipeline of two panels on
 * two streams -- a panel's update levels that read no block of the previous panel run while that panel finishes -- 0
 * one panel after another (the graph replay is always one after another).  "update_tile_policy": the block updates' tile height per dependency level, for
 * factorizations made afterwards -- 0 (default) 64-row tiles for real levels with at least 4 tiles per compute unit and
 * 32-row tiles otherwise, 1 all 32-row, 2 all 64-row, 3 and 4 the real rule for real and complex with the bound at 4 and
 * 16 tiles per compute unit (fplus_amd.autotune times them).  "host_threads": the host threads of each factorization's
 * pool (1 to 64; 0 the default, 15), for many factorizations at once on several threads.  "update_atomic_rows": 1 (default) makes only the update
 * tiles whose rows another task of their dependency level also writes use atomics, 0 every tile of such a level (for
 * factorizations made afterwards).  "deterministic_updates": 1 splits every dependency level of the block updates so that no
 * two of its tasks write one row (each row updated in task order: bit-reproducible factors, at the cost of more
 * launches), 0 (default) lets tasks of a level share rows through atomics, whose order of arrival sets the rounding
 * (for factorizations made afterwards).  "solve_dataflow": 1 runs the triangular solves with A as
 * two persistent launches whose items wait for their inputs through per-block counters, 0 (default) as one launch per
 * dependency level and step.  "solve_noalloc_limit_bytes": the bytes fplus_amd_solve_device_noalloc may use of its
 * array (0: all of it; for tests of the limit).  Unknown names are FPLUS_ERR_BADARG. */
int fplus_amd_tune(const char *name, int64_t value);

/* Per-kernel timing with HIP events on the library's stream (benchmarks): on = 1 starts recording and clears the sums,
 * on = 0 stops.  fplus_amd_profile_get returns, for a kernel name, the summed milliseconds, the launches and the summed
 * algorithmic work (flops, or bytes for the streaming kernels) of the launches recorded so far. */
int fplus_amd_profile(int on);
int fplus_amd_profile_get(const char *kernel, double *ms, int64_t *launches, double *work);

#ifdef __cplusplus
}
#endif
#endif /* FPLUS_AMD_H */
