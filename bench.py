"""Benchmark: factor + solve seconds per system at backward error <= 1e-13 -- learned mixed precision vs the plain FP64
GPU path vs the CPU.

Workload: the complex shift-invert pencil A = L - sigma M, 64 modes x 4 fields x 1024 coefficients (n = 262144),
sigma = 0.3 + 0.7i, row and column scales 10^U(-8, 8), 16 right-hand sides N(0,1), values and right-hand sides
resident in HBM when a timed step starts.  The symbolic analysis (orders, supernode partition) is made once, before the
timed region.  Legs, each a whole system per step:

  fp64_gpu             the plain FP64 drop-in: the pivoting factorization (fplus_amd_factor_device) + the direct solve
                       of the 16 right-hand sides (fplus_amd_solve_device).
  fp64_refactor        fplus_amd_refactor_solve with FP64 updates: the factorization replayed in the pivot order of a
                       factorization of the pencil at the neighbouring shift sigma - 0.01i (made once, before the timed
                       region, as in a shift sweep), then the refined solve (which stops at once when the direct solve
                       already meets the target).
  fp32_refactor_ir     the same with every block update on the FP32 matrix cores, GMRES-IR to the target.
  learned_refactor_ir  the same with the learned selector deciding per block (fplus_amd/data/selector_shiftinvert.npz,
                       trained by fplus_amd/train_selector.py), GMRES-IR, and the FP64 refactorization as the fallback
                       when the refinement stalls.  This is `value`.

The target is the equilibrated normwise backward error ||Dr (b - A x)|| / (||Dr A Dc||_1 ||x / Dc|| + ||Dr b||) <= 1e-13 for
every right-hand side (the refined solves run to it); every leg's result is checked on the host with it, the normwise
one (||A x - b|| / (||A|| ||x|| + ||b||), which the 16 orders of scaling make vacuous) and the componentwise one.

The line also carries the kernels of one FP64 step timed with HIP events on the library's stream (an untimed step run
after the timed ones: events around the ~170k launches slow a step by about a tenth), the roofline of the dominant
kernel, and a CPU baseline: oracle/cpu_panel.c (the same numeric path, panel-blocked, OpenMP on the host threads, no
BLAS is installed) timed directly on a smaller member of the same family, median of 3 runs, scaled to the workload by
the factor-flops ratio (at most 50x).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--legs a,b,..] [--no-cpu] [--verbose]
    python bench.py --copy-probe      # 3 device-to-device copies of 1 GiB: the HBM counters' calibration run
"""
import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import fplus_amd  # noqa: E402
from fplus_amd import _lib, gen  # noqa: E402

WORKLOADS = {
    "shiftinvert-64x4x1024": dict(modes=64, nfields=4, ncoef=1024),
    "shiftinvert-16x4x1024": dict(modes=16, nfields=4, ncoef=1024),
    "shiftinvert-8x4x256": dict(modes=8, nfields=4, ncoef=256),
    # the precision configuration: 64^3 7-point convection-diffusion, Peclet 10, scales 10^U(-4, 4), real, 8 right-hand
    # sides, FP32 and BF16 hi/lo updates against FP64, GMRES-IR to 1e-14
    "convdiff3d-64": dict(g=64),
    # the learned selector's configuration: the shift-invert family with 17 orders of scaling, 10^U(-8, 9), n = 524288,
    # 64 right-hand sides
    "shiftinvert-128x4x1024-s17": dict(modes=128, nfields=4, ncoef=1024, scale_exp=8.0, scale_hi=9.0),
    # the batching configuration's sweep: one n = 131072 pattern refactored at the 32 shifts sigma_k = 0.3 + i k / 32
    # (k = 1 .. 32), 4 right-hand sides each, in the pivot order of one pivoting factorization at sigma = 0.3 + 0.5i
    "sweep-32x4x1024": dict(modes=32, nfields=4, ncoef=1024),
    # the batching configuration's independent systems: 256 members of the shift-invert family, each its own seed, with
    # mode counts drawn so that n is log-uniform in [4096, 65536], 4 right-hand sides each, factored and solved one
    # after another, and side by side on several host threads with a stream each
    "batch-256": dict(systems=256, nfields=4, ncoef=1024),
    # the row kernel (rook / full / partial pivoting, incomplete LU): the 96 x 96 convection-diffusion matrix (n = 9216)
    # in its analysed order, equilibrated, factored by fplus_factor_row once per leg
    "rowlu-convdiff2d-96": dict(g=96),
    # many independent systems through the row kernel: 256 convection-diffusion matrices on g x g grids, g log-uniform
    # in [32, 96] (n 1024 .. 9216), each its own seed, partial-pivoting ILU (drop_tol 1e-3)
    "rowlu-batch-256": dict(systems=256, gmin=32, gmax=96, strategy=1, drop_tol=1e-3),
}
# the row kernel's legs: (pivot strategy 0 full / 1 partial / 2 rook, drop_tol)
ROW_LEGS = {"row_partial_ilu": (1, 1e-3), "row_rook_ilu": (2, 1e-3), "row_full_ilu": (0, 1e-2), "row_partial": (1, 0.0)}
# the legs each workload runs (the default: the headline's four)
WORKLOAD_LEGS = {"convdiff3d-64": ("fp64_gpu", "fp32_factor_ir", "bf16_factor_ir"),
                 "sweep-32x4x1024": ("sweep", "sweep_graph"), "batch-256": ("batch", "batch_streams"),
                 "rowlu-convdiff2d-96": tuple(ROW_LEGS), "rowlu-batch-256": ("row_batch", "row_one_by_one")}
WORKLOAD_NRHS = {"convdiff3d-64": 8, "shiftinvert-128x4x1024-s17": 64, "sweep-32x4x1024": 4, "batch-256": 4}
WORKLOAD_TOL = {"convdiff3d-64": 1e-14}
DEFAULT_WORKLOAD = "shiftinvert-64x4x1024"   # n = 262144, the plain FP64 GPU drop-in configuration
SIGMA = 0.3 + 0.7j
SIGMA_PREV = 0.3 + 0.69j                     # the neighbouring shift whose pivot order the refactorizations replay
CPU_SAMPLE = dict(modes=8, nfields=4, ncoef=1024)   # n = 32768, the CPU baseline's member of the family
PEAK_FP64_MFMA = 78.6e12      # MI355X dense FP64 matrix, FLOP/s
PEAK_HBM = 8.0e12             # MI355X HBM3E, B/s
NRHS_DEFAULT = 16
TOL_DEFAULT = 1e-13
LEGS = ("fp64_gpu", "fp64_refactor", "fp32_refactor_ir", "learned_refactor_ir")
ALL_LEGS = (LEGS + ("fp32_factor_ir", "bf16_factor_ir", "sweep", "sweep_graph", "batch", "batch_streams") + tuple(ROW_LEGS)
            + ("row_batch", "row_one_by_one"))
SELECTOR = os.path.join(ROOT, "fplus_amd", "data", "selector_shiftinvert.npz")
# the dominant kernel's HBM bytes per launch, from the rocprofv3 FETCH_SIZE and WRITE_SIZE passes summarized there
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "blk_upd-hbm_traffic-shiftinvert-64x4x1024.txt")


def measured_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed counter summary, or None"""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    vals = {}
    for line in open(TRAFFIC_FILE):
        if "=" in line and not line.startswith(" "):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    if vals.get("workload") != workload or vals.get("kernel") != kernel:
        return None
    return float(vals["hbm_bytes_per_launch"])


def copy_probe():
    """three 1 GiB device-to-device copies through the library (the HBM counters' known-size reference)"""
    L = _lib.lib()
    nb = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    if L.fplus_amd_dev_alloc(nb, C.byref(src)) != 0 or L.fplus_amd_dev_alloc(nb, C.byref(dst)) != 0:
        raise RuntimeError("device allocation failed")
    for _ in range(3):
        if L.fplus_amd_dev_copy(dst, src, nb, 2) != 0:
            raise RuntimeError("copy failed")
    L.fplus_amd_dev_free(src)
    L.fplus_amd_dev_free(dst)
    print(json.dumps({"copy_probe": {"copies": 3, "bytes_each": nb}}))


def dev_upload(a):
    p = C.c_void_p()
    L = _lib.lib()
    if L.fplus_amd_dev_alloc(a.nbytes, C.byref(p)) != 0:
        raise RuntimeError("device allocation failed")
    if L.fplus_amd_dev_copy(p, C.c_void_p(a.ctypes.data), a.nbytes, 0) != 0:
        raise RuntimeError("upload failed")
    return p


def dev_download(p, a):
    if _lib.lib().fplus_amd_dev_copy(C.c_void_p(a.ctypes.data), p, a.nbytes, 1) != 0:
        raise RuntimeError("download failed")


def dev_copy(dst, src, nbytes):
    if _lib.lib().fplus_amd_dev_copy(dst, src, nbytes, 2) != 0:
        raise RuntimeError("device copy failed")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(workload_flops, runs=3):
    """oracle/cpu_panel.c on the family's CPU member, all host threads of the box's share (at most 16), median of `runs`"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n, Ap, Ai, Ax = gen.shift_invert(**CPU_SAMPLE, sigma=SIGMA)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    b = gen.rhs(n, NRHS_DEFAULT, True)
    threads = min(16, os.cpu_count() or 1)
    # the sample's factor flops as the GPU path counts them (the same structure as the workload's count; the blocked
    # CPU path's own count is a little lower: it leaves out rows its updates leave exactly zero)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    sample_flops = fplus_amd.stats(num)["flops"]
    del num
    times, info = [], None
    for r in range(runs):
        t0 = time.perf_counter()
        x, info = O.cpu_panel_factor_solve(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, b, panel=256, threads=threads)
        times.append(time.perf_counter() - t0)
        log(f"cpu baseline run {r + 1}/{runs}: {times[-1]:.2f} s")
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    comp = acc.componentwise(A, x, b)
    med = float(np.median(times))
    ratio = workload_flops / sample_flops
    model = platform.processor() or "x86_64"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    desc = ", ".join(f"{k}={v}" for k, v in CPU_SAMPLE.items())
    return {"value": med * ratio, "unit": "s/system", "cores": threads, "kind": "port",
            "sample": f"shift_invert({desc}), n={n}, sigma {SIGMA}: permutation + Ruiz + panel-blocked supernodal LU "
                      f"(oracle/cpu_panel.c, panel 256, relax 32, OpenMP on {threads} threads, own blocked loops: no BLAS "
                      f"installed) + {NRHS_DEFAULT} solves, median {med:.2f} s of {runs} runs ({', '.join(f'{t:.2f}' for t in times)}; "
                      f"factor {info['factor_s']:.2f} s at {info['flops'] / info['factor_s'] / 1e9:.1f} GFLOP/s, componentwise "
                      f"backward error {comp:.1e}), scaled by the factor-flops ratio {workload_flops:.3e}/{sample_flops:.3e} = "
                      f"{ratio:.1f}; CPU {model}, {os.cpu_count()} logical CPUs visible",
            "sample_s": med, "flops_ratio": ratio, "sample_flops": sample_flops, "sample_flops_cpu_count": info["flops"]}


FLOP_KERNELS = {"blk_upd", "blk_inverse", "blk_uinv", "gemm_sub", "getf2"}
BYTE_KERNELS = {"lfwd_gather", "lfwd_pull", "lfwd_push", "ub_dots", "ub_red", "ub_prod", "utdots", "ut_prod", "ltdots", "lt_prod",
                "ruiz_rowmax", "ruiz_scale_colmax", "ruiz_scale_cols"}
PROFILED = ("blk_upd", "blk_inverse", "blk_uinv", "gemm_sub", "getf2", "trsm_sub", "lfwd_gather", "lfwd_pull", "lfwd_push", "ub_dots",
            "ub_red", "ub_prod", "utdots", "ut_prod", "ltdots", "lt_prod", "ruiz_rowmax", "ruiz_scale_colmax", "ruiz_scale_cols")


def kernel_rate(name, sec, launches, work):
    """seconds and launches of one kernel, with its rate: TFLOP/s (and the fraction of the FP64 matrix peak) for the
    factor's arithmetic kernels, GB/s of algorithmic bytes (and the fraction of HBM peak) for Ruiz and the solves"""
    r = {"s": sec, "launches": launches}
    if work > 0 and sec > 0:
        if name in FLOP_KERNELS:
            r.update({"TFLOP/s": work / sec / 1e12, "frac_fp64_mfma": work / sec / PEAK_FP64_MFMA})
        elif name in BYTE_KERNELS:
            r.update({"GB/s": work / sec / 1e9, "frac_hbm": work / sec / PEAK_HBM})
    return r


def sweep_main(args, kw, legs, nrhs):
    """the shift sweep: one pivoting factorization (untimed set-up), then the 32 shifts refactored and solved in one
    fplus_amd_refactor_sweep call from device-resident values and right-hand sides, launched kernel by kernel ("sweep")
    and through the captured graph ("sweep_graph"); seconds per shift, and per system with the set-up shared out"""
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    shifts = [0.3 + 1j * k / 32 for k in range(1, 33)]
    n, Ap, Ai, Ax0 = gen.shift_invert(**kw, sigma=0.3 + 0.5j)
    # the pencil's values are L - sigma M: the shifts' values from L and M (one generation at sigma 0 and one at 1)
    _, _, _, AxL = gen.shift_invert(**kw, sigma=0.0)
    _, _, _, AxL1 = gen.shift_invert(**kw, sigma=1.0)
    AxM = AxL - AxL1
    axs = np.ascontiguousarray(np.stack([AxL - sg * AxM for sg in shifts]))
    nnz = int(Ap[-1])
    b = np.stack([gen.rhs(n, nrhs, True, seed=100 + k).T for k in range(len(shifts))])   # shift, column, row
    b = np.ascontiguousarray(b)
    log(f"sweep: n {n}, nnz {nnz}, {len(shifts)} shifts, {nrhs} right-hand sides each")
    L = _lib.lib()
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    opts = fplus_amd.default_options()
    err = C.create_string_buffer(512)
    P64 = _lib.I64P
    d_ax0 = dev_upload(np.ascontiguousarray(Ax0))
    d_axs = dev_upload(axs)
    d_b0 = dev_upload(b)
    d_b = dev_upload(b)
    aopt = fplus_amd.amd_options(refactor=1)
    num = C.c_void_p()
    t0 = time.perf_counter()
    _lib.check(L.fplus_amd_factor_ex(sym.ptr, Ap.ctypes.data_as(P64), Ai.ctypes.data_as(P64), d_ax0, 1, 1, C.byref(opts),
                                     C.byref(aopt), C.byref(num), err, 512), err)
    L.fplus_amd_dev_sync()
    setup_s = time.perf_counter() - t0
    ro = _lib.RefineOptions()
    L.fplus_amd_default_refine_options(C.byref(ro))
    ro.tol = 1e-13
    ro.inner_tol = 1e-2
    ns = len(shifts)
    out = {}
    for leg in legs:
        fplus_amd.tune("refactor_graph", 1 if leg == "sweep_graph" else 0)
        info = (_lib.RefineInfo * ns)()
        fb = (C.c_int * ns)()
        times = []
        CH = 8   # shifts per call: a progress line between calls
        vb = nrhs * n * 16        # bytes of one shift's right-hand sides
        va = nnz * 16             # bytes of one shift's values
        for rep in range(args.warmup + args.steps):
            dev_copy(d_b, d_b0, b.nbytes)
            L.fplus_amd_dev_sync()
            t1 = time.perf_counter()
            for k0 in range(0, ns, CH):
                kc = min(CH, ns - k0)
                _lib.check(L.fplus_amd_refactor_sweep(num, C.c_void_p(d_axs.value + k0 * va), kc, C.c_void_p(d_b.value + k0 * vb),
                                                      nrhs, n, 1, 0, C.byref(ro),
                                                      C.c_void_p(C.addressof(info) + k0 * C.sizeof(_lib.RefineInfo)),
                                                      C.cast(C.addressof(fb) + k0 * C.sizeof(C.c_int), C.POINTER(C.c_int)),
                                                      err, 512), err)
                L.fplus_amd_dev_sync()
                log(f"{leg} run {rep + 1}: shifts {k0 + 1}-{k0 + kc} done, {time.perf_counter() - t1:.1f} s")
            if rep >= args.warmup:
                times.append(time.perf_counter() - t1)
        x = np.empty_like(b)
        dev_download(d_b, x)
        Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(num), (n,)).copy()
        Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(num), (n,)).copy()
        A = sp.csc_matrix((axs[-1], Ai, Ap), shape=(n, n))
        meas = acc.all_measures(A, x[-1].T, b[-1].T, Dr, Dc)   # the last shift (Dr, Dc are the last refactorization's)
        sec = float(np.median(times))
        out[leg] = {"sweep_s": sec, "s_per_shift": sec / ns, "s_per_system_with_setup": (sec + setup_s) / ns,
                    "steps_s": times, "outer_max": max(i.outer for i in info), "converged": all(i.converged for i in info),
                    "fallbacks": int(sum(fb)), "last_shift_backward_error": meas}
        log(f"{leg}: {sec:.3f} s for {ns} shifts ({sec / ns:.3f} s/shift), last shift {meas}")
    fplus_amd.tune("refactor_graph", 0)
    L.fplus_numeric_free(num)
    for p in (d_ax0, d_axs, d_b0, d_b):
        L.fplus_amd_dev_free(p)
    head = min(out, key=lambda k: out[k]["sweep_s"])   # the faster launch form
    res = {"metric": "factor+solve s/system at bwd err <=1e-13: mixed/learned vs FP64 GPU vs CPU",
           "value": out[head]["s_per_system_with_setup"], "unit": "s/system", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": out[head]["sweep_s"] * 1e3, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": args.workload, "matrix": "complex shift-invert pencil", **kw, "n": n, "nnz": nnz,
                      "shifts": "0.3 + i k / 32, k = 1..32", "nrhs_per_shift": nrhs, "value_leg": head,
                      "pivot_order_from": "one pivoting factorization at sigma 0.3 + 0.5i (set-up, shared out in value)"},
           "setup_s": setup_s, "legs": out, "roofline": None, "cpu_baseline": None}
    print(json.dumps(res))


def batch_main(args, kw, nrhs):
    """many independent systems: each one's symbolic analysis made before the timed region, then every system's
    factorization and solve from device-resident values and right-hand sides -- one after another (leg "batch"), and
    side by side on --batch-threads host threads, each with its own stream (fplus_amd_thread_stream) and a share of the
    host cores for its factorizations' pools (leg "batch_streams")"""
    import threading
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    rng = np.random.default_rng(2024)
    L = _lib.lib()
    opts = fplus_amd.default_options()
    P64 = _lib.I64P
    systems = []
    t0 = time.perf_counter()
    for i in range(kw["systems"]):
        modes = int(np.clip(np.rint(2.0 ** rng.uniform(0.0, 4.0)), 1, 16))
        n, Ap, Ai, Ax = gen.shift_invert(modes, kw["nfields"], kw["ncoef"], seed=5000 + i)
        sym = fplus_amd.Symbolic(n, Ap, Ai)
        b = np.asfortranarray(gen.rhs(n, nrhs, True, seed=6000 + i))
        systems.append((n, Ap, Ai, Ax, sym, b, dev_upload(np.ascontiguousarray(Ax)), dev_upload(b)))
        if (i + 1) % 32 == 0:
            log(f"batch: {i + 1} systems analysed, {time.perf_counter() - t0:.1f} s")
    prep_s = time.perf_counter() - t0
    ns = [x[0] for x in systems]
    legs = [x for x in (args.legs or "batch,batch_streams").split(",") if x]

    def run(nthreads):
        """every system factored and solved, on nthreads host threads taking systems in turn; (wall seconds, per-system
        seconds, singular systems, worst checked backward errors)"""
        times, singular, checks = [0.0] * len(systems), [], []
        nxt, mu = [0], threading.Lock()
        errs = []

        def worker():
            err = C.create_string_buffer(512)
            if nthreads > 1:
                _lib.check(L.fplus_amd_thread_stream(1), None)
            try:
                while True:
                    with mu:
                        i = nxt[0]
                        nxt[0] += 1
                    if i >= len(systems):
                        return
                    n, Ap, Ai, Ax, sym, b, d_ax, d_b = systems[i]
                    t1 = time.perf_counter()
                    num = C.c_void_p()
                    rc = L.fplus_amd_factor_device(sym.ptr, Ap.ctypes.data_as(P64), Ai.ctypes.data_as(P64), d_ax, 1, C.byref(opts),
                                                   C.byref(num), err, 512)
                    if rc == _lib.FPLUS_ERR_SINGULAR:   # a member whose pivot fails the threshold test: reported, as a caller would see it
                        times[i] = time.perf_counter() - t1
                        with mu:
                            singular.append(i)
                        continue
                    _lib.check(rc, err)
                    _lib.check(L.fplus_amd_solve_device(num, d_b, nrhs, n, 0, err, 512), err)
                    L.fplus_amd_dev_sync()
                    times[i] = time.perf_counter() - t1
                    if i % 16 == 0:   # every 16th system's solutions kept for the check (outside the per-system time)
                        x = np.empty_like(b)
                        dev_download(d_b, x)
                        Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(num), (n,)).copy()
                        Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(num), (n,)).copy()
                        with mu:
                            checks.append((i, x, Dr, Dc))
                    L.fplus_numeric_free(num)
            except Exception as e:   # noqa: BLE001 -- reported by the caller
                errs.append(e)
            finally:
                if nthreads > 1:
                    L.fplus_amd_thread_stream(0)

        for x in systems:   # the right-hand sides afresh (a run overwrites them with the solutions)
            if L.fplus_amd_dev_copy(x[7], C.c_void_p(x[5].ctypes.data), x[5].nbytes, 0) != 0:
                raise RuntimeError("upload failed")
        tune_threads = max(1, 15 // nthreads)
        fplus_amd.tune("host_threads", tune_threads if nthreads > 1 else 0)
        t_all = time.perf_counter()
        th = [threading.Thread(target=worker) for _ in range(nthreads)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        total = time.perf_counter() - t_all
        fplus_amd.tune("host_threads", 0)
        if errs:
            raise errs[0]
        worst = {}
        for i, x, Dr, Dc in checks:
            n, Ap, Ai, Ax, _, b = systems[i][:6]
            m = acc.all_measures(sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)), x, b, Dr, Dc)
            worst = max(worst, m, key=lambda d: d.get("equilibrated", 0.0)) if worst else m
        return total, times, sorted(singular), worst

    out = {}
    for leg in legs:
        nt = 1 if leg == "batch" else args.batch_threads
        total, times, singular, worst = run(nt)
        out[leg] = {"s_per_system": total / len(systems), "total_s": total, "host_threads": nt,
                    "per_system_s": {"min": min(times), "median": float(np.median(times)), "max": max(times)},
                    "singular_systems": singular, "backward_error_worst_checked": worst,
                    "meets_target": bool(worst and worst["equilibrated"] <= 1e-13)}
        log(f"{leg}: {nt} host threads, {total:.2f} s for {len(systems)} systems ({total / len(systems):.4f} s/system), "
            f"worst checked {worst}")
    for x in systems:
        L.fplus_amd_dev_free(x[6])
        L.fplus_amd_dev_free(x[7])
    head = "batch_streams" if "batch_streams" in out else legs[0]
    total = out[head]["total_s"]
    res = {"metric": "factor+solve s/system at bwd err <=1e-13: mixed/learned vs FP64 GPU vs CPU",
           "value": total / len(systems), "unit": "s/system", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": total * 1e3,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128", "data": "synthetic",
           "config": {"workload": args.workload, "matrix": "complex shift-invert pencil family, own seed per system", **kw,
                      "n_min": min(ns), "n_max": max(ns), "n_median": int(np.median(ns)), "nrhs": nrhs, "value_leg": head,
                      "batching": f"{args.batch_threads} host threads, each with its own stream" if head == "batch_streams"
                      else "none: one system after another"},
           "legs": out, "analysis_s": prep_s,
           "speedup_vs_one_after_another": (out["batch"]["total_s"] / out["batch_streams"]["total_s"]
                                            if "batch" in out and "batch_streams" in out else None),
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(res))


def rowlu_main(args, kw, legs):
    """the row kernel: the matrix in its analysed order (the symbolic orders), equilibrated by the device Ruiz and laid
    out by rows before the timed region; per leg, fplus_factor_row's seconds (the values uploaded, the elimination on
    the device, the factors back on the host), median of --steps after --warmup, and the device elimination's row
    merges and their algorithmic bytes; beside it the reference's own row kernel (src/factor_row.c built with OpenMP,
    oracle/_ref/libfplus_ref_row_omp.so) on the host's cores with the same input, and whether both factorizations agree
    (p, q, patterns; values)"""
    n, Rp, Ri, Rx = row_input(*gen.convdiff2d(**kw))
    log(f"workload rowlu: n {n}, nnz {len(Ri)}")
    threads = min(16, os.cpu_count() or 1)
    ref = None
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libfplus_ref_row_omp.so")
    if not args.no_cpu and os.path.exists(ref_path):
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        ref = C.CDLL(ref_path)
        ref.fplus_factor_row.argtypes = [C.c_int64, _lib.I64P, _lib.I64P, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.POINTER(_lib.Factors), C.c_char_p, C.c_size_t]
    out = {}
    for leg in legs:
        strategy, dt = ROW_LEGS[leg]
        for _ in range(args.warmup):
            fplus_amd.factor_row(n, Rp, Ri, Rx, dt, strategy)
        ts, st = [], None
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            d = fplus_amd.factor_row(n, Rp, Ri, Rx, dt, strategy)
            ts.append(time.perf_counter() - t0)
            st = fplus_amd.row_stats()
        g = float(np.median(ts))
        r = {"strategy": ["full", "partial", "rook"][strategy], "drop_tol": dt, "s": g, "device_s": st["t_device"],
             "runs": st["runs"], "nnz_LU": int(d["Lp"][-1] + d["Up"][-1]), "merges": st["merges"],
             "merge_GB": st["merge_bytes"] / 1e9, "merge_GB/s": st["merge_bytes"] / st["t_device"] / 1e9,
             "frac_hbm": st["merge_bytes"] / st["t_device"] / PEAK_HBM, "n_pivots_replaced": d["n_pivots_replaced"]}
        log(f"{leg}: {g:.3f} s (device {st['t_device']:.3f} s, {st['runs']} runs), nnz(L+U) {r['nnz_LU']}")
        if ref is not None:
            cts = []
            for _ in range(max(1, args.cpu_runs)):
                f = _lib.Factors()
                e = C.create_string_buffer(256)
                t0 = time.perf_counter()
                rc = ref.fplus_factor_row(n, Rp.ctypes.data_as(_lib.I64P), Ri.ctypes.data_as(_lib.I64P), Rx.ctypes.data, dt, 0,
                                          0, strategy, _lib.FPLUS_PIVOT_REL_DEFAULT, C.byref(f), e, 256)
                cts.append(time.perf_counter() - t0)
                if rc != 0:
                    raise SystemExit(f"reference row kernel: {rc} {e.value.decode()}")
                nL, nU = f.Lp[n], f.Up[n]
                same = (nL + nU == r["nnz_LU"] and np.array_equal(np.ctypeslib.as_array(f.p, (n,)), d["p"])
                        and np.array_equal(np.ctypeslib.as_array(f.q, (n,)), d["q"])
                        and np.array_equal(np.ctypeslib.as_array(f.Ui, (nU,)), d["Ui"])
                        and np.array_equal(np.ctypeslib.as_array(C.cast(f.Ux, _lib.DP), (nU,)), d["Ux"]))
                _lib.lib().fplus_factors_free(C.byref(f))   # the factors' arrays are plain malloc'ed blocks
            c = float(np.median(cts))
            r.update({"cpu_s": c, "speedup_vs_cpu": c / g, "same_factors_as_reference": bool(same)})
            log(f"{leg}: reference CPU {c:.3f} s ({threads} threads), identical factors: {same}")
        out[leg] = r
    head = out[legs[0]]
    line = {"metric": "row-kernel factorization seconds", "value": head["s"], "unit": "s/system", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["s"] * 1e3, "higher_is_better": False,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": args.workload, "n": n, "nnz": int(len(Ri)), "legs": legs}, "legs": out,
            "roofline": {"bound": "hbm", "achieved": head["merge_GB/s"], "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                         "frac": head["frac_hbm"], "traffic": None,
                         "note": "one workgroup per system: the bound is its chain of dependent loads per step, not HBM"}}
    if ref is not None:
        line["cpu_baseline"] = {"value": head["cpu_s"], "unit": "s/system", "cores": threads, "kind": "reference",
                                "sample": f"the same matrix and leg: the reference's src/factor_row.c built with OpenMP "
                                          f"(-O3 -march=x86-64-v3), OMP_NUM_THREADS={os.environ.get('OMP_NUM_THREADS')}, "
                                          f"median of {args.cpu_runs} runs"}
    print(json.dumps(line), flush=True)


def row_input(n, Ap, Ai, Ax):
    """the row kernel's input as its pipelines make it: the matrix in its analysed order, equilibrated (device Ruiz),
    laid out by rows"""
    import scipy.sparse as sp
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    B = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))[sym.row_perm][:, sym.col_perm].tocsc()
    B.sort_indices()
    Bp, Bi = B.indptr.astype(np.int64), B.indices.astype(np.int64)
    Bx, _, _ = fplus_amd.ruiz(n, Bp, Bi, B.data)
    R = sp.csc_matrix((Bx, Bi, Bp), shape=(n, n)).tocsr()
    R.sort_indices()
    return n, R.indptr.astype(np.int64), R.indices.astype(np.int64), np.ascontiguousarray(R.data)


def rowlu_batch_main(args, kw, legs):
    """many small systems through the row kernel: each system's input prepared before the timed region; leg row_batch
    factors all of them in one fplus_amd_factor_row_batch call (one workgroup per system in shared launches), leg
    row_one_by_one with one fplus_factor_row call per system; beside them the reference's own row kernel (OpenMP build)
    one system after another on the host's cores, and on as many host threads at once, one system per thread with
    OpenMP off inside; both device legs checked against the reference system by system (p, q, U pattern and values)"""
    import threading
    rng = np.random.default_rng(77)
    systems = []
    for i in range(kw["systems"]):
        g = int(np.rint(np.exp(rng.uniform(np.log(kw["gmin"]), np.log(kw["gmax"])))))
        systems.append(row_input(*gen.convdiff2d(g, seed=100 + i)))
    ns = [s[0] for s in systems]
    log(f"workload rowlu-batch: {len(systems)} systems, n {min(ns)} .. {max(ns)}, total {sum(ns)}")
    strategy, dt = kw["strategy"], kw["drop_tol"]
    out = {}
    res = {}
    for leg in legs:
        for _ in range(args.warmup):
            if leg == "row_batch":
                fplus_amd.factor_row_batch(systems, dt, strategy)
            else:
                for sm in systems:
                    fplus_amd.factor_row(*sm, dt, strategy)
        ts = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            if leg == "row_batch":
                got = fplus_amd.factor_row_batch(systems, dt, strategy)
                st = fplus_amd.row_stats()
            else:
                got = [fplus_amd.factor_row(*sm, dt, strategy) for sm in systems]
            ts.append(time.perf_counter() - t0)
        res[leg] = got
        g = float(np.median(ts))
        out[leg] = {"s": g, "systems_per_s": len(systems) / g}
        if leg == "row_batch":
            out[leg].update({"device_s": st["t_device"], "runs": st["runs"], "merges": st["merges"],
                             "merge_GB/s": st["merge_bytes"] / st["t_device"] / 1e9})
        log(f"{leg}: {g:.3f} s for {len(systems)} systems")
    threads = min(16, os.cpu_count() or 1)
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libfplus_ref_row_omp.so")
    cpu = None
    if not args.no_cpu and os.path.exists(ref_path):
        os.environ.setdefault("OMP_NUM_THREADS", str(threads))
        ref = C.CDLL(ref_path)
        ref.fplus_factor_row.argtypes = [C.c_int64, _lib.I64P, _lib.I64P, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                         C.c_double, C.POINTER(_lib.Factors), C.c_char_p, C.c_size_t]
        gomp = C.CDLL("libgomp.so.1")

        def one(sm, check):
            n, Rp, Ri, Rx = sm
            f = _lib.Factors()
            e = C.create_string_buffer(256)
            rc = ref.fplus_factor_row(n, Rp.ctypes.data_as(_lib.I64P), Ri.ctypes.data_as(_lib.I64P), Rx.ctypes.data, dt, 0, 0,
                                      strategy, _lib.FPLUS_PIVOT_REL_DEFAULT, C.byref(f), e, 256)
            same = None
            if rc == 0 and check is not None:
                nU = f.Up[n]
                same = all(not isinstance(d, Exception) and np.array_equal(np.ctypeslib.as_array(f.p, (n,)), d["p"])
                           and np.array_equal(np.ctypeslib.as_array(f.Ui, (nU,)), d["Ui"])
                           and np.array_equal(np.ctypeslib.as_array(C.cast(f.Ux, _lib.DP), (nU,)), d["Ux"]) for d in check)
            if rc == 0:
                _lib.lib().fplus_factors_free(C.byref(f))
            return rc, same

        t0 = time.perf_counter()
        checks = [one(sm, [res[l][i] for l in res]) for i, sm in enumerate(systems)]
        seq = time.perf_counter() - t0
        same = all(c[1] for c in checks)
        nxt, mu = [0], threading.Lock()

        def worker():
            gomp.omp_set_num_threads(1)
            while True:
                with mu:
                    i = nxt[0]
                    nxt[0] += 1
                if i >= len(systems):
                    return
                one(systems[i], None)

        t0 = time.perf_counter()
        ths = [threading.Thread(target=worker) for _ in range(threads)]
        for t in ths:
            t.start()
        for t in ths:
            t.join()
        par = time.perf_counter() - t0
        cpu = {"sequential_s": seq, "threads_s": par}
        log(f"reference CPU: {seq:.3f} s one after another (OpenMP {threads}), {par:.3f} s on {threads} threads one system "
            f"each; device factors identical: {same}")
        for leg in out:
            out[leg].update({"speedup_vs_cpu_sequential": seq / out[leg]["s"], "speedup_vs_cpu_threads": par / out[leg]["s"],
                             "same_factors_as_reference": bool(same)})
    head = out[legs[0]]
    line = {"metric": "row-kernel batch seconds", "value": head["s"], "unit": "s/batch", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": head["s"] * 1e3, "higher_is_better": False, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": args.workload, "systems": len(systems), "n_min": min(ns), "n_max": max(ns),
                       "strategy": ["full", "partial", "rook"][strategy], "drop_tol": dt, "legs": legs}, "legs": out}
    if cpu is not None:
        line["cpu_baseline"] = {"value": min(cpu["sequential_s"], cpu["threads_s"]), "unit": "s/batch", "cores": threads,
                                "kind": "reference",
                                "sample": f"the same {len(systems)} systems through the reference's src/factor_row.c built with "
                                          f"OpenMP: {cpu['sequential_s']:.3f} s one after another with {threads} OpenMP threads, "
                                          f"{cpu['threads_s']:.3f} s on {threads} host threads one system each (value: the faster)"}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--legs", default=None, help="comma-separated legs (default: the workload's)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--max-outer", type=int, default=4, help="refinement steps at most in the refined legs")
    ap.add_argument("--verbose", action="store_true", help="the factorization's progress lines (for long profiled runs)")
    ap.add_argument("--copy-probe", action="store_true", help="only the HBM counters' calibration copies")
    ap.add_argument("--batch-threads", type=int, default=4, help="host threads of the batch workload's side-by-side leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the path would shard by "
              "independent systems, or by elimination subtrees with separator updates exchanged between GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    legs = [x for x in (args.legs or ",".join(WORKLOAD_LEGS.get(args.workload, LEGS))).split(",") if x]
    for x in legs:
        if x not in ALL_LEGS:
            raise SystemExit(f"unknown leg {x}")
    kw = WORKLOADS[args.workload]
    NRHS = WORKLOAD_NRHS.get(args.workload, NRHS_DEFAULT)
    TOL = WORKLOAD_TOL.get(args.workload, TOL_DEFAULT)
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    if args.workload.startswith("sweep"):
        return sweep_main(args, kw, legs, NRHS)
    if args.workload.startswith("batch"):
        return batch_main(args, kw, NRHS)
    if args.workload.startswith("rowlu-batch"):
        return rowlu_batch_main(args, kw, legs)
    if args.workload.startswith("rowlu"):
        return rowlu_main(args, kw, legs)
    shift = args.workload.startswith("shiftinvert")
    if shift:
        n, Ap, Ai, Ax = gen.shift_invert(**kw, sigma=SIGMA)
        _, _, _, Ax_prev = gen.shift_invert(**kw, sigma=SIGMA_PREV)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(**kw)
        Ax_prev = Ax
    cplx = bool(np.iscomplexobj(Ax))
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    log(f"workload {args.workload}: n {n}, nnz {Ap[-1]}")
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    L = _lib.lib()
    opts = fplus_amd.default_options(verbose=2 if args.verbose else 0)
    d_Ax = dev_upload(np.ascontiguousarray(Ax))
    b = gen.rhs(n, NRHS, cplx)
    bf = np.asfortranarray(b)
    d_b0 = dev_upload(bf)              # the right-hand sides, kept
    d_b = dev_upload(bf)               # overwritten with x by each step
    err = C.create_string_buffer(512)
    P64 = _lib.I64P

    def reset_b():
        dev_copy(d_b, d_b0, bf.nbytes)
        L.fplus_amd_dev_sync()

    def check_x(Dr, Dc):
        x = np.empty_like(bf)
        dev_download(d_b, x)
        return acc.all_measures(A, x, b, Dr, Dc)

    out_legs = {}
    kern, st64, profiled_s = {}, None, None
    if "fp64_gpu" in legs:
        def step():
            num = C.c_void_p()
            _lib.check(L.fplus_amd_factor_device(sym.ptr, Ap.ctypes.data_as(P64), Ai.ctypes.data_as(P64), d_Ax, int(cplx),
                                                 C.byref(opts), C.byref(num), err, 512), err)
            _lib.check(L.fplus_amd_solve_device(num, d_b, NRHS, n, 0, err, 512), err)
            return num

        for _ in range(args.warmup):
            reset_b()
            L.fplus_numeric_free(step())
        times, num = [], None
        for _ in range(args.steps):
            if num is not None:
                L.fplus_numeric_free(num)
            reset_b()
            t0 = time.perf_counter()
            num = step()
            L.fplus_amd_dev_sync()
            times.append(time.perf_counter() - t0)
        Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(num), (n,)).copy()
        Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(num), (n,)).copy()
        meas = check_x(Dr, Dc)
        # the solve alone on this factorization, with one and with 16 right-hand sides
        solve_s = {}
        for k in (1, NRHS):
            reset_b()
            ts = time.perf_counter()
            _lib.check(L.fplus_amd_solve_device(num, d_b, k, n, 0, err, 512), err)
            L.fplus_amd_dev_sync()
            solve_s[f"nrhs{k}"] = time.perf_counter() - ts
        L.fplus_numeric_free(num)
        # one more step with HIP events around every launch: the kernels' times for the breakdown and the roofline
        reset_b()
        L.fplus_amd_profile(1)
        t2 = time.perf_counter()
        num = step()
        L.fplus_amd_dev_sync()
        profiled_s = time.perf_counter() - t2
        L.fplus_amd_profile(0)
        st64 = _lib.Stats()
        L.fplus_amd_numeric_stats(num, C.byref(st64))
        L.fplus_numeric_free(num)
        for name in PROFILED:
            ms, cnt, work = C.c_double(), C.c_int64(), C.c_double()
            L.fplus_amd_profile_get(name.encode(), C.byref(ms), C.byref(cnt), C.byref(work))
            if cnt.value:
                kern[name] = (ms.value / 1e3, cnt.value, work.value)
        out_legs["fp64_gpu"] = {"s_per_system": float(np.median(times)), "steps_s": times, "solve_s": solve_s,
                                "backward_error": meas, "meets_target": meas["equilibrated"] <= TOL}
        log(f"fp64_gpu: {out_legs['fp64_gpu']['s_per_system']:.3f} s/system, {meas}")

    # low-precision pivoting factorizations + GMRES-IR (the precision configuration)
    ro_ir = _lib.RefineOptions()
    L.fplus_amd_default_refine_options(C.byref(ro_ir))
    ro_ir.tol = TOL
    ro_ir.inner_tol = 1e-2
    ro_ir.max_outer = args.max_outer
    for leg, prec in (("fp32_factor_ir", 1), ("bf16_factor_ir", 4)):
        if leg not in legs:
            continue
        aopt = fplus_amd.amd_options(precision=prec)

        def step():
            num = C.c_void_p()
            _lib.check(L.fplus_amd_factor_ex(sym.ptr, Ap.ctypes.data_as(P64), Ai.ctypes.data_as(P64), d_Ax, 1, int(cplx),
                                             C.byref(opts), C.byref(aopt), C.byref(num), err, 512), err)
            info = _lib.RefineInfo()
            _lib.check(L.fplus_amd_solve_refined(num, d_b, 1, NRHS, n, C.byref(ro_ir), C.byref(info), err, 512), err)
            return num, info

        for _ in range(args.warmup):
            reset_b()
            L.fplus_numeric_free(step()[0])
        times, num = [], None
        for _ in range(args.steps):
            if num is not None:
                L.fplus_numeric_free(num)
            reset_b()
            t0 = time.perf_counter()
            num, info = step()
            L.fplus_amd_dev_sync()
            times.append(time.perf_counter() - t0)
        Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(num), (n,)).copy()
        Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(num), (n,)).copy()
        meas = check_x(Dr, Dc)
        st = _lib.Stats()
        L.fplus_amd_numeric_stats(num, C.byref(st))
        L.fplus_numeric_free(num)
        out_legs[leg] = {"s_per_system": float(np.median(times)), "steps_s": times, "backward_error": meas,
                         "meets_target": meas["equilibrated"] <= TOL, "factor_s": st.t_total,
                         "refine": {"outer": info.outer, "inner": info.inner, "converged": info.converged,
                                    "bwd_first": info.bwd_first, "bwd_last": info.bwd_last}}
        log(f"{leg}: {out_legs[leg]['s_per_system']:.3f} s/system, refine {out_legs[leg]['refine']}, {meas}")

    ref_legs = [x for x in legs if x in ("fp64_refactor", "fp32_refactor_ir", "learned_refactor_ir")]
    setup = None
    if ref_legs:
        # the pivot order the refactorizations replay: a factorization of the pencil at the neighbouring shift, made once
        sel = fplus_amd.Selector.load(SELECTOR)
        d_Ax_prev = dev_upload(np.ascontiguousarray(Ax_prev))
        aopt = fplus_amd.amd_options(refactor=1, selector=sel)
        num = C.c_void_p()
        t0 = time.perf_counter()
        _lib.check(L.fplus_amd_factor_ex(sym.ptr, Ap.ctypes.data_as(P64), Ai.ctypes.data_as(P64), d_Ax_prev, 1, int(cplx),
                                         C.byref(opts), C.byref(aopt), C.byref(num), err, 512), err)
        L.fplus_amd_dev_sync()
        setup = {"pivoting_factorization_with_record_s": time.perf_counter() - t0,
                 "values": f"sigma {SIGMA_PREV}" if shift else "the workload's own"}
        L.fplus_amd_dev_free(d_Ax_prev)
        st = _lib.Stats()
        L.fplus_amd_numeric_stats(num, C.byref(st))
        setup["replay_bytes"] = st.replay_bytes
        log(f"refactorization setup: {setup}")
        ro = _lib.RefineOptions()
        L.fplus_amd_default_refine_options(C.byref(ro))
        ro.tol = TOL
        ro.inner_tol = 1e-2     # GMRES-IR's inner solves need only modest accuracy: fewer triangular solves per step
        ro.max_outer = args.max_outer
        prec = {"fp64_refactor": 0, "fp32_refactor_ir": 1, "learned_refactor_ir": 2}
        for leg in ref_legs:
            def step():
                info = _lib.RefineInfo()
                fb = C.c_int(0)
                _lib.check(L.fplus_amd_refactor_solve(num, d_Ax, 1, d_b, 1, NRHS, n, prec[leg], C.byref(ro), C.byref(info),
                                                      C.byref(fb), err, 512), err)
                return info, fb.value

            for _ in range(args.warmup):
                reset_b()
                step()
            times = []
            for _ in range(args.steps):
                reset_b()
                t0 = time.perf_counter()
                info, fb = step()
                L.fplus_amd_dev_sync()
                times.append(time.perf_counter() - t0)
            Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(num), (n,)).copy()
            Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(num), (n,)).copy()
            meas = check_x(Dr, Dc)
            st = _lib.Stats()
            L.fplus_amd_numeric_stats(num, C.byref(st))
            r = {"s_per_system": float(np.median(times)), "steps_s": times, "backward_error": meas,
                 "meets_target": meas["equilibrated"] <= TOL,
                 "refine": {"outer": info.outer, "inner": info.inner, "converged": info.converged, "bwd_first": info.bwd_first,
                            "bwd_last": info.bwd_last},
                 "fallback_to_fp64": fb, "refactor_s": st.t_total, "pivots_replaced": st.refactor_replaced,
                 "min_pivot_ratio": st.refactor_min_ratio, "f32_blocks": st.f32_blocks, "blocks": st.nblocks,
                 "f32_update_flop_share": st.f32_flop_share}
            out_legs[leg] = r
            log(f"{leg}: {r['s_per_system']:.3f} s/system, refine {r['refine']}, fallback {fb}, {meas}")
        L.fplus_numeric_free(num)
        del sel

    roof = None
    if kern:
        dom = max(kern, key=lambda k: kern[k][0])
        t_dom, cnt_dom, work_dom = kern[dom]
        bound = "mfma" if dom in FLOP_KERNELS else "hbm"
        achieved = work_dom / t_dom
        peak = PEAK_FP64_MFMA if bound == "mfma" else PEAK_HBM
        roof = {"kernel": dom, "bound": bound, "achieved": achieved / (1e12 if bound == "mfma" else 1e9),
                "peak": peak / (1e12 if bound == "mfma" else 1e9), "unit": "TFLOP/s" if bound == "mfma" else "GB/s",
                "frac": achieved / peak, "traffic": measured_traffic(args.workload, dom), "launches": cnt_dom,
                "avg_launch_ms": 1e3 * t_dom / cnt_dom, "work_per_launch": work_dom / cnt_dom}
        if dom == "blk_upd":
            roof["algorithmic_bytes_per_launch"] = st64.gemm_bytes / cnt_dom
    cpu = None
    if not args.no_cpu and st64 is not None and args.workload == DEFAULT_WORKLOAD:
        cpu = cpu_baseline(st64.flops, args.cpu_runs)
    head = next((x for x in ("learned_refactor_ir", "bf16_factor_ir") if x in out_legs), legs[0] if legs else None)
    value = out_legs[head]["s_per_system"] if head else None
    base = out_legs.get("fp64_gpu", {}).get("s_per_system")
    out = {
        "metric": "factor+solve s/system at bwd err <=1e-13: mixed/learned vs FP64 GPU vs CPU",
        "value": value, "unit": "s/system", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": value * 1e3 if value else None, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if cplx else "f64", "data": "synthetic",
        "config": {"workload": args.workload,
                   "matrix": "complex shift-invert pencil" if shift else "3D 7-point convection-diffusion, Peclet 10",
                   **kw, "n": n, "nnz": int(Ap[-1]), "sigma": str(SIGMA) if shift else None,
                   "scales": "10^U(-8,8)" if shift else "10^U(-4,4)", "nrhs": NRHS, "value_leg": head,
                   "backward_error_target": TOL, "backward_error_measure": "equilibrated normwise, every right-hand side",
                   "ordering": "fplus_analyze defaults", "relax_rows": 32},
        "legs": out_legs,
        "speedup_vs_fp64_gpu": {k: base / v["s_per_system"] for k, v in out_legs.items()} if base else None,
        "refactor_setup": setup,
        "backward_error_max": max((v["backward_error"]["equilibrated"] for v in out_legs.values()), default=None),
        "profiled_step_s": profiled_s,
        "factor": None if st64 is None else {
            "flops": st64.flops, "gemm_flops": st64.gemm_flops, "gemm_bytes": st64.gemm_bytes, "nnzL": st64.nnzL,
            "nnzU": st64.nnzU, "supernodes": st64.nsupernodes, "panels": st64.panels, "levels": st64.levels,
            "t_ruiz": st64.t_ruiz, "t_factor": st64.t_factor, "t_host": st64.t_host,
            "gflops_fp64_gpu": st64.flops / out_legs["fp64_gpu"]["s_per_system"] / 1e9} if "fp64_gpu" in out_legs else None,
        "kernels": {k: kernel_rate(k, *v) for k, v in sorted(kern.items(), key=lambda kv: -kv[1][0])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    for p in (d_Ax, d_b0, d_b):
        L.fplus_amd_dev_free(p)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
