"""Benchmark of the FP64 GPU factor + solve path: seconds per system.

A step is one factorization of the workload's matrix from values already resident in HBM (the permutation into the
analysed order, Ruiz, the supernodal LU) plus one solve of 16 right-hand sides, also resident in HBM, through the C
interface (fplus_amd_factor_device, fplus_amd_solve_device).  The symbolic analysis is done once, before the timed
region (it is reused across refactorizations of one pattern).  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu] [--verbose]
    python bench.py --copy-probe      # 3 device-to-device copies of 1 GiB: the HBM counters' calibration run

The line carries the backward error of every right-hand side of the last step, the solve alone on that step's
factorization with 1 and with 16 right-hand sides (solve_s), the roofline of the dominant kernel, and a CPU baseline.  The kernels are timed with HIP events on the library's stream around every launch, in one more
step run right after the timed ones: events on the ~175k launches of a step slow it by about a tenth, so the timed
steps run without them (the profiled step's own time is in the line as profiled_step_s).  The CPU baseline is the reference's own
supernodal kernel (oracle/_ref, compiled from the reference's sources; no CBLAS is installed, so its loop form) or the
oracle, on one host thread, timed on a smaller member of the same matrix family and scaled to the workload by the
ratio of the factor flops.
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

# name: (generator, keyword arguments, description); "sample" is the CPU baseline's smaller member
WORKLOADS = {
    "shiftinvert-16x4x256": (gen.shift_invert, dict(modes=16, nfields=4, ncoef=256), "complex shift-invert pencil"),
    "shiftinvert-32x4x512": (gen.shift_invert, dict(modes=32, nfields=4, ncoef=512), "complex shift-invert pencil"),
    "shiftinvert-64x4x1024": (gen.shift_invert, dict(modes=64, nfields=4, ncoef=1024), "complex shift-invert pencil"),
    "convdiff2d-96": (gen.convdiff2d, dict(g=96), "2D convection-diffusion"),
    "convdiff3d-32": (gen.convdiff3d, dict(g=32), "3D convection-diffusion"),
}
CPU_SAMPLE = {
    "shiftinvert": (gen.shift_invert, dict(modes=8, nfields=4, ncoef=128)),
    "convdiff2d": (gen.convdiff2d, dict(g=96)),
    "convdiff3d": (gen.convdiff3d, dict(g=16)),
}
DEFAULT_WORKLOAD = "shiftinvert-64x4x1024"   # n = 262144, the plain FP64 GPU drop-in configuration
PEAK_FP64_MFMA = 78.6e12      # MI355X dense FP64 matrix, FLOP/s
PEAK_HBM = 8.0e12             # MI355X HBM3E, B/s
NRHS = 16
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


def factor_flops(n, Ap, Ai, Ax, sym):
    """c * sum_k l_k u_k of the structure (the same count for any device or host factorization of these orders)"""
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    return fplus_amd.stats(num)["flops"], num


def cpu_baseline(workload_flops, family):
    """the reference column kernel (or the oracle) on one thread, on the family's CPU sample, scaled by flops"""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    fn, kw = CPU_SAMPLE[family]
    n, Ap, Ai, Ax = fn(**kw)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    fl, _ = factor_flops(n, Ap, Ai, Ax, sym)
    Bp, Bi, Bx = O.permute(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm)
    kind = "reference" if O.ref_available() else "port"
    b = gen.rhs(n, NRHS, np.iscomplexobj(Ax))
    t0 = time.perf_counter()
    if kind == "reference":
        x, Dr, Dc = O.ref_ruiz(n, Bp, Bi, Bx)
        f = O.ref_factor_sn(n, Bp, Bi, x, relax_rows=32, panel=512)
        for r in range(NRHS):
            xs = np.ascontiguousarray(b[:, r])
            O.ref().fplus_tri_solve(n, O._p64(f["Lp"]), O._p64(f["Li"]), f["Lx"].ctypes.data, O._p64(f["Up"]), O._p64(f["Ui"]),
                                    f["Ux"].ctypes.data, xs.ctypes.data, 0, int(np.iscomplexobj(Ax)))
    else:
        num = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm)
        num.solve(b)
    t = time.perf_counter() - t0
    scaled = t * workload_flops / fl
    desc = ", ".join(f"{k}={v}" for k, v in kw.items())
    return {"value": scaled, "unit": "s/system", "cores": 1, "kind": kind,
            "sample": f"{family}({desc}), n={n}: factor (Ruiz + supernodal panel kernel, panel 512, relax 32, no BLAS) + "
                      f"{NRHS} solves in {t:.2f} s on one "
                      f"thread, scaled by the factor flops ratio {workload_flops:.3e}/{fl:.3e}; CPU {platform.processor() or 'x86_64'}, "
                      f"{os.cpu_count()} logical CPUs on the host"}


FLOP_KERNELS = {"blk_upd", "blk_inverse", "blk_uinv", "gemm_sub", "getf2"}
BYTE_KERNELS = {"lfwd_gather", "lfwd_pull", "lfwd_push", "ub_dots", "ub_red", "ub_prod", "utdots", "ut_prod", "ltdots", "lt_prod",
                "ruiz_rowmax", "ruiz_scale_colmax", "ruiz_scale_cols"}


def kernel_rate(name, sec, launches, work):
    """seconds and launches of one kernel over the timed region, with its rate: TFLOP/s (and the fraction of the
    FP64 matrix peak) for the factor's arithmetic kernels, GB/s of algorithmic bytes (and the fraction of HBM peak)
    for Ruiz and the solves"""
    r = {"s": sec, "launches": launches}
    if work > 0 and sec > 0:
        if name in FLOP_KERNELS:
            r.update({"TFLOP/s": work / sec / 1e12, "frac_fp64_mfma": work / sec / PEAK_FP64_MFMA})
        elif name in BYTE_KERNELS:
            r.update({"GB/s": work / sec / 1e9, "frac_hbm": work / sec / PEAK_HBM})
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default=DEFAULT_WORKLOAD, choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--verbose", action="store_true", help="the factorization's progress lines (for long profiled runs)")
    ap.add_argument("--copy-probe", action="store_true", help="only the HBM counters' calibration copies")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the path would shard by "
              "independent systems, or by elimination subtrees with separator updates exchanged between GPUs)", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    fn, kw, desc = WORKLOADS[args.workload]
    n, Ap, Ai, Ax = fn(**kw)
    cplx = np.iscomplexobj(Ax)
    vs = 2 if cplx else 1
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    L = _lib.lib()
    opts = fplus_amd.default_options(verbose=2 if args.verbose else 0)
    d_Ax = dev_upload(np.ascontiguousarray(Ax))
    b = gen.rhs(n, NRHS, cplx)
    bf = np.asfortranarray(b)
    d_b = dev_upload(bf)
    err = C.create_string_buffer(512)

    def step():
        num = C.c_void_p()
        rc = L.fplus_amd_factor_device(sym.ptr, Ap.ctypes.data_as(_lib.I64P), Ai.ctypes.data_as(_lib.I64P), d_Ax, int(cplx),
                                       C.byref(opts), C.byref(num), err, 512)
        _lib.check(rc, err)
        rc = L.fplus_amd_solve_device(num, d_b, NRHS, n, 0, err, 512)
        _lib.check(rc, err)
        return num

    for _ in range(args.warmup):
        num = step()
        L.fplus_numeric_free(num)
        L.fplus_amd_dev_copy(d_b, C.c_void_p(bf.ctypes.data), bf.nbytes, 0)
    L.fplus_amd_dev_sync()
    t0 = time.perf_counter()
    nums = []
    for _ in range(args.steps):
        if nums:   # a refactorization loop keeps one factorization: the previous one goes before the next is made
            L.fplus_numeric_free(nums.pop())
        nums.append(step())
    L.fplus_amd_dev_sync()
    t1 = time.perf_counter()
    sec = (t1 - t0) / args.steps
    # one more step with HIP events around every launch: the kernels' times for the breakdown and the roofline
    L.fplus_amd_dev_copy(d_b, C.c_void_p(bf.ctypes.data), bf.nbytes, 0)
    L.fplus_amd_profile(1)
    t2 = time.perf_counter()
    nums.append(step())
    L.fplus_amd_dev_sync()
    t3 = time.perf_counter()
    L.fplus_amd_profile(0)
    # the solve alone on the last factorization, with one and with 16 right-hand sides (wall clock, vectors in HBM)
    solve_s = {}
    for k in (1, NRHS):
        L.fplus_amd_dev_copy(d_b, C.c_void_p(bf.ctypes.data), bf.nbytes, 0)
        L.fplus_amd_dev_sync()
        ts = time.perf_counter()
        _lib.check(L.fplus_amd_solve_device(nums[-1], d_b, k, n, 0, err, 512), err)
        L.fplus_amd_dev_sync()
        solve_s[f"nrhs{k}"] = time.perf_counter() - ts
    # correctness of the last step: solve the 16 right-hand sides again from b and check the backward errors
    L.fplus_amd_dev_copy(d_b, C.c_void_p(bf.ctypes.data), bf.nbytes, 0)
    _lib.check(L.fplus_amd_solve_device(nums[-1], d_b, NRHS, n, 0, err, 512), err)
    x = np.empty_like(bf)
    dev_download(d_b, x)
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    Dr = np.ctypeslib.as_array(L.fplus_numeric_Dr(nums[-1]), (n,)).copy()
    Dc = np.ctypeslib.as_array(L.fplus_numeric_Dc(nums[-1]), (n,)).copy()
    bwd = acc.all_measures(A, x, b, Dr, Dc)
    st = _lib.Stats()
    L.fplus_amd_numeric_stats(nums[-1], C.byref(st))
    for nm in nums:
        L.fplus_numeric_free(nm)
    # the kernels' sums over the timed region
    kern = {}
    for name in ("blk_upd", "blk_inverse", "blk_uinv", "gemm_sub", "getf2", "trsm_sub", "lfwd_gather", "lfwd_pull", "lfwd_push", "ub_dots", "ub_red", "ub_prod", "utdots", "ut_prod", "ltdots",
                 "lt_prod", "ruiz_rowmax", "ruiz_scale_colmax", "ruiz_scale_cols"):
        ms, cnt, work = C.c_double(), C.c_int64(), C.c_double()
        L.fplus_amd_profile_get(name.encode(), C.byref(ms), C.byref(cnt), C.byref(work))
        if cnt.value:
            kern[name] = (ms.value / 1e3, cnt.value, work.value)
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
        roof["algorithmic_bytes_per_launch"] = st.gemm_bytes / cnt_dom
    family = args.workload.split("-")[0]
    cpu = None if args.no_cpu else cpu_baseline(st.flops, family)
    out = {
        "metric": "factor+solve s/system at bwd err <=1e-13: mixed/learned vs FP64 GPU vs CPU",
        "value": sec, "unit": "s/system", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sec * 1e3, "profiled_step_s": t3 - t2, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "c128" if cplx else "f64", "data": "synthetic",
        "config": {"workload": args.workload, "matrix": desc, **kw, "n": n, "nnz": int(Ap[-1]), "nrhs": NRHS,
                   "precision": "fp64 (plain GPU drop-in)", "ordering": "fplus_analyze defaults", "relax_rows": 32},
        # the headline's stated measure (the reference tests' normwise one) and the two scale-invariant ones, each the
        # maximum over the right-hand sides of the last step
        "backward_error_max": bwd["normwise"],
        "backward_error": bwd,
        "solve_s": solve_s,
        "factor": {"flops": st.flops, "gemm_flops": st.gemm_flops, "gemm_bytes": st.gemm_bytes, "nnzL": st.nnzL, "nnzU": st.nnzU,
                   "supernodes": st.nsupernodes, "panels": st.panels, "levels": st.levels,
                   "t_ruiz": st.t_ruiz, "t_factor": st.t_factor, "t_host": st.t_host, "gflops": st.flops / sec / 1e9},
        "kernels": {k: kernel_rate(k, *v) for k, v in sorted(kern.items(), key=lambda kv: -kv[1][0])},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
