"""ctypes bindings of libfplus_amd.so, the FPLUS C interface of the MI355X build (include/fplus.h).

The library is built in-tree (``make -C fplus_amd`` or ``__graft_entry__.build()``); importing the bindings without it
raises: there is no Python or CPU fallback of the numeric path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfplus_amd.so")

FPLUS_OK = 0
FPLUS_ERR_NOMEM = -1
FPLUS_ERR_SINGULAR = -2
FPLUS_ERR_BADARG = -3
FPLUS_ERR_GPU = -4
FPLUS_ERR_INTERNAL = -5
FPLUS_PIVOT_FULL, FPLUS_PIVOT_PARTIAL, FPLUS_PIVOT_ROOK = 0, 1, 2
FPLUS_PIVOT_REL_DEFAULT = 1.4901161193847656e-08
FPLUS_ORDER_COLAMD, FPLUS_ORDER_AMD_ATA, FPLUS_ORDER_AUTO = 0, 1, 2

I64P = C.POINTER(C.c_int64)
DP = C.POINTER(C.c_double)
FP = C.POINTER(C.c_float)
VP = C.c_void_p


class Factors(C.Structure):
    """fplus_factors (include/fplus.h)"""
    _fields_ = [("n", C.c_int64), ("is_complex", C.c_int), ("csc", C.c_int), ("Lp", I64P), ("Li", I64P), ("Lx", VP),
                ("Up", I64P), ("Ui", I64P), ("Ux", VP), ("p", I64P), ("q", I64P), ("nsupernodes", C.c_int64),
                ("n_pivots_replaced", C.c_int64)]


class Options(C.Structure):
    """fplus_options (include/fplus.h), 14 fields in the reference's order"""
    _fields_ = [("pin_dense", C.c_int), ("tau_density", C.c_double), ("tau_min_nnz", C.c_int64), ("postorder", C.c_int),
                ("equilibrate", C.c_int), ("ruiz_max_iter", C.c_int), ("ruiz_tol", C.c_double), ("pivot_rel", C.c_double),
                ("pivot_strategy", C.c_int), ("drop_tol", C.c_double), ("relax_rows", C.c_int64), ("panel", C.c_int64),
                ("use_gpu", C.c_int), ("gpu_min_flops", C.c_double), ("verbose", C.c_int)]


class Stats(C.Structure):
    """fplus_amd_stats (include/fplus_amd.h)"""
    _fields_ = [("nnzL", C.c_int64), ("nnzU", C.c_int64), ("nsupernodes", C.c_int64), ("nblocks", C.c_int64),
                ("panels", C.c_int64), ("levels", C.c_int64), ("flops", C.c_double), ("gemm_flops", C.c_double),
                ("t_ruiz", C.c_double), ("t_factor", C.c_double), ("t_host", C.c_double), ("t_wait", C.c_double),
                ("t_total", C.c_double), ("ruiz_iters", C.c_int), ("gemm_bytes", C.c_double),
                ("shared_levels", C.c_int64), ("max_block_w", C.c_int64), ("coop_retried", C.c_int), ("precision", C.c_int),
                ("refactor_replaced", C.c_int64), ("refactor_min_ratio", C.c_double), ("replay_bytes", C.c_double),
                ("f32_blocks", C.c_int64), ("f32_flop_share", C.c_double), ("pivot_ties", C.c_int64)]


class RowInfo(C.Structure):
    """fplus_amd_row_info (include/fplus_amd.h)"""
    _fields_ = [("runs", C.c_int64), ("restarts", C.c_int64), ("merges", C.c_int64), ("pool_entries", C.c_int64),
                ("multipliers", C.c_int64), ("t_device", C.c_double), ("t_total", C.c_double), ("merge_bytes", C.c_double)]


class AmdOptions(C.Structure):
    """fplus_amd_options (include/fplus_amd.h)"""
    _fields_ = [("precision", C.c_int), ("refactor", C.c_int), ("static_pivots", C.c_int), ("reserved", C.c_int),
                ("selector", C.c_void_p)]


class RefineOptions(C.Structure):
    """fplus_amd_refine_options (include/fplus_amd.h)"""
    _fields_ = [("max_outer", C.c_int), ("restart", C.c_int), ("tol", C.c_double), ("inner_tol", C.c_double),
                ("verbose", C.c_int), ("reserved", C.c_int)]


class RefineInfo(C.Structure):
    """fplus_amd_refine_info (include/fplus_amd.h)"""
    _fields_ = [("outer", C.c_int), ("inner", C.c_int), ("converged", C.c_int), ("reserved", C.c_int),
                ("bwd_first", C.c_double), ("bwd_last", C.c_double)]


# every entry point include/fplus.h and include/fplus_amd.h declare, with its ctypes signature
SIGNATURES = {
    "fplus_factors_free": (None, [C.POINTER(Factors)]),
    "fplus_factor_sn": (C.c_int, [C.c_int64, I64P, I64P, VP, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int, C.c_double,
                                  C.c_double, C.POINTER(Factors), C.c_char_p, C.c_size_t]),
    "fplus_factor_row": (C.c_int, [C.c_int64, I64P, I64P, VP, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double,
                                   C.POINTER(Factors), C.c_char_p, C.c_size_t]),
    "fplus_ruiz": (C.c_int, [C.c_int64, I64P, I64P, VP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, DP, DP, C.c_char_p,
                             C.c_size_t]),
    "fplus_tri_solve": (C.c_int, [C.c_int64, I64P, I64P, VP, I64P, I64P, VP, VP, C.c_int, C.c_int]),
    "fplus_solve_plan_create": (C.c_int, [C.c_int64, I64P, I64P, I64P, I64P, C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "fplus_tri_solve_plan": (C.c_int, [VP, C.c_int64, I64P, I64P, VP, I64P, I64P, VP, VP, C.c_int64, C.c_int64, C.c_int,
                                       C.c_int]),
    "fplus_solve_plan_nsupernodes": (C.c_int64, [VP]),
    "fplus_solve_plan_free": (None, [VP]),
    "fplus_coletree_postorder": (C.c_int, [C.c_int64, C.c_int64, I64P, I64P, I64P]),
    "fplus_analyze": (C.c_int, [C.c_int64, I64P, I64P, C.c_int, C.c_double, C.c_int64, VP, VP, C.c_int, C.POINTER(VP),
                                C.c_char_p, C.c_size_t]),
    "fplus_analyze_ord": (C.c_int, [C.c_int64, I64P, I64P, C.c_int, C.c_double, C.c_int64, VP, VP, C.c_int, C.c_int,
                                    C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "fplus_symbolic_free": (None, [VP]),
    "fplus_symbolic_n": (C.c_int64, [VP]),
    "fplus_symbolic_nnz": (C.c_int64, [VP]),
    "fplus_symbolic_row_perm": (I64P, [VP]),
    "fplus_symbolic_col_perm": (I64P, [VP]),
    "fplus_symbolic_n_dense_rows": (C.c_int64, [VP]),
    "fplus_symbolic_n_dense_cols": (C.c_int64, [VP]),
    "fplus_factor_sn_perm": (C.c_int, [C.c_int64, I64P, I64P, VP, C.c_int, I64P, I64P, C.c_int, C.c_int, C.c_double, C.c_int,
                                       C.c_int64, C.c_int64, C.c_int, C.c_double, C.c_double, C.POINTER(Factors), DP, DP,
                                       C.c_char_p, C.c_size_t]),
    "fplus_factor_sn_sym": (C.c_int, [VP, I64P, I64P, VP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int64,
                                      C.c_int64, C.c_int, C.c_double, C.c_double, C.POINTER(Factors), DP, DP, C.c_char_p,
                                      C.c_size_t]),
    "fplus_factor_row_perm": (C.c_int, [C.c_int64, I64P, I64P, VP, C.c_int, I64P, I64P, C.c_int, C.c_int, C.c_double,
                                        C.c_int, C.c_double, C.c_int, C.c_double, C.POINTER(Factors), DP, DP, C.c_char_p,
                                        C.c_size_t]),
    "fplus_factor_row_sym": (C.c_int, [VP, I64P, I64P, VP, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double,
                                       C.c_int, C.c_double, C.POINTER(Factors), DP, DP, C.c_char_p, C.c_size_t]),
    "fplus_default_options": (None, [C.POINTER(Options)]),
    "fplus_analyze_opts": (C.c_int, [C.c_int64, I64P, I64P, C.POINTER(Options), C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "fplus_factor": (C.c_int, [VP, I64P, I64P, VP, C.c_int, C.POINTER(Options), C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "fplus_solve": (C.c_int, [VP, VP, C.c_int64, C.c_int64, C.c_int, C.c_char_p, C.c_size_t]),
    "fplus_numeric_factors": (C.POINTER(Factors), [VP]),
    "fplus_numeric_p": (I64P, [VP]),
    "fplus_numeric_q": (I64P, [VP]),
    "fplus_numeric_Dr": (DP, [VP]),
    "fplus_numeric_Dc": (DP, [VP]),
    "fplus_numeric_free": (None, [VP]),
    "fplus_amd_symbolic_from_orders": (C.c_int, [C.c_int64, I64P, I64P, I64P, I64P, C.POINTER(VP), C.c_char_p, C.c_size_t]),
    "fplus_amd_numeric_stats": (C.c_int, [VP, VP]),
    "fplus_amd_factor_device": (C.c_int, [VP, I64P, I64P, VP, C.c_int, C.POINTER(Options), C.POINTER(VP), C.c_char_p,
                                          C.c_size_t]),
    "fplus_amd_solve_device": (C.c_int, [VP, VP, C.c_int64, C.c_int64, C.c_int, C.c_char_p, C.c_size_t]),
    "fplus_amd_solve_device_noalloc": (C.c_int, [VP, VP, C.c_int64, C.c_int64, C.c_int, C.c_char_p, C.c_size_t]),
    "fplus_amd_dev_alloc": (C.c_int, [C.c_size_t, C.POINTER(VP)]),
    "fplus_amd_dev_free": (C.c_int, [VP]),
    "fplus_amd_dev_copy": (C.c_int, [VP, VP, C.c_size_t, C.c_int]),
    "fplus_amd_dev_sync": (C.c_int, []),
    "fplus_amd_thread_stream": (C.c_int, [C.c_int]),
    "fplus_amd_row_stats": (C.c_int, [VP]),
    "fplus_amd_factor_row_batch": (C.c_int, [C.c_int64, I64P, VP, VP, VP, C.c_int, C.c_double, C.c_int, C.c_double, VP,
                                             C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "fplus_amd_profile": (C.c_int, [C.c_int]),
    "fplus_amd_profile_get": (C.c_int, [C.c_char_p, DP, I64P, DP]),
    "fplus_amd_tune": (C.c_int, [C.c_char_p, C.c_int64]),
    "fplus_amd_default_options": (None, [VP]),
    "fplus_amd_factor_ex": (C.c_int, [VP, I64P, I64P, VP, C.c_int, C.c_int, C.POINTER(Options), VP, C.POINTER(VP), C.c_char_p,
                                      C.c_size_t]),
    "fplus_amd_refactor": (C.c_int, [VP, VP, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_size_t]),
    "fplus_amd_default_refine_options": (None, [VP]),
    "fplus_amd_solve_refined": (C.c_int, [VP, VP, C.c_int, C.c_int64, C.c_int64, VP, VP, C.c_char_p, C.c_size_t]),
    "fplus_amd_refactor_solve": (C.c_int, [VP, VP, C.c_int, VP, C.c_int, C.c_int64, C.c_int64, C.c_int, VP, VP,
                                           C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "fplus_amd_refactor_sweep": (C.c_int, [VP, VP, C.c_int64, VP, C.c_int64, C.c_int64, C.c_int, C.c_int, VP, VP,
                                           C.POINTER(C.c_int), C.c_char_p, C.c_size_t]),
    "fplus_amd_selector_create": (C.c_int, [C.c_uint64, C.POINTER(VP)]),
    "fplus_amd_selector_free": (None, [VP]),
    "fplus_amd_selector_train": (C.c_int, [VP, FP, FP, C.c_int64, C.c_int, C.c_float, C.c_uint64, FP]),
    "fplus_amd_selector_predict": (C.c_int, [VP, FP, C.c_int64, FP]),
    "fplus_amd_selector_set_threshold": (C.c_int, [VP, C.c_float]),
    "fplus_amd_selector_get_weights": (C.c_int, [VP, FP, C.c_int64]),
    "fplus_amd_selector_set_weights": (C.c_int, [VP, FP, C.c_int64]),
    "fplus_amd_selector_log": (C.c_int, [VP, FP, DP, DP, C.c_int64, C.POINTER(C.c_int64)]),
}

_lib = None


def lib():
    """the loaded library with every signature set; RuntimeError when it is not built"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C fplus_amd` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class FplusError(RuntimeError):
    def __init__(self, code, msg):
        names = {FPLUS_ERR_NOMEM: "NOMEM", FPLUS_ERR_SINGULAR: "SINGULAR", FPLUS_ERR_BADARG: "BADARG",
                 FPLUS_ERR_GPU: "GPU", FPLUS_ERR_INTERNAL: "INTERNAL"}
        super().__init__(f"FPLUS_ERR_{names.get(code, code)}: {msg}")
        self.code = code


def check(rc, err):
    if rc != FPLUS_OK:
        raise FplusError(rc, err.value.decode(errors="replace") if err is not None else "")
