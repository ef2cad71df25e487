"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle (liboracle.so, built by oracle/Makefile) and, when it was
built, of the reference's own numeric sources (_ref/libfplus_ref.so, built by oracle/ref.mk).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module, as the checker.  The product
(fplus_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)


def _p64(a):
    return a.ctypes.data_as(_I64P)


def _pd(a):
    return a.ctypes.data_as(_DP)


def _f64view(x):
    """float64 view of a float64 / complex128 contiguous array"""
    x = np.ascontiguousarray(x)
    return x.view(np.float64) if np.iscomplexobj(x) else x


class OrcLU(C.Structure):
    _fields_ = [("n", C.c_int64), ("cplx", C.c_int32), ("pad", C.c_int32), ("nsn", C.c_int64),
                ("Lp", _I64P), ("Li", _I64P), ("Lx", _DP), ("Up", _I64P), ("Ui", _I64P), ("Ux", _DP),
                ("p", _I64P), ("sn_first", _I64P)]


class OrcNum(C.Structure):
    _fields_ = [("lu", C.POINTER(OrcLU)), ("p", _I64P), ("q", _I64P), ("Dr", _DP), ("Dc", _DP), ("ruiz_iters", C.c_int)]


class OrcRowLU(C.Structure):
    _fields_ = [("n", C.c_int64), ("cplx", C.c_int32), ("pad", C.c_int32), ("Lp", _I64P), ("Li", _I64P), ("Lx", _DP),
                ("Up", _I64P), ("Ui", _I64P), ("Ux", _DP), ("p", _I64P), ("q", _I64P), ("n_replaced", C.c_int64),
                ("fail_step", C.c_int64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.orc_permute.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, _I64P, _I64P, _I64P, _I64P, _DP]
        L.orc_ruiz.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, C.c_int, C.c_int, C.c_double, _DP, _DP,
                               C.POINTER(C.c_int), C.c_char_p, C.c_size_t]
        L.orc_factor_sn.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, C.c_int64, C.c_double,
                                    C.POINTER(C.POINTER(OrcLU)), C.c_char_p, C.c_size_t]
        L.orc_lu_free.argtypes = [C.POINTER(OrcLU)]
        L.orc_factor_perm.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, _I64P, _I64P, C.c_int, C.c_int, C.c_double,
                                      C.c_int64, C.c_double, C.POINTER(C.POINTER(OrcNum)), C.c_char_p, C.c_size_t]
        L.orc_num_free.argtypes = [C.POINTER(OrcNum)]
        L.orc_solve.argtypes = [C.POINTER(OrcNum), _DP, C.c_int64, C.c_int64, C.c_int]
        L.orc_tri_solve.argtypes = [C.c_int64, _I64P, _I64P, _DP, _I64P, _I64P, _DP, _DP, C.c_int, C.c_int]
        L.orc_backward_error.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, _DP, _DP, C.c_int]
        L.orc_backward_error.restype = C.c_double
        L.orc_plan_nsupernodes.argtypes = [C.c_int64, _I64P, _I64P, _I64P, _I64P]
        L.orc_plan_nsupernodes.restype = C.c_int64
        L.orc_factor_row.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, C.c_double, C.c_int, C.c_double,
                                     C.POINTER(C.POINTER(OrcRowLU)), C.c_char_p, C.c_size_t]
        L.orc_rowlu_free.argtypes = [C.POINTER(OrcRowLU)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle error {code}: {msg}")
        self.code = code


def _csc(A):
    import scipy.sparse as sp
    A = sp.csc_matrix(A)
    A.sort_indices()
    return (A.shape[0], np.ascontiguousarray(A.indptr, dtype=np.int64), np.ascontiguousarray(A.indices, dtype=np.int64),
            np.ascontiguousarray(A.data))


def permute(n, Ap, Ai, Ax, row_perm, col_perm):
    cplx = int(np.iscomplexobj(Ax))
    Bp = np.empty(n + 1, np.int64)
    Bi = np.empty(len(Ai), np.int64)
    Bx = np.empty_like(np.ascontiguousarray(Ax))
    rc = lib().orc_permute(n, _p64(Ap), _p64(Ai), _pd(_f64view(Ax)), cplx, _p64(np.ascontiguousarray(row_perm, np.int64)),
                           _p64(np.ascontiguousarray(col_perm, np.int64)), _p64(Bp), _p64(Bi), _pd(_f64view(Bx)))
    if rc != 0:
        raise OracleError(rc, "permute")
    return Bp, Bi, Bx


def ruiz(n, ptr, idx, x, is_csc=True, max_iter=20, tol=1e-2):
    """returns (scaled values, Dr, Dc, iterations)"""
    x = np.array(x, copy=True)
    Dr = np.empty(n)
    Dc = np.empty(n)
    it = C.c_int(0)
    err = C.create_string_buffer(256)
    rc = lib().orc_ruiz(n, _p64(ptr), _p64(idx), _pd(_f64view(x)), int(np.iscomplexobj(x)), int(is_csc), max_iter, tol,
                        _pd(Dr), _pd(Dc), C.byref(it), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return x, Dr, Dc, it.value


def _lu_dict(f, cplx):
    n = f.n
    nL = f.Lp[n]
    nU = f.Up[n]
    vs = 2 if cplx else 1

    def arr(ptr, cnt, dt):
        return np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt, copy=True)

    Lx = arr(f.Lx, nL * vs, np.float64)
    Ux = arr(f.Ux, nU * vs, np.float64)
    if cplx:
        Lx = Lx.view(np.complex128)
        Ux = Ux.view(np.complex128)
    return dict(n=n, nsn=f.nsn, Lp=arr(f.Lp, n + 1, np.int64), Li=arr(f.Li, nL, np.int64), Lx=Lx,
                Up=arr(f.Up, n + 1, np.int64), Ui=arr(f.Ui, nU, np.int64), Ux=Ux, p=arr(f.p, n, np.int64),
                sn_first=arr(f.sn_first, f.nsn + 1, np.int64))


def factor_sn(n, Ap, Ai, Ax, relax_rows=32, pivot_rel=1.4901161193847656e-08):
    cplx = int(np.iscomplexobj(Ax))
    out = C.POINTER(OrcLU)()
    err = C.create_string_buffer(256)
    rc = lib().orc_factor_sn(n, _p64(Ap), _p64(Ai), _pd(_f64view(Ax)), cplx, relax_rows, pivot_rel, C.byref(out), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        return _lu_dict(out.contents, cplx)
    finally:
        lib().orc_lu_free(out)


class Numeric:
    """the composed factorization (permute, Ruiz, supernodal LU, p / q / Dr / Dc in original coordinates)"""

    def __init__(self, n, Ap, Ai, Ax, row_perm, col_perm, equilibrate=True, ruiz_max_iter=20, ruiz_tol=1e-2,
                 relax_rows=32, pivot_rel=1.4901161193847656e-08):
        self.n = n
        self.cplx = int(np.iscomplexobj(Ax))
        self._keep = (np.ascontiguousarray(Ap, np.int64), np.ascontiguousarray(Ai, np.int64), np.ascontiguousarray(Ax))
        self.ptr = C.POINTER(OrcNum)()
        err = C.create_string_buffer(256)
        rc = lib().orc_factor_perm(n, _p64(self._keep[0]), _p64(self._keep[1]), _pd(_f64view(self._keep[2])), self.cplx,
                                   _p64(np.ascontiguousarray(row_perm, np.int64)), _p64(np.ascontiguousarray(col_perm, np.int64)),
                                   int(equilibrate), ruiz_max_iter, ruiz_tol, relax_rows, pivot_rel, C.byref(self.ptr), err, 256)
        if rc != 0:
            raise OracleError(rc, err.value.decode())
        c = self.ptr.contents
        self.lu = _lu_dict(c.lu.contents, self.cplx)
        self.p = np.ctypeslib.as_array(c.p, shape=(n,)).copy()
        self.q = np.ctypeslib.as_array(c.q, shape=(n,)).copy()
        self.Dr = np.ctypeslib.as_array(c.Dr, shape=(n,)).copy()
        self.Dc = np.ctypeslib.as_array(c.Dc, shape=(n,)).copy()
        self.ruiz_iters = c.ruiz_iters

    def solve(self, b, trans=0):
        b = np.asarray(b)
        one = b.ndim == 1
        dt = np.complex128 if (self.cplx or np.iscomplexobj(b)) else np.float64
        x = np.array(b.reshape(self.n, -1), dtype=dt, copy=True)
        out = np.empty_like(x)
        for r in range(x.shape[1]):
            col = np.ascontiguousarray(x[:, r])
            rc = lib().orc_solve(self.ptr, _pd(_f64view(col)), 1, self.n, trans)
            if rc != 0:
                raise OracleError(rc, "solve")
            out[:, r] = col
        return out[:, 0] if one else out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_num_free(self.ptr)
            self.ptr = None


def tri_solve(L, U, b, trans=0):
    """the plain sweeps on CSC factors (reference src/solve.c:11-60), one right-hand side at a time"""
    Lp, Li, Lx = L
    Up, Ui, Ux = U
    cplx = np.iscomplexobj(Lx) or np.iscomplexobj(Ux) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    Lx, Ux = np.ascontiguousarray(Lx, dt), np.ascontiguousarray(Ux, dt)
    Lp, Li, Up, Ui = (np.ascontiguousarray(a, np.int64) for a in (Lp, Li, Up, Ui))
    b = np.asarray(b)
    n = len(Lp) - 1
    X = np.array(b.reshape(n, -1), dtype=dt, copy=True)
    for r in range(X.shape[1]):
        col = np.ascontiguousarray(X[:, r])
        rc = lib().orc_tri_solve(n, _p64(Lp), _p64(Li), _pd(_f64view(Lx)), _p64(Up), _p64(Ui), _pd(_f64view(Ux)),
                                 _pd(_f64view(col)), trans, int(cplx))
        if rc != 0:
            raise OracleError(rc, "tri_solve")
        X[:, r] = col
    return X[:, 0] if b.ndim == 1 else X


def plan_nsupernodes(Lp, Li, Up, Ui):
    Lp, Li, Up, Ui = (np.ascontiguousarray(a, np.int64) for a in (Lp, Li, Up, Ui))
    return lib().orc_plan_nsupernodes(len(Lp) - 1, _p64(Lp), _p64(Li), _p64(Up), _p64(Ui))


def backward_error(n, Ap, Ai, Ax, x, b, trans=0):
    cplx = int(np.iscomplexobj(Ax))
    x = np.ascontiguousarray(x, dtype=np.complex128 if cplx else np.float64)
    b = np.ascontiguousarray(b, dtype=np.complex128 if cplx else np.float64)
    return lib().orc_backward_error(n, _p64(Ap), _p64(Ai), _pd(_f64view(Ax)), cplx, _pd(_f64view(x)), _pd(_f64view(b)), trans)


def _row_factors(n, cplx, Lp, Li, Lx, Up, Ui, Ux, p, q, nrep):
    """CSR factors out of C arrays (copied)"""
    dt = np.complex128 if cplx else np.float64
    vs = 2 if cplx else 1

    def arr(ptr, cnt):
        return np.ctypeslib.as_array(ptr, shape=(cnt,)).copy() if cnt > 0 else np.zeros(0, np.int64)

    def vals(ptr, cnt):
        if cnt == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(C.cast(ptr, _DP), shape=(cnt * vs,)).copy().view(dt)

    Lp_, Up_ = arr(Lp, n + 1), arr(Up, n + 1)
    return dict(n=n, Lp=Lp_, Li=arr(Li, int(Lp_[n])), Lx=vals(Lx, int(Lp_[n])), Up=Up_, Ui=arr(Ui, int(Up_[n])),
                Ux=vals(Ux, int(Up_[n])), p=arr(p, n), q=arr(q, n), n_replaced=int(nrep))


def factor_row(n, Rp, Ri, Rx, drop_tol=0.0, pivot_strategy=1, pivot_rel=1.4901161193847656e-08):
    """the oracle's row kernel on a CSR matrix: dict of CSR factors (L unit diagonal first, U from the diagonal), p, q,
    n_replaced; OracleError (code, message; .fail_step on a singular matrix) otherwise"""
    Rp = np.ascontiguousarray(Rp, np.int64)
    Ri = np.ascontiguousarray(Ri, np.int64)
    Rx = np.ascontiguousarray(Rx)
    cplx = int(np.iscomplexobj(Rx))
    f = C.POINTER(OrcRowLU)()
    err = C.create_string_buffer(256)
    rc = lib().orc_factor_row(n, _p64(Rp), _p64(Ri), _pd(_f64view(Rx)), cplx, drop_tol, pivot_strategy, pivot_rel,
                              C.byref(f), err, 256)
    if rc != 0:
        e = OracleError(rc, err.value.decode())
        if f:
            e.fail_step = f.contents.fail_step
            lib().orc_rowlu_free(f)
        raise e
    try:
        F = f.contents
        return _row_factors(n, cplx, F.Lp, F.Li, F.Lx, F.Up, F.Ui, F.Ux, F.p, F.q, F.n_replaced)
    finally:
        lib().orc_rowlu_free(f)


# ---------------------------------------------------------------- the reference's own numeric sources (oracle/_ref)
class RefFactors(C.Structure):
    """fplus_factors of the reference header (include/fplus.h:37-46)"""
    _fields_ = [("n", C.c_int64), ("is_complex", C.c_int), ("csc", C.c_int), ("Lp", _I64P), ("Li", _I64P), ("Lx", C.c_void_p),
                ("Up", _I64P), ("Ui", _I64P), ("Ux", C.c_void_p), ("p", _I64P), ("q", _I64P), ("nsupernodes", C.c_int64),
                ("n_pivots_replaced", C.c_int64)]


_ref = None
REF_PATH = os.path.join(HERE, "_ref", "libfplus_ref.so")


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_PATH)
        L.fplus_ruiz.argtypes = [C.c_int64, _I64P, _I64P, C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                 _DP, _DP, C.c_char_p, C.c_size_t]
        L.fplus_factor_sn.argtypes = [C.c_int64, _I64P, _I64P, C.c_void_p, C.c_int, C.c_int, C.c_int64, C.c_int64, C.c_int,
                                      C.c_double, C.c_double, C.POINTER(RefFactors), C.c_char_p, C.c_size_t]
        L.fplus_factors_free.argtypes = [C.POINTER(RefFactors)]
        L.fplus_tri_solve.argtypes = [C.c_int64, _I64P, _I64P, C.c_void_p, _I64P, _I64P, C.c_void_p, C.c_void_p, C.c_int, C.c_int]
        L.fplus_solve_plan_create.argtypes = [C.c_int64, _I64P, _I64P, _I64P, _I64P, C.POINTER(C.c_void_p), C.c_char_p, C.c_size_t]
        L.fplus_solve_plan_nsupernodes.argtypes = [C.c_void_p]
        L.fplus_solve_plan_nsupernodes.restype = C.c_int64
        L.fplus_solve_plan_free.argtypes = [C.c_void_p]
        L.fplus_factor_row.argtypes = [C.c_int64, _I64P, _I64P, C.c_void_p, C.c_double, C.c_int, C.c_int, C.c_int,
                                       C.c_double, C.POINTER(RefFactors), C.c_char_p, C.c_size_t]
        _ref = L
    return _ref


def ref_plan_nsupernodes(Lp, Li, Up, Ui):
    """the reference's own solve plan (oracle/_ref) on these patterns: its supernode count"""
    Lp, Li, Up, Ui = (np.ascontiguousarray(a, np.int64) for a in (Lp, Li, Up, Ui))
    pl = C.c_void_p()
    err = C.create_string_buffer(256)
    rc = ref().fplus_solve_plan_create(len(Lp) - 1, _p64(Lp), _p64(Li), _p64(Up), _p64(Ui), C.byref(pl), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        return ref().fplus_solve_plan_nsupernodes(pl)
    finally:
        ref().fplus_solve_plan_free(pl)


def ref_ruiz(n, ptr, idx, x, is_csc=True, max_iter=20, tol=1e-2):
    x = np.array(x, copy=True)
    Dr = np.empty(n)
    Dc = np.empty(n)
    err = C.create_string_buffer(256)
    rc = ref().fplus_ruiz(n, _p64(ptr), _p64(idx), x.ctypes.data, int(np.iscomplexobj(x)), int(is_csc), max_iter, tol, 0,
                          _pd(Dr), _pd(Dc), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    return x, Dr, Dc


def ref_factor_sn(n, Ap, Ai, Ax, relax_rows=32, panel=1, pivot_rel=1.4901161193847656e-08):
    """the reference's fplus_factor_sn (column kernel: the library is built without a CBLAS)"""
    cplx = int(np.iscomplexobj(Ax))
    f = RefFactors()
    err = C.create_string_buffer(256)
    Ax = np.ascontiguousarray(Ax)
    rc = ref().fplus_factor_sn(n, _p64(Ap), _p64(Ai), Ax.ctypes.data, cplx, 0, relax_rows, panel, 0, 5e7, pivot_rel,
                               C.byref(f), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        nL = f.Lp[n]
        nU = f.Up[n]
        dt = np.complex128 if cplx else np.float64

        def arr(ptr, cnt):
            return np.ctypeslib.as_array(ptr, shape=(cnt,)).copy()

        Lx = np.ctypeslib.as_array(C.cast(f.Lx, _DP), shape=(nL * (2 if cplx else 1),)).copy().view(dt)
        Ux = np.ctypeslib.as_array(C.cast(f.Ux, _DP), shape=(nU * (2 if cplx else 1),)).copy().view(dt)
        return dict(n=n, nsn=f.nsupernodes, Lp=arr(f.Lp, n + 1), Li=arr(f.Li, nL), Lx=Lx, Up=arr(f.Up, n + 1),
                    Ui=arr(f.Ui, nU), Ux=Ux, p=arr(f.p, n), q=arr(f.q, n))
    finally:
        ref().fplus_factors_free(C.byref(f))


# ---------------------------------------------------------------- the OpenMP panel-blocked CPU baseline (cpu_panel.c)
_cp = None


def cpu_panel_lib():
    global _cp
    if _cp is None:
        path = os.path.join(HERE, "libcpupanel.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libcpupanel.so is not built (make -C oracle)")
        L = C.CDLL(path)
        L.cpb_factor_solve.argtypes = [C.c_int64, _I64P, _I64P, _DP, C.c_int, _I64P, _I64P, C.c_int64, C.c_double, C.c_int64,
                                       C.c_int, _DP, C.c_int64, _DP, C.c_char_p, C.c_size_t]
        L.cpb_factor_solve.restype = C.c_int
        _cp = L
    return _cp


def cpu_panel_factor_solve(n, Ap, Ai, Ax, row_perm, col_perm, b, relax_rows=32, pivot_rel=1.4901161193847656e-08, panel=256,
                           threads=None):
    """x = A^-1 b by the OpenMP panel-blocked CPU path, with its timings: (x, {"prep_s", "factor_s", "solve_s", "flops",
    "supernodes", "threads"})"""
    cplx = np.iscomplexobj(Ax)
    Ap = np.ascontiguousarray(Ap, np.int64)
    Ai = np.ascontiguousarray(Ai, np.int64)
    xv = _f64view(np.ascontiguousarray(Ax, np.complex128 if cplx else np.float64))
    rp = np.ascontiguousarray(row_perm, np.int64)
    cp = np.ascontiguousarray(col_perm, np.int64)
    X = np.array(np.asarray(b).reshape(n, -1), dtype=np.complex128 if cplx else np.float64, order="F", copy=True)
    out = np.zeros(6)
    err = C.create_string_buffer(512)
    nth = int(threads or os.cpu_count() or 1)
    rc = cpu_panel_lib().cpb_factor_solve(n, _p64(Ap), _p64(Ai), _pd(xv), int(cplx), _p64(rp), _p64(cp), int(relax_rows),
                                          float(pivot_rel), int(panel), nth, C.cast(X.ctypes.data, _DP),
                                          X.shape[1], _pd(out), err, 512)
    if rc != 0:
        raise OracleError(rc, err.value.decode(errors="replace"))
    info = {"prep_s": out[0], "factor_s": out[1], "solve_s": out[2], "flops": out[3], "supernodes": int(out[4]),
            "threads": int(out[5])}
    return (X[:, 0] if np.asarray(b).ndim == 1 else X), info


def ref_factor_row(n, Rp, Ri, Rx, drop_tol=0.0, pivot_strategy=1, pivot_rel=1.4901161193847656e-08):
    """the reference's own fplus_factor_row (oracle/_ref, built without OpenMP or a CBLAS) on a CSR matrix"""
    Rp = np.ascontiguousarray(Rp, np.int64)
    Ri = np.ascontiguousarray(Ri, np.int64)
    Rx = np.ascontiguousarray(Rx)
    cplx = int(np.iscomplexobj(Rx))
    f = RefFactors()
    err = C.create_string_buffer(256)
    rc = ref().fplus_factor_row(n, _p64(Rp), _p64(Ri), Rx.ctypes.data, drop_tol, cplx, 0, pivot_strategy, pivot_rel,
                                C.byref(f), err, 256)
    if rc != 0:
        raise OracleError(rc, err.value.decode())
    try:
        return _row_factors(n, cplx, f.Lp, f.Li, f.Lx, f.Up, f.Ui, f.Ux, f.p, f.q, f.n_pivots_replaced)
    finally:
        ref().fplus_factors_free(C.byref(f))
