"""fplus_amd -- the FPLUS sparse LU on MI355X.

The numeric factorization (Ruiz equilibration, the supernodal left-looking LU with partial pivoting) and the triangular
solves run as HIP kernels on one GPU behind the FPLUS C interface (include/fplus.h, libfplus_amd.so).  This module
binds that interface with ctypes and provides ``SparseLU``, the drop-in for ``scipy.sparse.linalg.splu`` of the
reference (reference sparse_lu_full_pivot.py:630-836): ``solve(b, trans)``, ``L``, ``U``, ``perm_r``, ``perm_c``,
``shape``, ``nnz`` (fplus_amd.splu, with the reference module's other functions).
"""
import ctypes as C

import numpy as np
import scipy.sparse as sp

from ._lib import (FPLUS_ORDER_AMD_ATA, FPLUS_ORDER_AUTO, FPLUS_ORDER_COLAMD, FPLUS_PIVOT_PARTIAL,  # noqa: F401
                   FPLUS_PIVOT_REL_DEFAULT, AmdOptions, Factors, FplusError, Options, RefineInfo, RefineOptions, RowInfo, Stats,
                   check, lib)

__all__ = ["SparseLU", "sparse_lu_full_pivot", "lu_solve", "lu_residual", "clear_ordering_cache", "Symbolic", "Numeric",
           "SolvePlan", "default_options", "amd_options", "FplusError", "ruiz", "factor_sn", "factor_sn_perm", "factor_sn_sym",
           "factor_row", "factor_row_sym", "factor_row_batch", "row_stats", "tri_solve", "tune", "Selector", "log_blocks"]

_I64P = C.POINTER(C.c_int64)
_DP = C.POINTER(C.c_double)


def _p64(a):
    return a.ctypes.data_as(_I64P)


def _vp(a):
    return C.c_void_p(a.ctypes.data)


def _canonical_csc(A):
    A = sp.csc_matrix(A)
    if not A.has_canonical_format:
        A = A.copy()
        A.sum_duplicates()
    A.sort_indices()
    n = A.shape[0]
    return n, np.ascontiguousarray(A.indptr, np.int64), np.ascontiguousarray(A.indices, np.int64), A.data


def default_options(**kw):
    o = Options()
    lib().fplus_default_options(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        setattr(o, k, v)
    return o


class Symbolic:
    """fplus_analyze: dense-line pinning, the fill-reducing order, the column-etree postorder"""

    def __init__(self, n, Ap, Ai, pin_dense=True, tau_density=0.5, tau_min_nnz=10, tau_row_mask=None, tau_col_mask=None,
                 postorder=True, ordering=FPLUS_ORDER_COLAMD):
        self.n = n
        self._Ap = np.ascontiguousarray(Ap, np.int64)
        self._Ai = np.ascontiguousarray(Ai, np.int64)
        self.ptr = C.c_void_p()
        err = C.create_string_buffer(512)
        rm = None if tau_row_mask is None else np.ascontiguousarray(tau_row_mask, np.uint8)
        cm = None if tau_col_mask is None else np.ascontiguousarray(tau_col_mask, np.uint8)
        self._masks = (rm, cm)
        rc = lib().fplus_analyze_ord(n, _p64(self._Ap), _p64(self._Ai), int(pin_dense), tau_density, tau_min_nnz,
                                     None if rm is None else _vp(rm), None if cm is None else _vp(cm), int(postorder),
                                     ordering, C.byref(self.ptr), err, 512)
        check(rc, err)
        self.row_perm = np.ctypeslib.as_array(lib().fplus_symbolic_row_perm(self.ptr), (n,)).copy() if n else np.zeros(0, np.int64)
        self.col_perm = np.ctypeslib.as_array(lib().fplus_symbolic_col_perm(self.ptr), (n,)).copy() if n else np.zeros(0, np.int64)
        self.n_dense_rows = lib().fplus_symbolic_n_dense_rows(self.ptr)
        self.n_dense_cols = lib().fplus_symbolic_n_dense_cols(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fplus_symbolic_free(self.ptr)
            self.ptr = None


def amd_options(**kw):
    """fplus_amd_options with the given fields (precision, refactor, static_pivots, selector: a Selector)"""
    o = AmdOptions()
    lib().fplus_amd_default_options(C.byref(o))
    o._selector = None
    for k, v in kw.items():
        if not hasattr(o, k):
            raise TypeError(f"unknown option {k}")
        if k == "selector":
            o._selector = v          # kept alive with the options (and the factorization made with them)
            v = v.ptr if v is not None else None
        setattr(o, k, v)
    return o


class Numeric:
    """fplus_factor on a symbolic analysis: the device factorization, its composed orders and scalings.  With
    ``amd_options`` (fplus_amd_options: update precision, keeping what a refactorization replays) it is made by
    fplus_amd_factor_ex."""

    def __init__(self, sym, Ap, Ai, Ax, options=None, amd_options=None):
        self.sym = sym
        self.n = sym.n
        self.is_complex = bool(np.iscomplexobj(Ax))
        self._Ax = np.ascontiguousarray(Ax, dtype=np.complex128 if self.is_complex else np.float64)
        Ap = np.ascontiguousarray(Ap, np.int64)
        Ai = np.ascontiguousarray(Ai, np.int64)
        self.ptr = C.c_void_p()
        err = C.create_string_buffer(512)
        o = options if options is not None else default_options()
        if amd_options is None:
            rc = lib().fplus_factor(sym.ptr, _p64(Ap), _p64(Ai), _vp(self._Ax), int(self.is_complex), C.byref(o),
                                    C.byref(self.ptr), err, 512)
        else:
            rc = lib().fplus_amd_factor_ex(sym.ptr, _p64(Ap), _p64(Ai), _vp(self._Ax), 0, int(self.is_complex), C.byref(o),
                                           C.byref(amd_options), C.byref(self.ptr), err, 512)
        check(rc, err)
        self._selector = getattr(amd_options, "_selector", None)
        self._refresh()

    def _refresh(self):
        n = self.n
        self.p = np.ctypeslib.as_array(lib().fplus_numeric_p(self.ptr), (n,)).copy()
        self.q = np.ctypeslib.as_array(lib().fplus_numeric_q(self.ptr), (n,)).copy()
        self.Dr = np.ctypeslib.as_array(lib().fplus_numeric_Dr(self.ptr), (n,)).copy()
        self.Dc = np.ctypeslib.as_array(lib().fplus_numeric_Dc(self.ptr), (n,)).copy()

    def refactor(self, Ax, precision=0, static_pivots=0):
        """fplus_amd_refactor: new values of the same pattern, factored in place in the recorded pivot order"""
        Ax = np.ascontiguousarray(Ax, dtype=np.complex128 if self.is_complex else np.float64)
        err = C.create_string_buffer(512)
        check(lib().fplus_amd_refactor(self.ptr, _vp(Ax), 0, int(precision), int(static_pivots), err, 512), err)
        self._Ax = Ax
        self._refresh()

    def solve_refined(self, b, max_outer=10, restart=8, tol=1e-14, inner_tol=1e-4, verbose=0):
        """fplus_amd_solve_refined: (x, info dict) for A x = b, b of shape (n,) or (n, k)"""
        b = np.asarray(b)
        one = b.ndim == 1
        dt = np.complex128 if self.is_complex else np.float64
        x = np.array(b.reshape(self.n, -1), dtype=dt, order="F", copy=True)
        ro = RefineOptions()
        lib().fplus_amd_default_refine_options(C.byref(ro))
        ro.max_outer, ro.restart, ro.tol, ro.inner_tol, ro.verbose = max_outer, restart, tol, inner_tol, verbose
        info = RefineInfo()
        err = C.create_string_buffer(512)
        check(lib().fplus_amd_solve_refined(self.ptr, C.c_void_p(x.ctypes.data), 0, x.shape[1], self.n, C.byref(ro),
                                            C.byref(info), err, 512), err)
        out = {k: getattr(info, k) for k, _ in RefineInfo._fields_ if k != "reserved"}
        return (x[:, 0] if one else x), out

    def refactor_solve(self, Ax, b, precision=2, max_outer=10, restart=8, tol=1e-14, inner_tol=1e-4):
        """fplus_amd_refactor_solve: refactorization at `precision` (static pivoting) and a refined solve, again with
        FP64 updates when the refinement stalls; returns (x, info dict with "fallback")"""
        Ax = np.ascontiguousarray(Ax, dtype=np.complex128 if self.is_complex else np.float64)
        b = np.asarray(b)
        one = b.ndim == 1
        dt = np.complex128 if self.is_complex else np.float64
        x = np.array(b.reshape(self.n, -1), dtype=dt, order="F", copy=True)
        ro = RefineOptions()
        lib().fplus_amd_default_refine_options(C.byref(ro))
        ro.max_outer, ro.restart, ro.tol, ro.inner_tol = max_outer, restart, tol, inner_tol
        info = RefineInfo()
        fb = C.c_int(0)
        err = C.create_string_buffer(512)
        check(lib().fplus_amd_refactor_solve(self.ptr, _vp(Ax), 0, C.c_void_p(x.ctypes.data), 0, x.shape[1], self.n,
                                             int(precision), C.byref(ro), C.byref(info), C.byref(fb), err, 512), err)
        self._Ax = Ax
        self._refresh()
        out = {k: getattr(info, k) for k, _ in RefineInfo._fields_ if k != "reserved"}
        out["fallback"] = fb.value
        return (x[:, 0] if one else x), out

    def refactor_sweep(self, Ax_list, b_list, precision=0, max_outer=10, restart=8, tol=1e-14, inner_tol=1e-4):
        """fplus_amd_refactor_sweep: each value set of Ax_list refactored and its right-hand sides b_list[k] (n or n x r,
        the same r for every shift) solved; returns ([x_k], [info dict per shift])"""
        dt = np.complex128 if self.is_complex else np.float64
        nsh = len(Ax_list)
        Ax = np.ascontiguousarray(np.stack([np.asarray(a, dtype=dt) for a in Ax_list], axis=0))
        bs = [np.asarray(x).reshape(self.n, -1) for x in b_list]
        r = bs[0].shape[1] if bs else 0
        B = np.empty((nsh, r, self.n), dtype=dt)   # shift-major, then column-major per shift
        for k, x in enumerate(bs):
            B[k] = x.T
        info = (RefineInfo * max(nsh, 1))()
        fb = (C.c_int * max(nsh, 1))()
        ro = RefineOptions()
        lib().fplus_amd_default_refine_options(C.byref(ro))
        ro.max_outer, ro.restart, ro.tol, ro.inner_tol = max_outer, restart, tol, inner_tol
        err = C.create_string_buffer(512)
        check(lib().fplus_amd_refactor_sweep(self.ptr, _vp(Ax), nsh, C.c_void_p(B.ctypes.data), r, self.n, 0, int(precision),
                                             C.byref(ro), info, fb, err, 512), err)
        self._Ax = Ax[-1] if nsh else self._Ax
        self._refresh()
        xs = [B[k].T.copy() for k in range(nsh)]
        xs = [x[:, 0] if np.asarray(b_list[k]).ndim == 1 else x for k, x in enumerate(xs)]
        infos = []
        for k in range(nsh):
            d = {f: getattr(info[k], f) for f, _ in RefineInfo._fields_ if f != "reserved"}
            d["fallback"] = fb[k]
            infos.append(d)
        return xs, infos

    def solve(self, b, trans=0):
        """solve in place on a copy: trans 0 A x = b, 1 A^T x = b, 2 A^H x = b; b of shape (n,) or (n, k)"""
        b = np.asarray(b)
        one = b.ndim == 1
        dt = np.complex128 if self.is_complex else (np.complex128 if np.iscomplexobj(b) else np.float64)
        if np.iscomplexobj(b) and not self.is_complex:   # a real factorization: real and imaginary parts apart
            xr = self.solve(b.real, trans)
            xi = self.solve(b.imag, trans)
            return xr + 1j * xi
        x = np.array(b.reshape(self.n, -1), dtype=dt, order="F", copy=True)
        err = C.create_string_buffer(512)
        rc = lib().fplus_solve(self.ptr, C.c_void_p(x.ctypes.data), x.shape[1], self.n, trans, err, 512)
        check(rc, err)
        return x[:, 0] if one else x

    def factors(self):
        """the CSC factors of diag(Dr) A[p, :][:, q] diag(Dc) = L U as scipy matrices, and the pivot order"""
        f = lib().fplus_numeric_factors(self.ptr)
        if not f:
            raise FplusError(-5, "factor export failed")
        f = f.contents
        n = f.n
        nL, nU = f.Lp[n], f.Up[n]
        dt = np.complex128 if f.is_complex else np.float64
        vs = 2 if f.is_complex else 1
        Lp = np.ctypeslib.as_array(f.Lp, (n + 1,)).copy()
        Li = np.ctypeslib.as_array(f.Li, (max(nL, 1),))[:nL].copy()
        Lx = np.ctypeslib.as_array(C.cast(f.Lx, _DP), (max(nL * vs, 1),))[:nL * vs].copy().view(dt)
        Up = np.ctypeslib.as_array(f.Up, (n + 1,)).copy()
        Ui = np.ctypeslib.as_array(f.Ui, (max(nU, 1),))[:nU].copy()
        Ux = np.ctypeslib.as_array(C.cast(f.Ux, _DP), (max(nU * vs, 1),))[:nU * vs].copy().view(dt)
        p = np.ctypeslib.as_array(f.p, (n,)).copy()
        L = sp.csc_matrix((Lx, Li, Lp), shape=(n, n))
        U = sp.csc_matrix((Ux, Ui, Up), shape=(n, n))
        return dict(L=L, U=U, p=p, nsupernodes=f.nsupernodes, Lp=Lp, Li=Li, Lx=Lx, Up=Up, Ui=Ui, Ux=Ux,
                    n_pivots_replaced=f.n_pivots_replaced)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fplus_numeric_free(self.ptr)
            self.ptr = None


class _IdentitySymbolic(Symbolic):
    """the identity orders (colamd=False), through fplus_amd_symbolic_from_orders"""

    def __init__(self, n, Ap, Ai):
        self.n = n
        self._Ap = np.ascontiguousarray(Ap, np.int64)
        self._Ai = np.ascontiguousarray(Ai, np.int64)
        self.row_perm = np.arange(n, dtype=np.int64)
        self.col_perm = np.arange(n, dtype=np.int64)
        self.n_dense_rows = self.n_dense_cols = 0
        self.ptr = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = lib().fplus_amd_symbolic_from_orders(n, _p64(self._Ap), _p64(self._Ai), _p64(self.row_perm), _p64(self.col_perm),
                                                  C.byref(self.ptr), err, 512)
        check(rc, err)


def symbolic_from_orders(n, Ap, Ai, row_perm, col_perm):
    """a Symbolic with the given orders (fplus_amd_symbolic_from_orders)"""
    s = _IdentitySymbolic.__new__(_IdentitySymbolic)
    s.n = n
    s._Ap = np.ascontiguousarray(Ap, np.int64)
    s._Ai = np.ascontiguousarray(Ai, np.int64)
    s.row_perm = np.ascontiguousarray(row_perm, np.int64)
    s.col_perm = np.ascontiguousarray(col_perm, np.int64)
    s.n_dense_rows = s.n_dense_cols = 0
    s.ptr = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = lib().fplus_amd_symbolic_from_orders(n, _p64(s._Ap), _p64(s._Ai), _p64(s.row_perm), _p64(s.col_perm), C.byref(s.ptr),
                                              err, 512)
    check(rc, err)
    return s


def stats(num):
    """fplus_amd_numeric_stats of a Numeric as a dict"""
    st = Stats()
    check(lib().fplus_amd_numeric_stats(num.ptr, C.byref(st)), None)
    return {k: getattr(st, k) for k, _ in Stats._fields_}


def row_stats():
    """fplus_amd_row_stats: the calling thread's last row-kernel factorization (device runs, seconds, row merges and
    their algorithmic bytes, pool entries and multipliers) as a dict"""
    st = RowInfo()
    check(lib().fplus_amd_row_stats(C.byref(st)), None)
    return {k: getattr(st, k) for k, _ in RowInfo._fields_}


def ruiz(n, indptr, indices, data, is_csc=True, max_iter=20, tol=1e-2):
    """fplus_ruiz on the device: (scaled data, Dr, Dc)"""
    x = np.array(data, copy=True)
    Dr = np.empty(n)
    Dc = np.empty(n)
    err = C.create_string_buffer(512)
    ip = np.ascontiguousarray(indptr, np.int64)
    ix = np.ascontiguousarray(indices, np.int64)
    rc = lib().fplus_ruiz(n, _p64(ip), _p64(ix), _vp(x), int(np.iscomplexobj(x)), int(is_csc), max_iter, tol, 0,
                          Dr.ctypes.data_as(_DP), Dc.ctypes.data_as(_DP), err, 512)
    check(rc, err)
    return x, Dr, Dc


def _factors_dict(f, n, cplx):
    """an fplus_factors as numpy arrays (copies); the struct is freed"""
    try:
        nL, nU = f.Lp[n], f.Up[n]
        dt = np.complex128 if cplx else np.float64
        vs = 2 if cplx else 1
        return dict(n=n, nsn=f.nsupernodes, Lp=np.ctypeslib.as_array(f.Lp, (n + 1,)).copy(),
                    Li=np.ctypeslib.as_array(f.Li, (max(nL, 1),))[:nL].copy(),
                    Lx=np.ctypeslib.as_array(C.cast(f.Lx, _DP), (max(nL * vs, 1),))[:nL * vs].copy().view(dt),
                    Up=np.ctypeslib.as_array(f.Up, (n + 1,)).copy(),
                    Ui=np.ctypeslib.as_array(f.Ui, (max(nU, 1),))[:nU].copy(),
                    Ux=np.ctypeslib.as_array(C.cast(f.Ux, _DP), (max(nU * vs, 1),))[:nU * vs].copy().view(dt),
                    p=np.ctypeslib.as_array(f.p, (n,)).copy(),
                    q=np.ctypeslib.as_array(f.q, (n,)).copy() if f.q else np.arange(n, dtype=np.int64),
                    csc=int(f.csc), n_pivots_replaced=int(f.n_pivots_replaced))
    finally:
        lib().fplus_factors_free(C.byref(f))


_PIVOT_CODES = {"full": 0, "partial": 1, "rook": 2, 0: 0, 1: 1, 2: 2}


def factor_row(n, Rp, Ri, Rx, drop_tol=0.0, pivot_strategy="full", pivot_rel=FPLUS_PIVOT_REL_DEFAULT, verbose=0):
    """fplus_factor_row (the reference's _sparse_lu_c.lu_factor): the row kernel on the device on a canonical CSR matrix,
    full / partial / rook pivoting, drop_tol > 0 the incomplete factorization; CSR factors (L unit diagonal first, U
    from the diagonal on), p, q and n_pivots_replaced"""
    Rx = np.ascontiguousarray(Rx)
    cplx = int(np.iscomplexobj(Rx))
    f = Factors()
    err = C.create_string_buffer(512)
    rc = lib().fplus_factor_row(n, _p64(np.ascontiguousarray(Rp, np.int64)), _p64(np.ascontiguousarray(Ri, np.int64)), _vp(Rx),
                                float(drop_tol), cplx, int(verbose), _PIVOT_CODES[pivot_strategy], float(pivot_rel), C.byref(f),
                                err, 512)
    check(rc, err)
    return _factors_dict(f, n, cplx)


def factor_row_batch(systems, drop_tol=0.0, pivot_strategy="full", pivot_rel=FPLUS_PIVOT_REL_DEFAULT):
    """fplus_amd_factor_row_batch: independent CSR systems [(n, Rp, Ri, Rx), ...] of one scalar type through the row
    kernel in shared launches, one workgroup per system; a list with, per system, the dict factor_row returns or the
    FplusError it would raise"""
    k = len(systems)
    cplx = int(any(np.iscomplexobj(s[3]) for s in systems))
    dt = np.complex128 if cplx else np.float64
    keep = [(int(n), np.ascontiguousarray(Rp, np.int64), np.ascontiguousarray(Ri, np.int64), np.ascontiguousarray(Rx, dt))
            for n, Rp, Ri, Rx in systems]
    ns = np.array([s[0] for s in keep] or [0], np.int64)
    P = (C.c_void_p * max(k, 1))(*[s[1].ctypes.data for s in keep])
    I = (C.c_void_p * max(k, 1))(*[s[2].ctypes.data for s in keep])
    X = (C.c_void_p * max(k, 1))(*[s[3].ctypes.data for s in keep])
    out = (Factors * max(k, 1))()
    rc = (C.c_int * max(k, 1))()
    err = C.create_string_buffer(512)
    check(lib().fplus_amd_factor_row_batch(k, _p64(ns), P, I, X, cplx, float(drop_tol), _PIVOT_CODES[pivot_strategy],
                                           float(pivot_rel), out, rc, err, 512), err)
    res = []
    for i in range(k):
        if rc[i] != 0:
            res.append(FplusError(rc[i], f"system {i}"))
        else:
            res.append(_factors_dict(out[i], keep[i][0], cplx))
    return res


def factor_row_sym(sym, Ap, Ai, Ax, drop_tol=0.0, pivot_strategy="full", equilibrate=True, pivot_rel=FPLUS_PIVOT_REL_DEFAULT,
                   ruiz_max_iter=20, ruiz_tol=1e-2):
    """fplus_factor_row_sym: the orders of a Symbolic, Ruiz, the row kernel; CSC factors of the permuted, equilibrated
    matrix, p, q (positions of that matrix), Dr, Dc (its order)"""
    Ax = np.ascontiguousarray(Ax)
    cplx = int(np.iscomplexobj(Ax))
    n = sym.n
    f = Factors()
    Dr = np.empty(n)
    Dc = np.empty(n)
    err = C.create_string_buffer(512)
    rc = lib().fplus_factor_row_sym(sym.ptr, _p64(np.ascontiguousarray(Ap, np.int64)), _p64(np.ascontiguousarray(Ai, np.int64)),
                                    _vp(Ax), cplx, int(equilibrate), ruiz_max_iter, ruiz_tol, 0, float(drop_tol),
                                    _PIVOT_CODES[pivot_strategy], float(pivot_rel), C.byref(f), Dr.ctypes.data_as(_DP),
                                    Dc.ctypes.data_as(_DP), err, 512)
    check(rc, err)
    out = _factors_dict(f, n, cplx)
    out.update(Dr=Dr, Dc=Dc)
    return out


def factor_sn(n, Ap, Ai, Ax, relax_rows=32, pivot_rel=FPLUS_PIVOT_REL_DEFAULT):
    """fplus_factor_sn: the supernodal factorization of the matrix as given (duplicates summed, no equilibration)"""
    Ax = np.ascontiguousarray(Ax)
    cplx = int(np.iscomplexobj(Ax))
    f = Factors()
    err = C.create_string_buffer(512)
    rc = lib().fplus_factor_sn(n, _p64(np.ascontiguousarray(Ap, np.int64)), _p64(np.ascontiguousarray(Ai, np.int64)), _vp(Ax),
                               cplx, 0, relax_rows, 512, 1, 5e7, pivot_rel, C.byref(f), err, 512)
    check(rc, err)
    return _factors_dict(f, n, cplx)


def factor_sn_perm(n, Ap, Ai, Ax, row_perm, col_perm, equilibrate=True, relax_rows=32, pivot_rel=FPLUS_PIVOT_REL_DEFAULT,
                   ruiz_max_iter=20, ruiz_tol=1e-2):
    """fplus_factor_sn_perm: the factors of the permuted, equilibrated matrix as numpy arrays, and Dr, Dc (in the
    permuted order)"""
    Ax = np.ascontiguousarray(Ax)
    cplx = int(np.iscomplexobj(Ax))
    f = Factors()
    Dr = np.empty(n)
    Dc = np.empty(n)
    err = C.create_string_buffer(512)
    rc = lib().fplus_factor_sn_perm(n, _p64(np.ascontiguousarray(Ap, np.int64)), _p64(np.ascontiguousarray(Ai, np.int64)),
                                    _vp(Ax), cplx, _p64(np.ascontiguousarray(row_perm, np.int64)),
                                    _p64(np.ascontiguousarray(col_perm, np.int64)), int(equilibrate), ruiz_max_iter, ruiz_tol, 0,
                                    relax_rows, 512, 1, 5e7, pivot_rel, C.byref(f), Dr.ctypes.data_as(_DP),
                                    Dc.ctypes.data_as(_DP), err, 512)
    check(rc, err)
    out = _factors_dict(f, n, cplx)
    out.update(Dr=Dr, Dc=Dc)
    return out


def factor_sn_sym(sym, Ap, Ai, Ax, equilibrate=True, relax_rows=32, pivot_rel=FPLUS_PIVOT_REL_DEFAULT, ruiz_max_iter=20,
                  ruiz_tol=1e-2):
    """fplus_factor_sn_sym: factor_sn_perm with the orders of a Symbolic"""
    Ax = np.ascontiguousarray(Ax)
    cplx = int(np.iscomplexobj(Ax))
    n = sym.n
    f = Factors()
    Dr = np.empty(n)
    Dc = np.empty(n)
    err = C.create_string_buffer(512)
    rc = lib().fplus_factor_sn_sym(sym.ptr, _p64(np.ascontiguousarray(Ap, np.int64)), _p64(np.ascontiguousarray(Ai, np.int64)),
                                   _vp(Ax), cplx, int(equilibrate), ruiz_max_iter, ruiz_tol, 0, relax_rows, 512, 1, 5e7,
                                   pivot_rel, C.byref(f), Dr.ctypes.data_as(_DP), Dc.ctypes.data_as(_DP), err, 512)
    check(rc, err)
    out = _factors_dict(f, n, cplx)
    out.update(Dr=Dr, Dc=Dc)
    return out


def tune(name, value):
    """fplus_amd_tune: a process-wide tuning knob ("panel_lu_max_workgroups")"""
    check(lib().fplus_amd_tune(name.encode(), int(value)), None)


class SolvePlan:
    """A solve plan for caller-held CSC factors (reference _sparse_lu_c.c ``solve_plan``, src/solve_plan.c:81-207): the
    patterns of L (unit lower, rows in pivot order) and U (upper, diagonal stored) go to the device once; ``solve``
    uploads the values and the right-hand sides for each call."""

    def __init__(self, Lp, Li, Up, Ui):
        self.n = len(Lp) - 1
        self._keep = [np.ascontiguousarray(a, np.int64) for a in (Lp, Li, Up, Ui)]
        self.ptr = C.c_void_p()
        err = C.create_string_buffer(256)
        check(lib().fplus_solve_plan_create(self.n, *[_p64(a) for a in self._keep], C.byref(self.ptr), err, 256), err)

    @property
    def nsupernodes(self):
        """the supernodes of L the reference's plan finds (-1 for a factor without the diagonal first in every column)"""
        return lib().fplus_solve_plan_nsupernodes(self.ptr)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().fplus_solve_plan_free(self.ptr)
            self.ptr = None


def tri_solve(L, U, b, trans=0, plan=None):
    """x = U^-1 L^-1 b (trans 0), L^-T U^-T b (1) or L^-H U^-H b (2) for CSC factors L = (Lp, Li, Lx), U = (Up, Ui, Ux);
    b of shape (n,) or (n, k).  With a plan: fplus_tri_solve_plan (the sizes must match the plan's, else ValueError),
    without: fplus_tri_solve per column (reference sparse_lu_full_pivot.py _tri_solve_c)."""
    Lp, Li, Lx = (np.ascontiguousarray(a) for a in L)
    Up, Ui, Ux = (np.ascontiguousarray(a) for a in U)
    Lp, Li, Up, Ui = (np.ascontiguousarray(a, np.int64) for a in (Lp, Li, Up, Ui))
    cplx = np.iscomplexobj(Lx) or np.iscomplexobj(Ux) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    Lx, Ux = np.ascontiguousarray(Lx, dt), np.ascontiguousarray(Ux, dt)
    b = np.asarray(b)
    one = b.ndim == 1
    n = len(Lp) - 1
    X = np.array(b.reshape(n, -1), dtype=dt, order="F", copy=True)
    k = X.shape[1]
    if plan is not None:
        if plan.n != n or plan._keep[0][-1] != Lp[-1] or plan._keep[2][-1] != Up[-1]:
            raise ValueError("plan does not match the factors")
        rc = lib().fplus_tri_solve_plan(plan.ptr, n, _p64(Lp), _p64(Li), _vp(Lx), _p64(Up), _p64(Ui), _vp(Ux), _vp(X), k, n,
                                        int(trans), int(cplx))
        check(rc, None)
    else:
        for r in range(k):
            col = np.ascontiguousarray(X[:, r])
            check(lib().fplus_tri_solve(n, _p64(Lp), _p64(Li), _vp(Lx), _p64(Up), _p64(Ui), _vp(Ux), _vp(col), int(trans),
                                        int(cplx)), None)
            X[:, r] = col
    return X[:, 0] if one else X


from .selector import Selector, log_blocks  # noqa: E402,F401
from .splu import SparseLU, clear_ordering_cache, lu_residual, lu_solve, sparse_lu_full_pivot  # noqa: E402,F401
