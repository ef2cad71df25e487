"""The Python surface of the reference's sparse_lu_full_pivot module over the device library: ``sparse_lu_full_pivot``,
``SparseLU`` (the drop-in for the object ``scipy.sparse.linalg.splu`` returns), ``lu_solve``, ``lu_residual`` and the
ordering cache with ``clear_ordering_cache`` (reference sparse_lu_full_pivot.py:147-171, :497-628, :630-836,
:921-1007).

The keyword set and its defaults are the reference's: ``colamd=False`` (the identity orders), ``pin_dense_rows=False``,
``ordering='auto'``, ``relax_rows=0``, ``panel=1``, ``equilibrate=True`` (an int k runs at most k Ruiz iterations, 0
none), ``pivot_strategy='partial'``, ``pivot_rel=None`` (sqrt(eps)), ``cache_ordering=True``.  ``gpu`` and
``gpu_min_flops`` are accepted: the device path is the only one.  ``pivot_strategy='rook'`` / ``'full'`` and
``drop_tol > 0`` run the row kernel on the device (one workgroup per system; the reference's fplus_factor dispatch,
src/numeric.c:51-64).
"""
import collections
import hashlib
import threading

import numpy as np
import scipy.sparse as sp

from ._lib import FPLUS_ORDER_AMD_ATA, FPLUS_ORDER_AUTO, FPLUS_ORDER_COLAMD, FPLUS_PIVOT_REL_DEFAULT

_PIVOTS = {"partial": 1, "rook": 2, "full": 0}
_ORDERINGS = {"colamd": FPLUS_ORDER_COLAMD, "amd_ata": FPLUS_ORDER_AMD_ATA, "auto": FPLUS_ORDER_AUTO}
_RUIZ_MAX_ITER = 20

# ---------------------------------------------------------------- the ordering cache
# The orders depend on the pattern and the analysis settings only, so a cached entry is exact.  Keyed by a digest of the
# canonical pattern (all of its indices) and the settings; bounded by the bytes of the stored orders.
_CACHE = collections.OrderedDict()
_CACHE_LOCK = threading.Lock()
_CACHE_MAX_BYTES = 64 << 20
_cache_bytes = 0


def clear_ordering_cache():
    """Forget every cached ordering."""
    global _cache_bytes
    with _CACHE_LOCK:
        _CACHE.clear()
        _cache_bytes = 0


def _pattern_key(n, Ap, Ai, settings):
    h = hashlib.blake2b(digest_size=20)
    h.update(np.ascontiguousarray(Ap, np.int64).tobytes())
    h.update(np.ascontiguousarray(Ai, np.int64).tobytes())
    for m in settings[-2:]:   # the masks, tagged so that a row mask cannot stand for a column mask
        h.update(b"mask:" + (b"none" if m is None else np.packbits(np.asarray(m, dtype=bool)).tobytes()))
    return (n, int(Ap[-1]), h.hexdigest()) + tuple(settings[:-2])


def _orders(n, Ap, Ai, pin_dense_rows, tau_density, tau_min_nnz, tau_row_mask, tau_col_mask, postorder, ordering, cache):
    """(row_perm, col_perm, n_dense_rows, n_dense_cols) of fplus_analyze_ord, through the cache"""
    global _cache_bytes
    from . import Symbolic
    settings = (bool(pin_dense_rows), float(tau_density), int(tau_min_nnz), bool(postorder), ordering, tau_row_mask, tau_col_mask)
    key = _pattern_key(n, Ap, Ai, settings) if cache else None
    if key is not None:
        with _CACHE_LOCK:
            hit = _CACHE.get(key)
            if hit is not None:
                _CACHE.move_to_end(key)
                return hit[0].copy(), hit[1].copy(), hit[2], hit[3]
    s = Symbolic(n, Ap, Ai, pin_dense=pin_dense_rows, tau_density=tau_density, tau_min_nnz=tau_min_nnz,
                 tau_row_mask=tau_row_mask, tau_col_mask=tau_col_mask, postorder=postorder, ordering=_ORDERINGS[ordering])
    out = (s.row_perm.copy(), s.col_perm.copy(), int(s.n_dense_rows), int(s.n_dense_cols))
    if key is not None:
        with _CACHE_LOCK:
            old = _CACHE.pop(key, None)
            if old is not None:
                _cache_bytes -= old[0].nbytes + old[1].nbytes
            _CACHE[key] = out
            _cache_bytes += out[0].nbytes + out[1].nbytes
            while len(_CACHE) > 1 and _cache_bytes > _CACHE_MAX_BYTES:
                _, ev = _CACHE.popitem(last=False)
                _cache_bytes -= ev[0].nbytes + ev[1].nbytes
    return out[0].copy(), out[1].copy(), out[2], out[3]


# ---------------------------------------------------------------- the factorization
def _factor(A, drop_tol=0.0, equilibrate=True, colamd=False, pin_dense_rows=False, tau_density=0.5, tau_min_nnz=10,
            tau_row_mask=None, tau_col_mask=None, pivot_strategy="partial", dtype=None, verbose=0, relax_rows=0, panel=1,
            cache_ordering=True, gpu=False, gpu_min_flops=5e7, pivot_rel=None, postorder=True, ordering="auto"):
    """validated arguments -> (Numeric, Symbolic, n, Ap, Ai, Ax)"""
    from . import Numeric, _canonical_csc, default_options, symbolic_from_orders
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"A must be square; got shape {A.shape}.")
    if pivot_strategy not in _PIVOTS:
        raise ValueError(f"pivot_strategy must be 'partial', 'rook', or 'full'; got {pivot_strategy!r}.")
    if ordering not in _ORDERINGS:
        raise ValueError(f"ordering must be 'colamd', 'amd_ata' or 'auto'; got {ordering!r}.")
    if not drop_tol >= 0.0:
        raise ValueError(f"drop_tol must be >= 0; got {drop_tol!r}.")
    n, Ap, Ai, Ax = _canonical_csc(A)
    cplx = np.iscomplexobj(Ax)
    if dtype is not None:
        dt = np.dtype(dtype)
        if dt == np.complex128:
            Ax = np.ascontiguousarray(Ax, np.complex128)
        elif dt == np.float64:
            if cplx:
                raise ValueError("dtype float64 for a complex matrix")
        else:
            raise ValueError(f"dtype must be float64 or complex128; got {dt}")
    iters = _RUIZ_MAX_ITER if equilibrate is True else (0 if equilibrate is False else int(equilibrate))
    if colamd:
        rp, cp, _, _ = _orders(n, Ap, Ai, pin_dense_rows, tau_density, tau_min_nnz, tau_row_mask, tau_col_mask, postorder,
                               ordering, cache_ordering)
    else:
        rp = cp = np.arange(n, dtype=np.int64)
    sym = symbolic_from_orders(n, Ap, Ai, rp, cp)
    o = default_options(equilibrate=int(iters > 0), ruiz_max_iter=max(iters, 1), relax_rows=int(relax_rows),
                        panel=int(panel), verbose=int(verbose), use_gpu=1, gpu_min_flops=float(gpu_min_flops),
                        pivot_rel=FPLUS_PIVOT_REL_DEFAULT if pivot_rel is None else float(pivot_rel),
                        pivot_strategy=_PIVOTS[pivot_strategy], drop_tol=float(drop_tol))
    num = Numeric(sym, Ap, Ai, Ax, o)
    if pivot_strategy != "partial" or drop_tol != 0.0:
        nrep = num.factors()["n_pivots_replaced"]
        if nrep > 0:
            import warnings
            warnings.warn(f"ILU breakdown: {nrep} zero or negligible pivots replaced by drop_tol * max|a_ij| (the exact "
                          f"factorization's trailing pivots are at or below drop_tol={drop_tol:g}); the factors are a "
                          "preconditioner of reduced quality", RuntimeWarning, stacklevel=3)
    return num, sym, n, Ap, Ai, Ax


def sparse_lu_full_pivot(A, *, _return_csc=False, **kw):
    """(L, U, p, q, Dr, Dc) with diag(Dr) A[p, :][:, q] diag(Dc) = L U in the original coordinates (p, q, Dr, Dc
    composed as the reference does, sparse_lu_full_pivot.py:453-493); L and U CSR, or CSC with _return_csc=True"""
    num, _, _, _, _, _ = _factor(A, **kw)
    f = num.factors()
    fmt = (lambda M: M) if _return_csc else (lambda M: M.tocsr())   # noqa: E731
    return fmt(f["L"]), fmt(f["U"]), num.p.copy(), num.q.copy(), num.Dr.copy(), num.Dc.copy()


def lu_solve(L, U, p, q, b, Dr=None, Dc=None):
    """x with A x = b from sparse_lu_full_pivot's factors (reference sparse_lu_full_pivot.py:497-545): b_eq = Dr[p] b[p],
    z = U^-1 L^-1 b_eq (the device sweeps on the CSC factors), x[q] = Dc[q] z; b of shape (n,) or (n, k)"""
    from . import tri_solve
    L = sp.csc_matrix(L)
    U = sp.csc_matrix(U)
    b = np.asarray(b)
    dt = np.result_type(L.dtype, b)
    bp = np.asarray(b, dtype=dt)[p]
    if Dr is not None:
        bp = (Dr[p] * bp.T).T
    z = tri_solve((L.indptr, L.indices, L.data), (U.indptr, U.indices, U.data), bp, 0)
    x = np.empty_like(z)
    x[q] = (Dc[q] * z.T).T if Dc is not None else z
    return x


def lu_residual(A, L, U, p, q, Dr=None, Dc=None):
    """||diag(Dr) A[p, :][:, q] diag(Dc) - L U||_F / ||diag(Dr) A[p, :][:, q] diag(Dc)||_F (reference
    sparse_lu_full_pivot.py:598-628), with sparse products"""
    A = sp.csc_matrix(A)
    if Dr is not None:
        A = sp.diags(Dr) @ A
    if Dc is not None:
        A = A @ sp.diags(Dc)
    PAQ = sp.csc_matrix(A)[p, :][:, q]
    D = PAQ - sp.csc_matrix(L) @ sp.csc_matrix(U)
    nrm = sp.linalg.norm(PAQ, "fro")
    return float(sp.linalg.norm(D, "fro") / (nrm if nrm > 0 else 1.0))


class SparseLU:
    """Drop-in for the object ``scipy.sparse.linalg.splu`` returns (reference sparse_lu_full_pivot.py:630-836).

    ``SparseLU(A, **kw)`` factors A on the GPU with sparse_lu_full_pivot's keyword arguments and defaults.  Attributes:
    ``L``, ``U`` (CSC, of the scaled permuted matrix), ``perm_r``, ``perm_c``, ``shape``, ``nnz``, and (beyond splu's)
    ``Dr``, ``Dc``.  ``solve(b, trans='N'|'T'|'H')`` for b of shape (n,) or (n, k): the device's solves on the factors it
    holds."""

    def __init__(self, A, **kw):
        self._num, self._sym, self._n, _, _, _ = _factor(A, **kw)
        self._f = None
        self._is_complex = self._num.is_complex

    def _factors(self):
        if self._f is None:
            self._f = self._num.factors()
        return self._f

    @property
    def L(self):
        return self._factors()["L"]

    @property
    def U(self):
        return self._factors()["U"]

    @property
    def perm_r(self):
        return self._num.p

    @property
    def perm_c(self):
        return self._num.q

    @property
    def Dr(self):
        return self._num.Dr

    @property
    def Dc(self):
        return self._num.Dc

    @property
    def shape(self):
        return (self._n, self._n)

    @property
    def nnz(self):
        f = self._factors()
        return f["L"].nnz + f["U"].nnz

    def solve(self, b, trans="N"):
        if trans not in ("N", "T", "H"):
            raise ValueError(f"trans must be 'N', 'T', or 'H'; got {trans!r}")
        tr = 0 if trans == "N" else (1 if trans == "T" or not self._is_complex else 2)
        return self._num.solve(b, tr)
