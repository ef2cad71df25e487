"""Backward errors of a computed solution, three ways.

* ``normwise``: the reference tests' measure, ||b - A x||_2 / (||A||_1 ||x||_2 + ||b||_2), ||A||_1 the largest absolute
  column sum (reference tests/c/test_fplus.c:74-89, test_sparse_lu_full_pivot.py:759-760).  It is the headline
  metric's stated measure, but on matrices whose rows and columns are scaled over many orders ||A||_1 is set by a few
  huge entries, and a solution with a large error in its small-scale components still scores far below 1e-13.
* ``componentwise``: the Oettli-Prager measure max_i |b - A x|_i / (|A| |x| + |b|)_i, invariant under row and
  column scaling (a row with a zero denominator and a nonzero residual scores inf).
* ``equilibrated``: the normwise measure of the equilibrated system the factorization solves: for A x = b that is
  (Dr A Dc) y = Dr b with y = x / Dc, so ||Dr (b - A x)||_2 / (||Dr A Dc||_1 ||x / Dc||_2 + ||Dr b||_2); for A^T x = b
  (and A^H) the roles of Dr and Dc swap.  Invariant under the row and column scaling Ruiz removes.

All three take a scipy sparse A, and x, b of shape (n,) or (n, k) (the maximum over the columns is returned).
"""
import numpy as np
import scipy.sparse as sp


def _op(A, trans):
    if trans in (0, "N"):
        return A
    if trans in (1, "T"):
        return A.T.tocsc()
    return A.conj().T.tocsc()


def _cols(a):
    a = np.asarray(a)
    return a.reshape(a.shape[0], -1)


def normwise(A, x, b, trans=0):
    """max over the right-hand sides of ||b - op(A) x|| / (||A||_1 ||x|| + ||b||)"""
    M = _op(A, trans)
    X, B = _cols(x), _cols(b)
    nA = float(abs(A).sum(axis=0).max()) if A.nnz else 0.0
    R = B - M @ X
    out = 0.0
    for k in range(X.shape[1]):
        den = nA * np.linalg.norm(X[:, k]) + np.linalg.norm(B[:, k])
        out = max(out, float(np.linalg.norm(R[:, k]) / den) if den > 0 else 0.0)
    return out


def componentwise(A, x, b, trans=0):
    """max over the right-hand sides and rows of |b - op(A) x|_i / (|op(A)| |x| + |b|)_i"""
    M = _op(A, trans)
    X, B = _cols(x), _cols(b)
    R = abs(B - M @ X)
    D = abs(M) @ abs(X) + abs(B)
    with np.errstate(divide="ignore", invalid="ignore"):
        q = np.where(D > 0, R / np.where(D > 0, D, 1.0), np.where(R > 0, np.inf, 0.0))
    return float(q.max()) if q.size else 0.0


def equilibrated(A, x, b, Dr, Dc, trans=0):
    """the normwise backward error of the system scaled by the factorization's Dr, Dc (a scale-invariant measure)"""
    left, right = (Dr, Dc) if trans in (0, "N") else (Dc, Dr)
    M = _op(A, trans)
    X, B = _cols(x), _cols(b)
    Ab = sp.diags(left) @ M @ sp.diags(right)
    nA = float(abs(Ab).sum(axis=0).max()) if Ab.nnz else 0.0
    R = left[:, None] * (B - M @ X)
    Y = X / right[:, None]
    Bs = left[:, None] * B
    out = 0.0
    for k in range(X.shape[1]):
        den = nA * np.linalg.norm(Y[:, k]) + np.linalg.norm(Bs[:, k])
        out = max(out, float(np.linalg.norm(R[:, k]) / den) if den > 0 else 0.0)
    return out


def all_measures(A, x, b, Dr, Dc, trans=0):
    """the three measures as a dict"""
    return {"normwise": normwise(A, x, b, trans), "componentwise": componentwise(A, x, b, trans),
            "equilibrated": equilibrated(A, x, b, Dr, Dc, trans)}
