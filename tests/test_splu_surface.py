"""The reference module's Python surface (fplus_amd.splu) on the CPU: keyword defaults, argument errors raised before
any device work, and the ordering cache (reference sparse_lu_full_pivot.py:147-171, :921-1007)."""
import inspect

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import gen, splu

# the reference's sparse_lu_full_pivot defaults (sparse_lu_full_pivot.py:147-171)
REFERENCE_DEFAULTS = dict(drop_tol=0.0, equilibrate=True, colamd=False, pin_dense_rows=False, tau_density=0.5, tau_min_nnz=10,
                          tau_row_mask=None, tau_col_mask=None, pivot_strategy="partial", dtype=None, verbose=0, relax_rows=0,
                          panel=1, cache_ordering=True, gpu=False, gpu_min_flops=5e7, pivot_rel=None, postorder=True,
                          ordering="auto")


def test_keyword_defaults_are_the_references():
    sig = inspect.signature(splu._factor)
    got = {k: v.default for k, v in sig.parameters.items() if k != "A"}
    assert got == REFERENCE_DEFAULTS


def test_argument_errors():
    A = sp.random(10, 12, density=0.3, format="csc")
    with pytest.raises(ValueError):
        fplus_amd.SparseLU(A)
    n, Ap, Ai, Ax = gen.convdiff2d(6)
    B = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    with pytest.raises(ValueError):
        fplus_amd.SparseLU(B, pivot_strategy="diagonal")
    with pytest.raises(ValueError):
        fplus_amd.SparseLU(B, ordering="metis")
    with pytest.raises(ValueError):
        fplus_amd.SparseLU(B, drop_tol=-1e-3)


def test_ordering_cache():
    n, Ap, Ai, Ax = gen.convdiff2d(12)
    splu.clear_ordering_cache()
    a = splu._orders(n, Ap, Ai, False, 0.5, 10, None, None, True, "colamd", True)
    assert len(splu._CACHE) == 1
    b = splu._orders(n, Ap, Ai, False, 0.5, 10, None, None, True, "colamd", True)
    assert len(splu._CACHE) == 1 and np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    s = fplus_amd.Symbolic(n, Ap, Ai, pin_dense=False)
    assert np.array_equal(a[1], s.col_perm) and np.array_equal(a[0], s.row_perm)
    splu._orders(n, Ap, Ai, False, 0.5, 10, None, None, False, "colamd", True)   # other settings: another entry
    assert len(splu._CACHE) == 2
    mask = np.zeros(n, bool)
    mask[3] = True
    r1 = splu._orders(n, Ap, Ai, True, 0.5, 10, mask, None, True, "colamd", True)
    r2 = splu._orders(n, Ap, Ai, True, 0.5, 10, None, mask, True, "colamd", True)
    assert len(splu._CACHE) == 4 and r1[2] == 1 and r2[3] == 1   # a row mask and a column mask are different keys
    splu.clear_ordering_cache()
    assert len(splu._CACHE) == 0


def test_lu_residual_closed_form():
    """lu_residual of exact factors is 0, and of a perturbed L its Frobenius norm ratio"""
    rng = np.random.default_rng(0)
    n = 8
    Lm = np.tril(rng.standard_normal((n, n)), -1) + np.eye(n)
    Um = np.triu(rng.standard_normal((n, n))) + 4 * np.eye(n)
    p, q = rng.permutation(n), rng.permutation(n)
    Dr, Dc = 10.0 ** rng.uniform(-2, 2, n), 10.0 ** rng.uniform(-2, 2, n)
    M = Lm @ Um                                   # = diag(Dr) A[p][:, q] diag(Dc)
    A = np.empty((n, n))
    A[np.ix_(p, q)] = M
    A = A / Dr[:, None] / Dc[None, :]
    assert splu.lu_residual(sp.csc_matrix(A), sp.csc_matrix(Lm), sp.csc_matrix(Um), p, q, Dr, Dc) < 1e-14
    E = np.zeros((n, n))
    E[5, 2] = 1e-3
    r = splu.lu_residual(sp.csc_matrix(A), sp.csc_matrix(Lm + E), sp.csc_matrix(Um), p, q, Dr, Dc)
    assert np.isclose(r, np.linalg.norm(E @ Um) / np.linalg.norm(M))


def test_ordering_modes():
    """FPLUS_ORDER_AUTO keeps the COLAMD-formulation order (the reference's AMD alternative is not in this build) and
    FPLUS_ORDER_AMD_ATA is refused with FPLUS_ERR_BADARG, not silently replaced"""
    from fplus_amd._lib import FPLUS_ERR_BADARG
    n, Ap, Ai, _ = gen.convdiff2d(12)
    s0 = fplus_amd.Symbolic(n, Ap, Ai, ordering=fplus_amd.FPLUS_ORDER_COLAMD)
    s2 = fplus_amd.Symbolic(n, Ap, Ai, ordering=fplus_amd.FPLUS_ORDER_AUTO)
    assert np.array_equal(s0.col_perm, s2.col_perm) and np.array_equal(s0.row_perm, s2.row_perm)
    with pytest.raises(fplus_amd.FplusError) as e:
        fplus_amd.Symbolic(n, Ap, Ai, ordering=fplus_amd.FPLUS_ORDER_AMD_ATA)
    assert e.value.code == FPLUS_ERR_BADARG and "AMD" in str(e.value)
