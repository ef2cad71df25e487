"""The oracle's row kernel (full / partial / rook pivoting, incomplete LU) pinned on the CPU: against the reference's own
src/factor_row.c compiled from its sources (oracle/_ref, bit for bit), against the reference's known answers
(tests/c/test_fplus.c:204-229), and against what the mathematics fixes -- P A Q = L U for the exact factorization,
the pivot bounds each strategy guarantees, and partial pivoting's pivots against LAPACK's GEPP on dense matrices."""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp

import oracle
from helpers import csr_arrays as csr, row_kernel_cases

STRATS = {0: "full", 1: "partial", 2: "rook"}


def dense_factors(F, n):
    L = sp.csr_matrix((F["Lx"], F["Li"], F["Lp"]), shape=(n, n)).toarray()
    U = sp.csr_matrix((F["Ux"], F["Ui"], F["Up"]), shape=(n, n)).toarray()
    return L, U


CASES = row_kernel_cases()


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref is not built (the reference tree is absent)")
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("strategy", [0, 1, 2], ids=["full", "partial", "rook"])
@pytest.mark.parametrize("drop_tol", [0.0, 1e-3, 1e-2])
def test_oracle_row_kernel_is_the_reference(case, strategy, drop_tol):
    _, n, Rp, Ri, Rx = case
    try:
        a = oracle.factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
    except oracle.OracleError as e:
        with pytest.raises(oracle.OracleError) as r:
            oracle.ref_factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
        assert r.value.args == e.args
        return
    b = oracle.ref_factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
    for key in ("p", "q", "Lp", "Li", "Up", "Ui"):
        assert np.array_equal(a[key], b[key]), key
    assert np.array_equal(a["Lx"], b["Lx"]) and np.array_equal(a["Ux"], b["Ux"])
    assert a["n_replaced"] == b["n_replaced"]


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("strategy", [0, 1, 2], ids=["full", "partial", "rook"])
def test_exact_row_factors_reproduce_the_matrix(case, strategy):
    """drop_tol 0: A[p][:, q] = L U, L unit lower with |L| <= 1, U upper; rook and full pivots dominate their U rows"""
    _, n, Rp, Ri, Rx = case
    F = oracle.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
    A = sp.csr_matrix((Rx, Ri, Rp), shape=(n, n)).toarray()
    L, U = dense_factors(F, n)
    assert np.allclose(np.tril(L), L) and np.allclose(np.triu(U), U)
    assert np.all(np.diag(L) == 1)
    PAQ = A[F["p"]][:, F["q"]]
    assert np.abs(PAQ - L @ U).max() <= 1e-12 * np.abs(A).max() * n
    assert np.abs(L).max() <= 1 + 1e-14
    if strategy != 1:
        d = np.abs(np.diag(U))
        assert np.all(np.abs(U).max(axis=1) <= d * (1 + 1e-14))
    else:
        assert np.array_equal(F["q"], np.arange(n))


def test_partial_pivots_are_lapacks_on_dense_matrices():
    """on dense random real matrices partial pivoting takes, step by step, the first largest entry of the column, as
    LAPACK's dgetrf does: the same row order and the same factors to rounding (zgetrf ranks complex entries by
    |re| + |im|, not by the modulus, so complex matrices are pinned by the bounds above instead)"""
    rng = np.random.default_rng(11)
    for n in (40, 64):
        A = rng.standard_normal((n, n))
        F = oracle.factor_row(*csr(A), 0.0, 1)
        P, Ll, Uu = sla.lu(A)
        perm = np.argmax(P, axis=0)          # A = P L U: row k of L U is row perm[k] of A
        assert np.array_equal(F["p"], perm)
        L, U = dense_factors(F, n)
        assert np.abs(L - Ll).max() < 1e-12 and np.abs(U - Uu).max() < 1e-12 * np.abs(Uu).max()


def test_ilu_drops_and_the_reference_breakdown_case():
    """[[1, 1], [1, 1.0005]] with full pivoting and drop_tol 1e-2: the second pivot cancels below drop_tol and is
    replaced by drop_tol * max|a_ij| = 0.010005 (test_fplus.c:204-222); [[1, 1], [1, 1]] without dropping is singular"""
    A = np.array([[1.0, 1.0], [1.0, 1.0005]])
    F = oracle.factor_row(*csr(A), 1e-2, 0)
    assert F["n_replaced"] == 1
    _, U = dense_factors(F, 2)
    assert abs(U[1, 1] - 0.010005) < 1e-15 and U[0, 0] == 1.0005
    with pytest.raises(oracle.OracleError) as e:
        oracle.factor_row(*csr(np.ones((2, 2))), 0.0, 0)
    assert e.value.code == -2 and "Matrix is singular" in str(e.value)
    # an ILU keeps less fill than the exact factorization and leaves no kept entry at or below drop_tol off the diagonal
    _, n, Rp, Ri, Rx = CASES[2]
    ex = oracle.factor_row(n, Rp, Ri, Rx, 0.0, 1)
    il = oracle.factor_row(n, Rp, Ri, Rx, 1e-2, 1)
    assert il["Lp"][-1] + il["Up"][-1] < ex["Lp"][-1] + ex["Up"][-1]
    Ui = np.repeat(np.arange(n), np.diff(il["Up"]))
    off = il["Ui"] != Ui
    assert np.all(np.abs(il["Ux"][off]) > 1e-2)
    Li = np.repeat(np.arange(n), np.diff(il["Lp"]))
    offl = il["Li"] != Li
    assert np.all(np.abs(il["Lx"][offl]) > 1e-2)


def test_rook_and_full_swap_columns_partial_does_not():
    n, Rp, Ri, Rx = CASES[0][1:]
    for s in (0, 2):
        F = oracle.factor_row(n, Rp, Ri, Rx, 0.0, s)
        assert not np.array_equal(F["q"], np.arange(n)), STRATS[s]
        assert np.array_equal(np.sort(F["q"]), np.arange(n))
