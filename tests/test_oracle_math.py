"""The oracle pinned by mathematics where the reference's fixtures leave it open: brute-force dense GEPP (modulus
argmax, the lowest row on ties) on small matrices, the factorization invariants, Ruiz's known answers and its
convergence property, the permutation against scipy indexing, the solves against scipy's sparse solver."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import oracle as O
from fplus_amd import gen


def dense_gepp(A):
    """textbook right-looking LU with partial pivoting by modulus; returns p (row of A pivoted at step j), L, U"""
    A = np.array(A, dtype=complex if np.iscomplexobj(A) else float)
    n = A.shape[0]
    rows = list(range(n))
    L = np.zeros_like(A)
    U = np.zeros_like(A)
    for j in range(n):
        mod = np.abs(A[j:, j])
        k = j + int(np.argmax(mod))
        A[[j, k]] = A[[k, j]]
        L[[j, k]] = L[[k, j]]
        rows[j], rows[k] = rows[k], rows[j]
        U[j, j:] = A[j, j:]
        L[j, j] = 1
        L[j + 1:, j] = A[j + 1:, j] / A[j, j]
        A[j + 1:, j:] -= np.outer(L[j + 1:, j], A[j, j:])
    return np.array(rows), L, U


def random_sparse(n, density, cplx, seed):
    rng = np.random.default_rng(seed)
    M = sp.random(n, n, density=density, random_state=rng, format="csc")
    if cplx:
        M = M + 1j * sp.random(n, n, density=density, random_state=rng, format="csc")
    M = M + sp.diags(0.1 + rng.random(n))
    M = sp.csc_matrix(M)
    M.sort_indices()
    return n, M.indptr.astype(np.int64), M.indices.astype(np.int64), np.ascontiguousarray(M.data)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("relax", [0, 3])
def test_factor_equals_dense_gepp(seed, cplx, relax):
    n, Ap, Ai, Ax = random_sparse(24, 0.12, cplx, seed)
    f = O.factor_sn(n, Ap, Ai, Ax, relax_rows=relax)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)).toarray()
    rows, L, U = dense_gepp(A)
    assert np.array_equal(f["p"], rows)
    Ls = sp.csc_matrix((f["Lx"], f["Li"], f["Lp"]), shape=(n, n)).toarray()
    Us = sp.csc_matrix((f["Ux"], f["Ui"], f["Up"]), shape=(n, n)).toarray()
    assert np.max(np.abs(Ls - L)) < 1e-12 * max(1.0, np.max(np.abs(L)))
    assert np.max(np.abs(Us - U)) < 1e-12 * max(1.0, np.max(np.abs(U)))


@pytest.mark.parametrize("mk", [lambda: gen.convdiff2d(20), lambda: gen.shift_invert(2, 4, 32), lambda: gen.convdiff3d(7)],
                         ids=["cd2d", "si", "cd3d"])
def test_factor_invariants(mk):
    n, Ap, Ai, Ax = mk()
    rp = np.random.default_rng(1).permutation(n)
    cp = np.random.default_rng(2).permutation(n)
    num = O.Numeric(n, Ap, Ai, Ax, rp, cp, relax_rows=8)
    f = num.lu
    L = sp.csc_matrix((f["Lx"], f["Li"], f["Lp"]), shape=(n, n))
    U = sp.csc_matrix((f["Ux"], f["Ui"], f["Up"]), shape=(n, n))
    assert np.all(f["Li"][f["Lp"][:-1]] == np.arange(n)), "unit diagonal first in every L column"
    assert np.all(f["Lx"][f["Lp"][:-1]] == 1)
    assert sp.triu(L, 1).nnz == 0 and sp.tril(U, -1).nnz == 0
    assert np.max(np.abs(L.data)) <= 1.0 + 1e-15, "partial pivoting bounds |L| by 1"
    assert np.all(f["Ui"][f["Up"][1:] - 1] == np.arange(n)), "the diagonal last in every U column"
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    B = sp.diags(num.Dr[num.p]) @ A[num.p, :][:, num.q] @ sp.diags(num.Dc[num.q])
    r = sp.linalg.norm(B - L @ U) / sp.linalg.norm(B)
    assert r < 1e-13


def test_ruiz_known_answers():
    """the reference's range-guard test (tests/c/test_fplus.c:249-266)"""
    n = 3
    p = np.array([0, 1, 2, 3], np.int64)
    i = np.array([0, 1, 2], np.int64)
    x0 = np.array([1e200, 1e100, 1e150])
    x, Dr, Dc, _ = O.ruiz(n, p, i, x0)
    assert np.all(np.abs(np.log(np.abs(x))) < 1e-2)
    assert np.all(np.abs(Dr * x0 * Dc - x) < 1e-14 * np.abs(x))
    with pytest.raises(O.OracleError) as e:
        O.ruiz(n, p, i, np.array([1e300, 1e-300, 1.0]))
    assert e.value.code == -3


@pytest.mark.parametrize("mk", [lambda: gen.convdiff2d(16), lambda: gen.shift_invert(2, 4, 32)], ids=["cd2d", "si"])
def test_ruiz_converges(mk):
    n, Ap, Ai, Ax = mk()
    x, Dr, Dc, it = O.ruiz(n, Ap, Ai, Ax, tol=1e-2)
    B = abs(sp.csc_matrix((x, Ai, Ap), shape=(n, n)))
    rmax = B.max(axis=1).toarray().ravel()
    cmax = B.max(axis=0).toarray().ravel()
    assert it < 20 and np.all(np.abs(np.log(rmax)) < 2e-2) and np.all(np.abs(np.log(cmax)) < 2e-2)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    assert sp.linalg.norm(sp.diags(Dr) @ A @ sp.diags(Dc) - sp.csc_matrix((x, Ai, Ap), shape=(n, n))) < 1e-13 * sp.linalg.norm(B)


@pytest.mark.parametrize("mk", [lambda: gen.convdiff2d(16), lambda: gen.shift_invert(2, 4, 32)], ids=["cd2d", "si"])
def test_ruiz_csr_route_equals_csc(mk):
    """the reference's CSR route (src/equil.c:186-441, lines = rows) and its CSC route scale the same matrix the same way:
    the same Dr and Dc bit for bit (maxima do not depend on the order they are taken in) and the same scaled entries"""
    n, Ap, Ai, Ax = mk()
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    R = A.tocsr()
    R.sort_indices()
    xc, Drc, Dcc, itc = O.ruiz(n, Ap, Ai, Ax)
    xr, Drr, Dcr, itr = O.ruiz(n, R.indptr.astype(np.int64), R.indices.astype(np.int64), R.data, is_csc=False)
    assert itc == itr
    assert np.array_equal(Drc, Drr) and np.array_equal(Dcc, Dcr)
    Bc = sp.csc_matrix((xc, Ai, Ap), shape=(n, n))
    Br = sp.csr_matrix((xr, R.indices, R.indptr), shape=(n, n))
    assert abs(Bc - Br).max() == 0.0


def test_ruiz_zero_line_is_singular():
    n = 3
    p = np.array([0, 1, 2, 2], np.int64)
    i = np.array([0, 1], np.int64)
    with pytest.raises(O.OracleError) as e:
        O.ruiz(n, p, i, np.array([1.0, 2.0]))
    assert e.value.code == -2


def test_permute_matches_scipy():
    n, Ap, Ai, Ax = gen.shift_invert(2, 4, 32)
    rp = np.random.default_rng(4).permutation(n)
    cp = np.random.default_rng(5).permutation(n)
    Bp, Bi, Bx = O.permute(n, Ap, Ai, Ax, rp, cp)
    ref = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))[rp, :][:, cp].tocsc()
    ref.sort_indices()
    assert np.array_equal(Bp, ref.indptr) and np.array_equal(Bi, ref.indices) and np.array_equal(Bx, ref.data)


@pytest.mark.parametrize("cplx", [False, True])
def test_solve_matches_scipy(cplx):
    n, Ap, Ai, Ax = gen.reference_test_matrix(120, cplx, 3.0)
    rp = np.random.default_rng(6).permutation(n)
    num = O.Numeric(n, Ap, Ai, Ax, rp, np.arange(n), relax_rows=4)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = np.random.default_rng(8).standard_normal(n) + (1j if cplx else 0) * np.random.default_rng(9).standard_normal(n)
    for tr, M in ((0, A), (1, A.T), (2, A.conj().T)):
        if tr == 2 and not cplx:
            continue
        x = num.solve(b, tr)
        xr = sla.spsolve(sp.csc_matrix(M), b)
        assert np.linalg.norm(x - xr) / np.linalg.norm(xr) < 1e-9
        assert O.backward_error(n, Ap, Ai, Ax, x, b, tr) < 1e-13


def test_singular_column():
    """a structurally zero pivot column after elimination: SINGULAR (src/factor_sn.c:318-326)"""
    A = sp.csc_matrix(np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 1.0]]))
    n, Ap, Ai, Ax = 3, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    with pytest.raises(O.OracleError) as e:
        O.factor_sn(n, Ap, Ai, Ax, relax_rows=0)
    assert e.value.code == -2
