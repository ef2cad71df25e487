"""Edge cases of the device path that the reference's own C tests hold (tests/c/test_fplus.c).

- Ruiz's range guard (`test_ruiz_range_guard`, test_fplus.c:249-266): entries near 1e200 are rescaled by an exact power
  of two instead of squaring to inf, and a range beyond double precision is a clear error -- on the device, bit-exact
  against the oracle;
- singular matrices: a zero column (Ruiz's zero-line test, src/equil.c:150-180) and a pivot column that elimination
  reduces to round-off (src/factor_sn.c:311-327, threshold pivot_rel * max |A(:, j)|) both give FPLUS_ERR_SINGULAR with a
  message, as the oracle does, and the next factorization in the same process is unaffected;
- `fplus_solve` with ldb > n and more right-hand sides than one device pass holds (the reference solves in groups of
  32, src/numeric.c:86-125): every column equals its own one-column solve bit for bit and the rows past n are untouched;
- the reference's high-level case with unscaled rows (`test_high_level(1, 0.0, PARTIAL)`, test_fplus.c:283);
- factorizations and solves on two host threads at once (include/fplus.h:4-5).
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import gen
from fplus_amd._lib import FPLUS_ERR_BADARG, FPLUS_ERR_SINGULAR, FplusError, lib
from helpers import norm_diff, sorted_csc

pytestmark = pytest.mark.gpu


def _bwd(A, x, b, op):
    nA = abs(A).sum(axis=0).max()
    M = {0: A, 1: A.T, 2: A.conj().T}[op]
    return float(np.linalg.norm(M @ x - b) / (nA * np.linalg.norm(x) + np.linalg.norm(b)))


def test_ruiz_range_guard_device():
    n = 3
    p = np.arange(4, dtype=np.int64)
    i = np.arange(3, dtype=np.int64)
    x0 = np.array([1e200, 1e100, 1e150])
    x, Dr, Dc = fplus_amd.ruiz(n, p, i, x0)
    xo, Dro, Dco, _ = O.ruiz(n, p, i, x0)
    assert np.array_equal(x, xo) and np.array_equal(Dr, Dro) and np.array_equal(Dc, Dco)
    assert np.all(np.abs(np.log(np.abs(x))) < 1e-2)
    assert np.all(np.abs(Dr * x0 * Dc - x) < 1e-14 * np.abs(x))
    with pytest.raises(FplusError) as e:
        fplus_amd.ruiz(n, p, i, np.array([1e300, 1e-300, 1.0]))
    assert e.value.code == FPLUS_ERR_BADARG


def _factor_device(n, Ap, Ai, Ax, relax=4):
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    return sym, fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=relax))


def _good_factor_after(cplx):
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, cplx, 6.0)
    sym, num = _factor_device(n, Ap, Ai, Ax)
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=4)
    assert np.array_equal(num.p, ref.p)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = np.random.default_rng(3).standard_normal(n)
    assert _bwd(A, num.solve(b), b, 0) < 1e-12


def test_zero_column_is_singular():
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, True, 6.0)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)).tolil()
    A[:, 150] = 0
    A = A.tocsc()
    A.eliminate_zeros()
    A.sort_indices()
    Bp, Bi, Bx = A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    sym = fplus_amd.Symbolic(n, Bp, Bi)
    with pytest.raises(O.OracleError) as eo:
        O.Numeric(n, Bp, Bi, Bx, sym.row_perm, sym.col_perm, relax_rows=4)
    with pytest.raises(FplusError) as e:
        fplus_amd.Numeric(sym, Bp, Bi, Bx, fplus_amd.default_options(relax_rows=4))
    assert e.value.code == FPLUS_ERR_SINGULAR == eo.value.code
    assert "singular" in str(e.value).lower()
    _good_factor_after(True)


@pytest.mark.parametrize("cplx", [False, True])
def test_dependent_columns_are_singular(cplx):
    """column 200 a copy of column 20: Ruiz scales both alike, and whichever of them is eliminated second is reduced to
    round-off, far below pivot_rel * its column maximum"""
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, cplx, 6.0)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)).tolil()
    A[:, 200] = A[:, 20]
    A = A.tocsc()
    A.sort_indices()
    Bp, Bi, Bx = A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    sym = fplus_amd.Symbolic(n, Bp, Bi)
    with pytest.raises(O.OracleError) as eo:
        O.Numeric(n, Bp, Bi, Bx, sym.row_perm, sym.col_perm, relax_rows=4)
    with pytest.raises(FplusError) as e:
        fplus_amd.Numeric(sym, Bp, Bi, Bx, fplus_amd.default_options(relax_rows=4))
    assert e.value.code == FPLUS_ERR_SINGULAR == eo.value.code
    assert "singular" in str(e.value).lower()
    _good_factor_after(cplx)


def test_small_exactly_singular():
    """the oracle's 3 x 3 case: the second pivot column eliminates to an exact zero"""
    A = sp.csc_matrix(np.array([[1.0, 2.0, 0.0], [2.0, 4.0, 0.0], [0.0, 0.0, 1.0]]))
    n, Ap, Ai, Ax = 3, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    sym = fplus_amd.symbolic_from_orders(n, Ap, Ai, np.arange(n), np.arange(n))
    with pytest.raises(FplusError) as e:
        fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=0))
    assert e.value.code == FPLUS_ERR_SINGULAR


@pytest.mark.parametrize("cplx", [False, True])
def test_solve_many_rhs_with_ldb(cplx):
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, cplx, 6.0)
    _, num = _factor_device(n, Ap, Ai, Ax)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    nrhs, ldb = 40, n + 5
    r = np.random.default_rng(11)
    dt = np.complex128 if cplx else np.float64
    B = r.standard_normal((ldb, nrhs)) + (1j * r.standard_normal((ldb, nrhs)) if cplx else 0)
    B = np.asfortranarray(B.astype(dt))
    for trans in ((0, 1, 2) if cplx else (0, 1)):
        X = B.copy(order="F")
        err = C.create_string_buffer(512)
        rc = lib().fplus_solve(num.ptr, C.c_void_p(X.ctypes.data), nrhs, ldb, trans, err, 512)
        assert rc == 0, err.value
        assert np.array_equal(X[n:], B[n:]), "rows past n were written"
        for k in range(nrhs):
            xk = num.solve(np.ascontiguousarray(B[:n, k]), trans)
            assert np.array_equal(X[:n, k], xk), f"trans {trans} column {k}: differs from its own solve"
            assert _bwd(A, X[:n, k], B[:n, k], trans) < 1e-12


def test_high_level_unscaled_complex():
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, True, 0.0)
    sym, num = _factor_device(n, Ap, Ai, Ax)
    assert sym.n_dense_rows >= 2 and sym.n_dense_cols >= 1
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=4)
    assert np.array_equal(num.Dr, ref.Dr) and np.array_equal(num.Dc, ref.Dc) and np.array_equal(num.p, ref.p)
    f = num.factors()
    Ld, Lo = sorted_csc(f["Lp"], f["Li"], f["Lx"]), sorted_csc(ref.lu["Lp"], ref.lu["Li"], ref.lu["Lx"])
    Ud, Uo = sorted_csc(f["Up"], f["Ui"], f["Ux"]), sorted_csc(ref.lu["Up"], ref.lu["Ui"], ref.lu["Ux"])
    assert norm_diff(Ld, Lo) < 1e-12 and norm_diff(Ud, Uo) < 1e-12
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    r = np.random.default_rng(5)
    for trans in (0, 1, 2):
        b = (2 * r.random((n, 3)) - 1) + 1j * (2 * r.random((n, 3)) - 1)
        x = num.solve(b, trans)
        for k in range(3):
            assert _bwd(A, x[:, k], b[:, k], trans) < 1e-12


def test_concurrent_factorizations():
    """the library keeps no per-call state outside its handles (reference include/fplus.h:4-5, separate calls may run
    on separate threads): two host threads factoring and solving at once, three times each, get the oracle's pivots
    and backward errors at round-off"""
    import threading

    cases = [(gen.reference_test_matrix(300, True, 6.0), 4), (gen.convdiff3d(10), 32)]
    results = [None, None]
    errors = []

    def work(i):
        try:
            (n, Ap, Ai, Ax), relax = cases[i]
            A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
            out = []
            for rep in range(3):
                sym, num = _factor_device(n, Ap, Ai, Ax, relax)
                b = np.random.default_rng(20 + rep).standard_normal(n)
                out.append((sym.row_perm.copy(), sym.col_perm.copy(), num.p.copy(), _bwd(A, num.solve(b), b, 0)))
            results[i] = out
        except Exception as e:   # reported below, in the test's own thread
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, ((n, Ap, Ai, Ax), relax) in enumerate(cases):
        rp, cp = results[i][0][0], results[i][0][1]
        ref = O.Numeric(n, Ap, Ai, Ax, rp, cp, relax_rows=relax)
        for rp_k, cp_k, p_k, be in results[i]:
            assert np.array_equal(rp_k, rp) and np.array_equal(cp_k, cp)
            assert np.array_equal(p_k, ref.p)
            assert be < 1e-12
