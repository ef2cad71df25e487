"""The solves on caller-held CSC factors (fplus_tri_solve, fplus_solve_plan_*, reference src/solve.c:11-74 and
src/solve_plan.c:81-719) against the oracle's plain sweeps: bit-identical for every trans, real and complex, with and
without a plan, for one and several right-hand sides, on the supernodal layout, the generic one (L's diagonal not
first), and factors with entries the sweeps must skip or apply late (L above its diagonal, U below it, repeated
diagonals).  The plan's supernode count equals the oracle's restatement of the reference's."""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import gen
from helpers import parity_matrices

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.ascontiguousarray(x).view(np.float64)


def _oracle_factors(n, Ap, Ai, Ax, relax):
    xs, _, _, _ = O.ruiz(n, Ap, Ai, Ax)
    f = O.factor_sn(n, Ap, Ai, xs, relax_rows=relax)
    return (f["Lp"], f["Li"], f["Lx"]), (f["Up"], f["Ui"], f["Ux"])


def _rhs(n, k, cplx, seed):
    r = np.random.default_rng(seed)
    b = r.standard_normal((n, k))
    return b + 1j * r.standard_normal((n, k)) if cplx else b


@pytest.mark.parametrize("case", parity_matrices(), ids=lambda c: c[0])
def test_tri_solve_bitexact(case):
    name, n, Ap, Ai, Ax, relax = case
    L, U = _oracle_factors(n, Ap, Ai, Ax, relax)
    cplx = np.iscomplexobj(Ax)
    b = _rhs(n, 1, cplx, 1)[:, 0]
    for trans in (0, 1, 2):
        x = fplus_amd.tri_solve(L, U, b, trans)
        xo = O.tri_solve(L, U, b, trans)
        assert np.array_equal(_bits(x), _bits(xo)), f"{name} trans {trans}: differs from the plain sweeps"


@pytest.mark.parametrize("case", parity_matrices(), ids=lambda c: c[0])
def test_plan_several_rhs(case):
    """a plan, 21 right-hand sides (two passes), equal to one-at-a-time plan-free solves and to the oracle"""
    name, n, Ap, Ai, Ax, relax = case
    L, U = _oracle_factors(n, Ap, Ai, Ax, relax)
    plan = fplus_amd.SolvePlan(L[0], L[1], U[0], U[1])
    assert plan.nsupernodes == O.plan_nsupernodes(L[0], L[1], U[0], U[1])
    cplx = np.iscomplexobj(Ax)
    B = _rhs(n, 21, cplx, 2)
    for trans in (0, 1, 2):
        X = fplus_amd.tri_solve(L, U, B, trans, plan)
        Xo = O.tri_solve(L, U, B, trans)
        assert np.array_equal(_bits(X), _bits(Xo)), f"{name} trans {trans}: plan solve differs from the plain sweeps"
        x5 = fplus_amd.tri_solve(L, U, B[:, 5], trans)
        assert np.array_equal(_bits(x5), _bits(X[:, 5]))


def test_device_factors_through_plan():
    """the factors of the device factorization, exported, solve through a plan as through the oracle's sweeps"""
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 48, seed=3)
    lu = fplus_amd.SparseLU(sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)))
    f = lu._factors()
    L, U = (f["Lp"], f["Li"], f["Lx"]), (f["Up"], f["Ui"], f["Ux"])
    plan = fplus_amd.SolvePlan(L[0], L[1], U[0], U[1])
    B = _rhs(n, 3, True, 4)
    for trans in (0, 1, 2):
        assert np.array_equal(_bits(fplus_amd.tri_solve(L, U, B, trans, plan)), _bits(O.tri_solve(L, U, B, trans)))


def test_generic_layout_falls_back():
    """L with every column reversed (the diagonal last): no supernodes (-1); the plain sweeps' solutions on the reversed
    factor bit for bit, which for trans 0 are the solutions on the original factor (reference test_solve_plan.py:89-102);
    the transposed sweeps sum L's columns in storage order, so there the reversal changes the rounding"""
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, True, 6.0)
    L, U = _oracle_factors(n, Ap, Ai, Ax, 4)
    Lp, Li, Lx = L
    Li2, Lx2 = Li.copy(), Lx.copy()
    for j in range(n):
        s = slice(Lp[j], Lp[j + 1])
        Li2[s] = Li[s][::-1]
        Lx2[s] = Lx[s][::-1]
    plan = fplus_amd.SolvePlan(Lp, Li2, U[0], U[1])
    assert plan.nsupernodes == -1 == O.plan_nsupernodes(Lp, Li2, U[0], U[1])
    b = _rhs(n, 2, True, 5)
    for trans in (0, 1, 2):
        x = fplus_amd.tri_solve((Lp, Li2, Lx2), U, b, trans, plan)
        assert np.array_equal(_bits(x), _bits(O.tri_solve((Lp, Li2, Lx2), U, b, trans)))
        if trans == 0:   # L y = b updates each row once per column: the order inside a column does not matter
            assert np.array_equal(_bits(x), _bits(O.tri_solve(L, U, b, trans)))


def _odd_factors(n, cplx, seed):
    """random CSC 'factors' with every irregularity the plain sweeps define: L entries above the diagonal, a missing
    or repeated unit diagonal, duplicate rows; U entries below the diagonal, repeated diagonals, unsorted rows"""
    r = np.random.default_rng(seed)

    def val(m):
        v = r.uniform(-1, 1, m)
        return v + 1j * r.uniform(-1, 1, m) if cplx else v

    Lp, Li, Up, Ui = [0], [], [0], []
    for j in range(n):
        rows = [j] if r.random() < 0.9 else []
        rows += list(r.choice(n, size=r.integers(0, 5), replace=True))
        r.shuffle(rows)
        Li += rows
        Lp.append(len(Li))
        urows = list(r.choice(n, size=r.integers(0, 5), replace=True)) + [j]
        if r.random() < 0.2:
            urows.insert(int(r.integers(0, len(urows))), j)
        r.shuffle(urows)
        Ui += urows
        Up.append(len(Ui))
    Lx, Ux = val(len(Li)) * 0.3, val(len(Ui)) * 0.3
    Ux[np.array(Ui) == np.repeat(np.arange(n), np.diff(Up))] += 2.0
    return ((np.array(Lp), np.array(Li), Lx), (np.array(Up), np.array(Ui), Ux))


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_irregular_factors(cplx, seed):
    n = 400
    L, U = _odd_factors(n, cplx, seed)
    plan = fplus_amd.SolvePlan(L[0], L[1], U[0], U[1])
    assert plan.nsupernodes == O.plan_nsupernodes(L[0], L[1], U[0], U[1])
    b = _rhs(n, 3, cplx, 10 + seed)
    b[7] = 0.0   # zero entries: the trans-0 sweeps skip them
    for trans in (0, 1, 2):
        x = fplus_amd.tri_solve(L, U, b, trans, plan)
        assert np.array_equal(_bits(x), _bits(O.tri_solve(L, U, b, trans))), f"trans {trans}"


def test_singular_and_mismatch():
    n, Ap, Ai, Ax = gen.convdiff2d(12)
    L, U = _oracle_factors(n, Ap, Ai, Ax, 32)
    b = np.ones(n)
    Ux = U[2].copy()
    j = 17
    Ux[U[0][j + 1] - 1] = 0.0        # U(j, j), stored last
    for trans in (0, 1):
        with pytest.raises(fplus_amd.FplusError) as e:
            fplus_amd.tri_solve(L, (U[0], U[1], Ux), b, trans)
        assert e.value.code == fplus_amd._lib.FPLUS_ERR_SINGULAR
    plan = fplus_amd.SolvePlan(L[0], L[1], U[0], U[1])
    n2, Ap2, Ai2, Ax2 = gen.convdiff2d(13)
    L2, U2 = _oracle_factors(n2, Ap2, Ai2, Ax2, 32)
    with pytest.raises(ValueError, match="plan does not match"):
        fplus_amd.tri_solve(L2, U2, np.ones(n2), 0, plan)
    # the C entry point checks the sizes itself
    import ctypes as C
    Lp, Li, Lx = (np.ascontiguousarray(a) for a in L)
    Up2 = np.ascontiguousarray(U[0]).copy()
    Up2[-1] -= 1
    x = np.ones(n)
    rc = fplus_amd.lib().fplus_tri_solve_plan(plan.ptr, n, fplus_amd._p64(Lp), fplus_amd._p64(Li), C.c_void_p(Lx.ctypes.data),
                                              fplus_amd._p64(Up2), fplus_amd._p64(U[1]), C.c_void_p(U[2].ctypes.data),
                                              C.c_void_p(x.ctypes.data), 1, n, 0, 0)
    assert rc == fplus_amd._lib.FPLUS_ERR_BADARG


def test_larger_factor_bitexact():
    """a factor with long dependency chains (convection-diffusion, n = 9216) and wide supernodes"""
    n, Ap, Ai, Ax = gen.convdiff2d(96)
    L, U = _oracle_factors(n, Ap, Ai, Ax, 32)
    plan = fplus_amd.SolvePlan(L[0], L[1], U[0], U[1])
    B = _rhs(n, 16, False, 6)
    for trans in (0, 1):
        X = fplus_amd.tri_solve(L, U, B, trans, plan)
        assert np.array_equal(_bits(X), _bits(O.tri_solve(L, U, B, trans)))
