"""The oracle's restatement of the solve plan's supernode count (reference src/solve_plan.c:97-199) and its plain
sweeps, pinned by closed forms and, where oracle/_ref is built, by the reference's own plan on the same patterns."""
import numpy as np
import pytest

import oracle as O
from helpers import parity_matrices


def _dense(n):
    Lp = np.arange(n + 1) * 0
    Li, Up, Ui = [], [0], []
    lp = [0]
    for j in range(n):
        Li += list(range(j, n))
        lp.append(len(Li))
        Ui += list(range(0, j + 1))
        Up.append(len(Ui))
    del Lp
    return np.array(lp), np.array(Li), np.array(Up), np.array(Ui)


def _identity(n):
    a = np.arange(n + 1)
    return a, np.arange(n), a.copy(), np.arange(n)


def _cases():
    out = [("dense-7", *_dense(7)), ("identity-9", *_identity(9))]
    Lp, Li, Up, Ui = _identity(3)
    # a row repeated above the diagonal of U column 2 and an entry below U's diagonal (blocked layout: generic)
    out.append(("repeated-row", Lp, Li, np.array([0, 2, 3, 6]), np.array([0, 1, 1, 0, 0, 2])))
    # the same repeat without the odd entry: plain sweeps, no overlap test
    out.append(("repeated-row-plain", Lp, Li, np.array([0, 1, 2, 5]), np.array([0, 1, 0, 0, 2])))
    Lp, Li, Up, Ui = _dense(5)
    out.append(("diag-last", Lp, np.concatenate([Li[Lp[j]:Lp[j + 1]][::-1] for j in range(5)]), Up, Ui))
    for name, n, Ap, Ai, Ax, relax in parity_matrices():
        xs, _, _, _ = O.ruiz(n, Ap, Ai, Ax)
        f = O.factor_sn(n, Ap, Ai, xs, relax_rows=relax)
        out.append((name, f["Lp"], f["Li"], f["Up"], f["Ui"]))
    return out


def test_plan_nsupernodes_closed_forms():
    assert O.plan_nsupernodes(*_dense(7)) == 1
    assert O.plan_nsupernodes(*_identity(9)) == 9
    c = {k[0]: k[1:] for k in _cases()}
    assert O.plan_nsupernodes(*c["repeated-row"]) == -1
    assert O.plan_nsupernodes(*c["repeated-row-plain"]) == 3
    assert O.plan_nsupernodes(*c["diag-last"]) == -1


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_plan_nsupernodes_matches_reference_build(case):
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    name, Lp, Li, Up, Ui = case
    assert O.plan_nsupernodes(Lp, Li, Up, Ui) == O.ref_plan_nsupernodes(Lp, Li, Up, Ui), name


def test_plain_sweeps_solve_the_factors():
    """x = U^-1 L^-1 b solves L U x = b (closed form: the products with the factors)"""
    import scipy.sparse as sp
    name, n, Ap, Ai, Ax, relax = parity_matrices()[0]
    xs, _, _, _ = O.ruiz(n, Ap, Ai, Ax)
    f = O.factor_sn(n, Ap, Ai, xs, relax_rows=relax)
    L = sp.csc_matrix((f["Lx"], f["Li"], f["Lp"]), shape=(n, n))
    U = sp.csc_matrix((f["Ux"], f["Ui"], f["Up"]), shape=(n, n))
    b = np.random.default_rng(0).standard_normal(n) + 0j
    for trans, op in ((0, lambda v: L @ (U @ v)), (1, lambda v: U.T @ (L.T @ v)), (2, lambda v: U.conj().T @ (L.conj().T @ v))):
        x = O.tri_solve((f["Lp"], f["Li"], f["Lx"]), (f["Up"], f["Ui"], f["Ux"]), b, trans)
        assert np.linalg.norm(op(x) - b) <= 1e-9 * np.linalg.norm(b), trans
