"""The backward-error measures (fplus_amd.accuracy) on the oracle's solutions, with mistakes planted in them on the CPU.

On a matrix whose rows and columns are scaled over 16 orders (the complex shift-invert family), the reference tests'
normwise measure scores a solution with a 1e-2 relative error in every component far below 1e-13; the componentwise and
the equilibrated measures see it.
"""
import numpy as np
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import accuracy as acc
from fplus_amd import gen


def _case():
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 64)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=32)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 2, True)
    return A, ref, b


def test_measures_of_a_good_solution():
    A, ref, b = _case()
    for trans in (0, 1, 2):
        x = ref.solve(b, trans)
        m = acc.all_measures(A, x, b, ref.Dr, ref.Dc, trans)
        assert m["normwise"] < 1e-20 and m["equilibrated"] < 1e-15 and m["componentwise"] < 1e-11, (trans, m)


def test_planted_errors_are_seen():
    A, ref, b = _case()
    x = ref.solve(b, 0)
    rng = np.random.default_rng(5)
    for rel in (1e-2, 1e-8):
        xp = x * (1 + rel * rng.standard_normal(x.shape))      # every component off by about rel
        m = acc.all_measures(A, xp, b, ref.Dr, ref.Dc, 0)
        assert m["normwise"] < 1e-13                            # the reference's measure passes it
        assert m["componentwise"] > rel / 100 and m["equilibrated"] > rel / 100, (rel, m)


def test_measures_match_closed_forms():
    """dense checks of the three formulas on a tiny matrix"""
    rng = np.random.default_rng(1)
    n = 6
    M = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = sp.csc_matrix(M)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    b = M @ x + 1e-3 * rng.standard_normal(n)
    Dr = 10.0 ** rng.uniform(-3, 3, n)
    Dc = 10.0 ** rng.uniform(-3, 3, n)
    r = b - M @ x
    assert np.isclose(acc.normwise(A, x, b), np.linalg.norm(r) / (np.abs(M).sum(0).max() * np.linalg.norm(x) + np.linalg.norm(b)))
    assert np.isclose(acc.componentwise(A, x, b), np.max(np.abs(r) / (np.abs(M) @ np.abs(x) + np.abs(b))))
    Mb = Dr[:, None] * M * Dc[None, :]
    e = np.linalg.norm(Dr * r) / (np.abs(Mb).sum(0).max() * np.linalg.norm(x / Dc) + np.linalg.norm(Dr * b))
    assert np.isclose(acc.equilibrated(A, x, b, Dr, Dc), e)
    # A^H x = b: the scalings swap sides
    rh = b - M.conj().T @ x
    MbH = Dc[:, None] * M.conj().T * Dr[None, :]
    eh = np.linalg.norm(Dc * rh) / (np.abs(MbH).sum(0).max() * np.linalg.norm(x / Dr) + np.linalg.norm(Dc * b))
    assert np.isclose(acc.equilibrated(A, x, b, Dr, Dc, 2), eh)
