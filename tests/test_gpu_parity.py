"""Parity of the HIP path (through the C interface) with the oracle, on the same seeded inputs.

Bit-exact: Ruiz (Dr, Dc and the scaled values), the pivot order, the supernode partition, the L and U patterns.
To 1e-12 relative (normwise over each factor) the L and U values, and to 1e-10 column by column (the summation order
differs, and a column built by cancellation carries the round-off of its larger terms); the solutions to 1e-10 of the
oracle's; backward errors below 1e-12.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from helpers import norm_diff, parity_matrices, rel_diff, same_pattern, sorted_csc

pytestmark = pytest.mark.gpu

TOL_VALUES = 1e-12       # normwise, per factor
TOL_COLUMNS = 1e-10      # per column


@pytest.mark.parametrize("case", parity_matrices(), ids=lambda c: c[0])
def test_ruiz_bitexact(case):
    name, n, Ap, Ai, Ax, _ = case
    for is_csc in (True, False):
        if is_csc:
            ptr, idx, x = Ap, Ai, Ax
        else:
            A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)).tocsr()
            A.sort_indices()
            ptr, idx, x = A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
        xd, Drd, Dcd = fplus_amd.ruiz(n, ptr, idx, x, is_csc=is_csc)
        xo, Dro, Dco, _ = O.ruiz(n, ptr, idx, x, is_csc=is_csc)
        assert np.array_equal(Drd, Dro) and np.array_equal(Dcd, Dco), f"{name} csc={is_csc}: Dr / Dc differ"
        assert np.array_equal(xd.view(np.float64), xo.view(np.float64)), f"{name} csc={is_csc}: scaled values differ"


def _compare(name, n, Ap, Ai, Ax, relax, sym):
    opts = fplus_amd.default_options(relax_rows=relax)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, opts)
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=relax)
    assert np.array_equal(num.Dr, ref.Dr) and np.array_equal(num.Dc, ref.Dc), f"{name}: Dr / Dc differ"
    assert np.array_equal(num.p, ref.p), f"{name}: pivot order differs"
    f = num.factors()
    assert f["nsupernodes"] == ref.lu["nsn"], f"{name}: supernodes {f['nsupernodes']} vs {ref.lu['nsn']}"
    Ld, Lo = sorted_csc(f["Lp"], f["Li"], f["Lx"]), sorted_csc(ref.lu["Lp"], ref.lu["Li"], ref.lu["Lx"])
    Ud, Uo = sorted_csc(f["Up"], f["Ui"], f["Ux"]), sorted_csc(ref.lu["Up"], ref.lu["Ui"], ref.lu["Ux"])
    assert same_pattern(Ld, Lo), f"{name}: L pattern differs"
    assert same_pattern(Ud, Uo), f"{name}: U pattern differs"
    nl, nu = norm_diff(Ld, Lo), norm_diff(Ud, Uo)
    dl, du = rel_diff(Ld, Lo), rel_diff(Ud, Uo)
    msg = f"{name}: L / U differ normwise by {nl:.2e} / {nu:.2e}, per column by {dl:.2e} / {du:.2e}"
    print(msg)
    assert nl < TOL_VALUES and nu < TOL_VALUES, msg
    assert dl < TOL_COLUMNS and du < TOL_COLUMNS, msg
    cplx = np.iscomplexobj(Ax)
    rng = np.random.default_rng(7)
    b = rng.standard_normal((n, 3)) + (1j * rng.standard_normal((n, 3)) if cplx else 0)
    for trans in ((0, 1, 2) if cplx else (0, 1)):
        x = num.solve(b, trans)
        xo = ref.solve(b, trans)
        d = np.linalg.norm(x - xo) / np.linalg.norm(xo)
        assert d < 1e-10, f"{name} trans {trans}: solution differs from the oracle's by {d:.2e}"
        for r in range(3):
            be = O.backward_error(n, Ap, Ai, Ax, x[:, r], b[:, r], trans)
            assert be < 1e-12, f"{name} trans {trans}: backward error {be:.2e}"


@pytest.mark.parametrize("case", parity_matrices(), ids=lambda c: c[0])
def test_factor_parity(case):
    name, n, Ap, Ai, Ax, relax = case
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    _compare(name, n, Ap, Ai, Ax, relax, sym)


@pytest.mark.parametrize("fixture", ["relaxed_panel_case3", "relaxed_panel_case50"])
@pytest.mark.parametrize("relax", [2, 8])
def test_relaxed_panel_fixtures(fixture, relax):
    """the reference's regression matrices (stored pre-ordered): identity orders, relaxed joins across panels"""
    import os
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", fixture + ".npz"))
    A = sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=(1200, 1200)).tocsc()
    A.sort_indices()
    n, Ap, Ai, Ax = 1200, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data
    sym = fplus_amd.symbolic_from_orders(n, Ap, Ai, np.arange(n), np.arange(n))
    _compare(fixture, n, Ap, Ai, Ax, relax, sym)
