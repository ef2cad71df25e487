"""The solves with A as two persistent launches (tuning knob "solve_dataflow") against the level-by-level sweeps (the
default) and the oracle: items taken by an atomic ticket, each waiting for its inputs through per-block counters.

- the forward sweep's results are those of the level sweeps (same operations, same order); the backward sweep groups a
  block's dot products differently (at most 8 groups, their partial sums added in group order), so the solutions differ
  from the level sweeps' by round-off times the conditioning: 1e-14 on the small cases, 3e-12 on the n = 8192
  shift-invert member (whose solution differs from the oracle's by 5e-11 either way); the bound is 1e-10, every column;
- a column solved alone equals the same column solved among 21 (two passes) bit for bit, and two solves give the same
  bits (no atomics in any sum);
- the solutions meet the equilibrated backward error 1e-14 and agree with the oracle's to 1e-10.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import accuracy as acc
from fplus_amd import gen
from helpers import parity_matrices

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore_knob():
    yield
    fplus_amd.tune("solve_dataflow", 0)


def _bits(x):
    return np.ascontiguousarray(x).view(np.float64)


def _cases():
    out = [c for c in parity_matrices() if c[0] in ("reftest-z", "convdiff3d-10", "shiftinvert-4x4x32")]
    out.append(("shiftinvert-16x4x128", *gen.shift_invert(16, 4, 128), 32))
    return out


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c[0])
def test_dataflow_against_level_sweeps(case):
    name, n, Ap, Ai, Ax, relax = case
    cplx = np.iscomplexobj(Ax)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=relax))
    rng = np.random.default_rng(5)
    b = rng.standard_normal((n, 21)) + (1j * rng.standard_normal((n, 21)) if cplx else 0)
    fplus_amd.tune("solve_dataflow", 0)
    xl = num.solve(b)
    fplus_amd.tune("solve_dataflow", 1)
    xd = num.solve(b)
    xd2 = num.solve(b)
    x1 = num.solve(b[:, 3])
    assert np.array_equal(_bits(xd), _bits(xd2)), f"{name}: two persistent solves differ"
    assert np.array_equal(_bits(x1), _bits(xd[:, 3])), f"{name}: one column alone differs from the same column among 21"
    d = np.max(np.linalg.norm(xd - xl, axis=0) / np.linalg.norm(xl, axis=0))
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    be = acc.equilibrated(A, xd, b, num.Dr, num.Dc)
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=relax) if n <= 20000 else None
    do = 0.0
    if ref is not None:
        xo = ref.solve(b[:, :2])
        do = np.max(np.linalg.norm(xd[:, :2] - xo, axis=0) / np.linalg.norm(xo, axis=0))
    print(f"{name}: n {n}, persistent vs level sweeps {d:.2e}, vs the oracle {do:.2e}, equilibrated backward error {be:.2e}")
    assert d < 1e-10 and be < 1e-14 and do < 1e-10, name


def test_dataflow_transposed_untouched():
    """trans 1 and 2 keep the level sweeps: the same bits whatever the knob"""
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 32)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    b = gen.rhs(n, 3, True)
    for trans in (1, 2):
        fplus_amd.tune("solve_dataflow", 0)
        xl = num.solve(b, trans)
        fplus_amd.tune("solve_dataflow", 1)
        assert np.array_equal(_bits(num.solve(b, trans)), _bits(xl)), trans
