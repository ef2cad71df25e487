"""The path at the benchmark's own configurations on the device.

- the CPU-runnable configuration (2D convection-diffusion on a 96 x 96 grid, n = 9216): full parity with the oracle
  (pivots, supernodes, patterns bit-exact, values normwise to 1e-12, backward error < 1e-12 for both solutions;
  the matrix's scaling over 12 orders makes the solutions themselves agree only to its conditioning);
- the plain FP64 drop-in configuration (complex shift-invert pencil, 64 modes x 4 fields x 1024 coefficients,
  n = 262144) and the 3D convection-diffusion configuration (64^3, n = 262144): size-independent properties -- the
  pivot order is a permutation, the factors reproduce every right-hand side to backward error < 1e-12 (with A, A^T and
  A^H), and several right-hand sides at once equal one at a time bit for bit;
- the splu drop-in surface (SparseLU) against the oracle.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import gen
from helpers import norm_diff, rel_diff, same_pattern, sorted_csc

pytestmark = pytest.mark.gpu


def _bwd(A, x, b, op):
    nA = abs(A).sum(axis=0).max()
    M = {"N": A, "T": A.T, "H": A.conj().T}[op]
    return float(np.linalg.norm(M @ x - b) / (nA * np.linalg.norm(x) + np.linalg.norm(b)))


def test_convdiff2d_96_parity():
    n, Ap, Ai, Ax = gen.convdiff2d(96)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=32))
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=32)
    assert np.array_equal(num.Dr, ref.Dr) and np.array_equal(num.Dc, ref.Dc)
    assert np.array_equal(num.p, ref.p)
    f = num.factors()
    assert f["nsupernodes"] == ref.lu["nsn"]
    Ld, Lo = sorted_csc(f["Lp"], f["Li"], f["Lx"]), sorted_csc(ref.lu["Lp"], ref.lu["Li"], ref.lu["Lx"])
    Ud, Uo = sorted_csc(f["Up"], f["Ui"], f["Ux"]), sorted_csc(ref.lu["Up"], ref.lu["Ui"], ref.lu["Ux"])
    assert same_pattern(Ld, Lo) and same_pattern(Ud, Uo)
    assert norm_diff(Ld, Lo) < 1e-12 and norm_diff(Ud, Uo) < 1e-12
    assert rel_diff(Ld, Lo) < 1e-10 and rel_diff(Ud, Uo) < 1e-10
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    r = np.random.default_rng(0)
    xt = r.standard_normal(n)
    b = A @ xt                                  # one right-hand side b = A x_true, x_true ~ N(0, 1), seed 0
    x = num.solve(b)
    xo = ref.solve(b)
    # rows and columns scaled over 12 orders: the solutions agree to the problem's conditioning, the backward errors
    # of both are at round-off
    assert _bwd(A, x, b, "N") < 1e-12 and _bwd(A, xo, b, "N") < 1e-12


@pytest.mark.parametrize("name", ["shiftinvert-64x4x1024", "convdiff3d-64"])
def test_full_size_properties(name):
    if name.startswith("shift"):
        n, Ap, Ai, Ax = gen.shift_invert(64, 4, 1024)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(64)
    cplx = np.iscomplexobj(Ax)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    t0 = time.time()

    def step(what):
        print(f"[{name}] {what} at {time.time() - t0:.1f} s", flush=True)

    sym = fplus_amd.Symbolic(n, Ap, Ai)
    step("analysed")
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    step("factored")
    assert np.array_equal(np.sort(num.p), np.arange(n)) and np.array_equal(np.sort(num.q), np.arange(n))
    assert np.all(np.isfinite(num.Dr)) and np.all(num.Dr > 0) and np.all(np.isfinite(num.Dc)) and np.all(num.Dc > 0)
    B = gen.rhs(n, 16, cplx)
    X = num.solve(B)
    step("solved 16 right-hand sides")
    for k in range(16):
        assert _bwd(A, X[:, k], B[:, k], "N") < 1e-12, f"right-hand side {k}"
    x3 = num.solve(B[:, 3])
    assert np.array_equal(x3.view(np.float64), np.ascontiguousarray(X[:, 3]).view(np.float64))   # 1 rhs = 16 at once
    for op, tr in (("T", 1), ("H", 2)):
        x = num.solve(B[:, 0], tr)
        step(f"solved with A^{op}")
        assert _bwd(A, x, B[:, 0], op) < 1e-12, op


def test_sparselu_drop_in():
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 48, seed=9)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    lu = fplus_amd.SparseLU(A, relax_rows=32)
    assert lu.shape == (n, n) and lu.nnz == lu.L.nnz + lu.U.nnz
    assert sorted(lu.perm_r) == list(range(n)) and sorted(lu.perm_c) == list(range(n))
    b = gen.rhs(n, 2, True)
    for op in ("N", "T", "H"):
        x = lu.solve(b, trans=op)
        assert x.shape == (n, 2)
        for k in range(2):
            assert _bwd(A, x[:, k], b[:, k], op) < 1e-12, op
        x0 = lu.solve(b[:, 0], trans=op)
        assert x0.shape == (n,) and np.array_equal(x0, x[:, 0])
