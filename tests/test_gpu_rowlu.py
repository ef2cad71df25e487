"""The row kernel on the device (rook / full / partial pivoting, incomplete LU; reference src/factor_row.c:1191-1685,
src/pipeline.c:129-225, the fplus_factor dispatch of src/numeric.c:51-64) against the oracle, whose row kernel is
bit-identical to the reference's own (tests/test_oracle_row.py).

- fplus_factor_row on the equilibrated CSR cases, 3 strategies x drop_tol 0 / 1e-3 / 1e-2: the same p, q and factor
  patterns; real values bit-identical (every value is formed with the reference's operations in its order), complex
  values within 1e-12 (the device divides complex numbers by Smith's method, the C library by its own);
- a matrix whose factorization fills in completely, so that the kernel outgrows its entry pool and multiplier log and
  is run again with larger ones;
- n = 0, exactly singular and structurally singular matrices (the oracle's failing step and message);
- the pipelines (fplus_factor_row_sym, and fplus_factor with rook / full / drop_tol > 0) against the oracle's permute,
  Ruiz and row kernel composed as src/pipeline.c and src/numeric.c do; the solves they give;
- the reference's C tests: the ILU breakdown case (test_fplus.c:204-229) and the ILU-preconditioned Richardson
  iteration of every strategy (test_fplus.c:166-199);
- the device-factorization extensions refuse a row-kernel handle.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle
from fplus_amd import accuracy as acc
from fplus_amd import gen
from fplus_amd._lib import FPLUS_ERR_BADARG, FPLUS_ERR_SINGULAR, FplusError, lib
from helpers import csr_arrays, row_kernel_cases

pytestmark = pytest.mark.gpu

CASES = row_kernel_cases()
NAMES = {0: "full", 1: "partial", 2: "rook"}


def assert_same_row_factors(d, o, cplx, what=""):
    for key in ("p", "q", "Lp", "Li", "Up", "Ui"):
        assert np.array_equal(d[key], o[key]), (what, key)
    assert d["n_pivots_replaced"] == o["n_replaced"], what
    for key in ("Lx", "Ux"):
        if cplx:
            scale = max(1.0, float(np.abs(o[key]).max(initial=0.0)))
            assert np.abs(d[key] - o[key]).max(initial=0.0) <= 1e-12 * scale, (what, key)
        else:
            assert np.array_equal(d[key], o[key]), (what, key)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("strategy", [0, 1, 2], ids=["full", "partial", "rook"])
@pytest.mark.parametrize("drop_tol", [0.0, 1e-3, 1e-2])
def test_row_kernel_against_the_oracle(case, strategy, drop_tol):
    name, n, Rp, Ri, Rx = case
    print(f"row kernel {name} {NAMES[strategy]} drop_tol {drop_tol}", flush=True)
    try:
        o = oracle.factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
    except oracle.OracleError as e:
        with pytest.raises(FplusError) as r:
            fplus_amd.factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
        assert r.value.code == e.code
        return
    d = fplus_amd.factor_row(n, Rp, Ri, Rx, drop_tol, strategy)
    assert d["csc"] == 0
    assert_same_row_factors(d, o, np.iscomplexobj(Rx), f"{name} {strategy} {drop_tol}")


def test_pools_grow_when_the_fill_outgrows_them():
    """an arrow matrix pivoting on its dense row first: every later row fills in completely (n^2 / 2 entries of L and of
    U against 3n of A), so the kernel reports its entry pool and multiplier log full and runs again with larger ones"""
    n = 400
    rng = np.random.default_rng(3)
    A = sp.lil_matrix((n, n))
    A[0, :] = rng.uniform(0.5, 1.0, n)
    A[:, 0] = rng.uniform(0.5, 1.0, (n, 1))
    A[0, 0] = 10.0
    A.setdiag(np.concatenate([[10.0], rng.uniform(1.0, 2.0, n - 1)]))
    n, Rp, Ri, Rx = csr_arrays(A.tocsr())
    for strategy in (1, 2):
        print(f"fill growth {NAMES[strategy]}", flush=True)
        o = oracle.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
        assert o["Up"][-1] > 16 * len(Ri)
        d = fplus_amd.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
        assert_same_row_factors(d, o, False, "fill growth")


def test_empty_and_singular():
    print("row kernel: n = 0, singular matrices", flush=True)
    d = fplus_amd.factor_row(0, np.zeros(1, np.int64), np.zeros(0, np.int64), np.zeros(0), 0.0, 0)
    assert d["Lp"].tolist() == [0] and d["Up"].tolist() == [0]
    n, Rp, Ri, Rx = csr_arrays(np.ones((2, 2)))
    for strategy in (0, 1, 2):
        with pytest.raises(FplusError) as e:
            fplus_amd.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
        assert e.value.code == FPLUS_ERR_SINGULAR and "Matrix is singular" in str(e.value)
    # a zero column: the oracle's failing step and message
    A = sp.random(60, 60, density=0.1, random_state=np.random.default_rng(2), format="lil") + sp.eye(60)
    A[:, 17] = 0
    n, Rp, Ri, Rx = csr_arrays(sp.csr_matrix(A))
    for strategy in (1, 2):
        with pytest.raises(oracle.OracleError) as eo:
            oracle.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
        with pytest.raises(FplusError) as ed:
            fplus_amd.factor_row(n, Rp, Ri, Rx, 0.0, strategy)
        assert ed.value.code == FPLUS_ERR_SINGULAR
        assert str(eo.value).split(": ", 1)[1] in str(ed.value)


def oracle_pipeline(n, Ap, Ai, Ax, rp, cp, drop_tol, strategy, equilibrate=True):
    """the oracle's permute, Ruiz (CSC route), CSR layout and row kernel, the factors back in CSC (src/pipeline.c:171-209)"""
    Bp, Bi, Bx = oracle.permute(n, Ap, Ai, Ax, rp, cp)
    if equilibrate:
        Bx, Dr, Dc, _ = oracle.ruiz(n, Bp, Bi, Bx)
    else:
        Dr, Dc = np.ones(n), np.ones(n)
    R = sp.csc_matrix((Bx, Bi, Bp), shape=(n, n)).tocsr()
    R.sort_indices()
    o = oracle.factor_row(n, R.indptr.astype(np.int64), R.indices.astype(np.int64), R.data, drop_tol, strategy)
    L = sp.csr_matrix((o["Lx"], o["Li"], o["Lp"]), shape=(n, n)).tocsc()
    U = sp.csr_matrix((o["Ux"], o["Ui"], o["Up"]), shape=(n, n)).tocsc()
    L.sort_indices()
    U.sort_indices()
    return o, L, U, Dr, Dc


def pipeline_cases():
    out = []
    for cplx in (False, True):
        out.append((f"reftest200-{'z' if cplx else 'd'}", *gen.reference_test_matrix(200, cplx, 6.0)))
    out.append(("convdiff2d-14", *gen.convdiff2d(14)))
    return out


PCASES = pipeline_cases()


@pytest.mark.parametrize("case", PCASES, ids=[c[0] for c in PCASES])
@pytest.mark.parametrize("strategy,drop_tol", [(0, 0.0), (2, 0.0), (1, 1e-2), (2, 1e-3)])
def test_row_pipelines_against_the_oracle(case, strategy, drop_tol):
    name, n, Ap, Ai, Ax = case
    print(f"row pipeline {name} {NAMES[strategy]} drop_tol {drop_tol}", flush=True)
    cplx = np.iscomplexobj(Ax)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    rp, cp = sym.row_perm, sym.col_perm
    o, Lo, Uo, Dro, Dco = oracle_pipeline(n, Ap, Ai, Ax, rp, cp, drop_tol, strategy)
    # the low-level pipeline: CSC factors of the permuted matrix, its Dr, Dc
    d = fplus_amd.factor_row_sym(sym, Ap, Ai, Ax, drop_tol, strategy)
    assert d["csc"] == 1
    assert np.array_equal(d["Dr"], Dro) and np.array_equal(d["Dc"], Dco)
    assert np.array_equal(d["p"], o["p"]) and np.array_equal(d["q"], o["q"])
    for key, M in (("L", Lo), ("U", Uo)):
        assert np.array_equal(d[key + "p"], M.indptr) and np.array_equal(d[key + "i"], M.indices), key
        if cplx:
            assert np.abs(d[key + "x"] - M.data).max() <= 1e-12 * max(1.0, np.abs(M.data).max()), key
        else:
            assert np.array_equal(d[key + "x"], M.data), key
    # fplus_factor: the orders and scalings composed (src/numeric.c:69-78), the solve through the plan
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(pivot_strategy=strategy, drop_tol=drop_tol))
    assert np.array_equal(num.p, rp[o["p"]]) and np.array_equal(num.q, cp[o["q"]])
    Dr = np.empty(n); Dc = np.empty(n)
    Dr[rp] = Dro
    Dc[cp] = Dco
    assert np.array_equal(num.Dr, Dr) and np.array_equal(num.Dc, Dc)
    f = num.factors()
    assert np.array_equal(f["Lp"], Lo.indptr) and np.array_equal(f["Ui"], Uo.indices)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    rng = np.random.default_rng(7)
    b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
    x = num.solve(b)
    if drop_tol == 0.0:
        m = acc.all_measures(A, x, b, num.Dr, num.Dc)
        assert m["equilibrated"] < 1e-14, m
    for trans in (1, 2):
        xt = num.solve(b, trans)
        if drop_tol == 0.0:
            assert acc.equilibrated(A, xt, b, num.Dr, num.Dc, trans) < 1e-14


def test_ilu_breakdown_case_of_the_reference():
    """[[1, 1], [1, 1.0005]], full pivoting, drop_tol 1e-2, no equilibration, no pinning: one pivot replaced by
    0.010005 (test_fplus.c:204-222); without dropping the exactly singular [[1, 1], [1, 1]] fails"""
    print("ILU breakdown case", flush=True)
    Ap = np.array([0, 2, 4], np.int64)
    Ai = np.array([0, 1, 0, 1], np.int64)
    Ax = np.array([1.0, 1.0, 1.0, 1.0005])
    opt = dict(pin_dense=0, equilibrate=0)
    sym = fplus_amd.Symbolic(2, Ap, Ai, pin_dense=False)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(pivot_strategy=0, drop_tol=1e-2, **opt))
    f = num.factors()
    assert f["n_pivots_replaced"] == 1
    U = f["U"].toarray()
    assert abs(U[1, 1] - 0.010005) < 1e-15
    with pytest.raises(FplusError) as e:
        fplus_amd.Numeric(sym, Ap, Ai, np.ones(4), fplus_amd.default_options(pivot_strategy=0, drop_tol=0.0, **opt))
    assert e.value.code == FPLUS_ERR_SINGULAR
    with pytest.warns(RuntimeWarning, match="ILU breakdown"):
        fplus_amd.SparseLU(sp.csc_matrix((Ax, Ai, Ap), shape=(2, 2)), pivot_strategy="full", drop_tol=1e-2, equilibrate=False)


def grid_convdiff(g, cplx):
    """the reference test's 5-point convection-diffusion operator on a g x g grid, row i's coefficient of unknown j
    -1.3 / -1.2 / 4 (+0.5i complex) / -0.8 / -0.7 for the neighbours below, left, itself, right, above
    (test_fplus.c:128-151)"""
    n = g * g
    rows, cols, vals = [], [], []
    for j in range(n):
        gx, gy = j % g, j // g
        for r, v, ok in ((j - g, -1.3, gy > 0), (j - 1, -1.2, gx > 0), (j, 4.0 + (0.5j if cplx else 0), True),
                         (j + 1, -0.8, gx < g - 1), (j + g, -0.7, gy < g - 1)):
            if ok:
                rows.append(r); cols.append(j); vals.append(v)
    A = sp.csc_matrix((np.array(vals, np.complex128 if cplx else np.float64), (rows, cols)), shape=(n, n))
    A.sort_indices()
    return n, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data


@pytest.mark.parametrize("cplx", [False, True])
def test_ilu_preconditioned_richardson(cplx):
    """every strategy with drop_tol 1e-2 keeps less fill than the exact factorization and, as a preconditioner of
    Richardson's iteration, reaches a backward error below 1e-12 within 60 iterations, on the reference's 20 x 20 grid
    (test_fplus.c:153-199)"""
    n, Ap, Ai, Ax = grid_convdiff(20, cplx)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    exact = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    fe = exact.factors()
    nnz_exact = fe["L"].nnz + fe["U"].nnz
    rng = np.random.default_rng(1)
    b = (2 * rng.random(n) - 1) + (1j * (2 * rng.random(n) - 1) if cplx else 0)
    for strategy in (0, 1, 2):
        print(f"ILU Richardson {'complex' if cplx else 'real'} {NAMES[strategy]}", flush=True)
        ilu = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(pivot_strategy=strategy, drop_tol=1e-2))
        f = ilu.factors()
        assert f["L"].nnz + f["U"].nnz < nnz_exact
        x = np.zeros(n, dtype=b.dtype)
        be = 1.0
        for it in range(60):
            x = x + ilu.solve(b - A @ x)
            be = acc.normwise(A, x, b)
            if be < 1e-13:
                break
        assert be < 1e-12, (strategy, be, it)


def test_extensions_refuse_a_row_handle():
    n, Ap, Ai, Ax = gen.convdiff2d(8)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(pivot_strategy=2))
    with pytest.raises(FplusError) as e:
        num.solve_refined(np.ones(n))
    assert e.value.code == FPLUS_ERR_BADARG and "row kernel" in str(e.value)
    err = C.create_string_buffer(256)
    rc = lib().fplus_amd_solve_device(num.ptr, None, 1, n, 0, err, 256)
    assert rc == FPLUS_ERR_BADARG
    with pytest.raises(FplusError):
        fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(pivot_strategy=2), fplus_amd.amd_options(refactor=1))


def test_batched_systems_match_one_by_one():
    """fplus_amd_factor_row_batch: independent systems of mixed sizes in shared launches (one workgroup each) give what
    fplus_factor_row gives for each alone, bit for bit, with per-system errors (a singular system, a non-canonical one)
    and per-system pool growth (an arrow matrix that fills in completely)"""
    rng = np.random.default_rng(9)
    systems = [c[1:] for c in CASES if not np.iscomplexobj(c[4])]
    for g in (6, 10, 17, 23):
        n, Ap, Ai, Ax = gen.convdiff2d(g, seed=g)
        systems.append((n, Ap, Ai, Ax))   # CSC arrays read as the CSR of the transpose
    n = 150
    A = sp.lil_matrix((n, n))
    A[0, :] = rng.uniform(0.5, 1.0, n)
    A[:, 0] = rng.uniform(0.5, 1.0, (n, 1))
    A.setdiag(np.concatenate([[10.0], rng.uniform(1.0, 2.0, n - 1)]))
    A[0, 0] = 10.0
    systems.append(csr_arrays(A.tocsr()))
    systems.append(csr_arrays(np.ones((3, 3))))                       # singular
    bad = list(csr_arrays(np.eye(3)))
    bad[2] = np.array([2, 0, 1], np.int64)                             # not canonical: row 0 holds columns 2, 0
    bad[1] = np.array([0, 2, 2, 3], np.int64)
    systems.append(tuple(bad))
    for strategy, dt in ((1, 1e-3), (2, 0.0), (0, 1e-2)):
        print(f"batch of {len(systems)} systems, {NAMES[strategy]} drop_tol {dt}", flush=True)
        got = fplus_amd.factor_row_batch(systems, dt, strategy)
        assert len(got) == len(systems)
        assert fplus_amd.row_stats()["runs"] >= len(systems) - 1   # every valid system ran at least once
        for i, (sysm, g) in enumerate(zip(systems, got)):
            try:
                one = fplus_amd.factor_row(*sysm, dt, strategy)
            except FplusError as e:
                assert isinstance(g, FplusError) and g.code == e.code, (i, g)
                continue
            assert not isinstance(g, FplusError), (i, g)
            for key in ("p", "q", "Lp", "Li", "Lx", "Up", "Ui", "Ux"):
                assert np.array_equal(g[key], one[key]), (i, key)
        if dt == 0.0:   # with dropping, the all-ones matrix's missing pivots are replaced instead (an ILU breakdown)
            assert isinstance(got[-2], FplusError) and got[-2].code == FPLUS_ERR_SINGULAR
        assert isinstance(got[-1], FplusError) and got[-1].code == FPLUS_ERR_BADARG
