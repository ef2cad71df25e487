"""Entry points and code paths of the device factorization beyond fplus_factor's default route, each against the oracle.

- fplus_factor_sn_perm (the entry point a reference-side binding calls, reference _sparse_lu_c.c:495-545) and
  fplus_factor_sn_sym with seeded orders, fplus_factor_sn on input with duplicate entries (summed, reference
  src/factor_sn.c:277), the automatic relaxation rule (relax_rows = -1, reference src/factor_sn.c:636-649, the golden
  fixture written from the reference build);
- the one-workgroup panel LU (the fallback of the cooperative one, forced through the tuning knob);
- the mid-size member of the complex shift-invert family (n = 16384: 64 panels of 256 columns, shared dependency
  levels, 256-column blocks) against digests written from the reference build;
- two factorizations of one matrix: the same structure, values equal to round-off.

Tolerances as in test_gpu_parity: pivot order, partition and patterns exact; L and U values to 1e-12 normwise and 1e-10
per column.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
import oracle as O
from fplus_amd import accuracy as acc
from fplus_amd import gen
from helpers import norm_diff, rel_diff, same_pattern, sorted_csc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _same_lu(d, r, name):
    """d, r: dicts with p, nsn, Lp/Li/Lx, Up/Ui/Ux"""
    assert np.array_equal(d["p"], r["p"]), f"{name}: pivot order differs"
    assert d["nsn"] == r["nsn"], f"{name}: supernodes {d['nsn']} vs {r['nsn']}"
    Ld, Lo = sorted_csc(d["Lp"], d["Li"], d["Lx"]), sorted_csc(r["Lp"], r["Li"], r["Lx"])
    Ud, Uo = sorted_csc(d["Up"], d["Ui"], d["Ux"]), sorted_csc(r["Up"], r["Ui"], r["Ux"])
    assert same_pattern(Ld, Lo) and same_pattern(Ud, Uo), f"{name}: patterns differ"
    nl, nu, dl, du = norm_diff(Ld, Lo), norm_diff(Ud, Uo), rel_diff(Ld, Lo), rel_diff(Ud, Uo)
    print(f"{name}: L / U differ normwise by {nl:.2e} / {nu:.2e}, per column by {dl:.2e} / {du:.2e}")
    assert nl < 1e-12 and nu < 1e-12 and dl < 1e-10 and du < 1e-10, name


def test_factor_sn_perm_and_sym():
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 48, seed=3)
    rng = np.random.default_rng(17)
    rp, cp = rng.permutation(n), rng.permutation(n)
    d = fplus_amd.factor_sn_perm(n, Ap, Ai, Ax, rp, cp, relax_rows=8)
    ref = O.Numeric(n, Ap, Ai, Ax, rp, cp, relax_rows=8)
    assert np.array_equal(d["Dr"], ref.Dr[rp]) and np.array_equal(d["Dc"], ref.Dc[cp])   # permuted order
    _same_lu(d, ref.lu, "factor_sn_perm")
    sym = fplus_amd.symbolic_from_orders(n, Ap, Ai, rp, cp)
    d2 = fplus_amd.factor_sn_sym(sym, Ap, Ai, Ax, relax_rows=8)
    assert np.array_equal(d2["Dr"], d["Dr"]) and np.array_equal(d2["Dc"], d["Dc"])
    _same_lu(d2, ref.lu, "factor_sn_sym")
    # the analysis' own orders through fplus_factor_sn_sym
    sym3 = fplus_amd.Symbolic(n, Ap, Ai)
    d3 = fplus_amd.factor_sn_sym(sym3, Ap, Ai, Ax, relax_rows=32)
    ref3 = O.Numeric(n, Ap, Ai, Ax, sym3.row_perm, sym3.col_perm, relax_rows=32)
    _same_lu(d3, ref3.lu, "factor_sn_sym (analysis orders)")


def test_factor_sn_sums_duplicates():
    """raw fplus_factor_sn: entries given twice (in a column out of row order) are summed, as the reference kernel does
    while it scatters; no equilibration"""
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, True, 0.0, seed=12345)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    rng = np.random.default_rng(2)
    coo = A.tocoo()
    split = rng.random(coo.nnz) < 0.3
    w = rng.uniform(0.2, 0.8, coo.nnz)
    rows = np.concatenate([coo.row, coo.row[split]])
    cols = np.concatenate([coo.col, coo.col[split]])
    vals = np.concatenate([np.where(split, w * coo.data, coo.data), ((1 - w) * coo.data)[split]])
    order = np.lexsort((rng.random(len(rows)), cols))           # columns in order, rows shuffled inside each
    rows, cols, vals = rows[order], cols[order], vals[order]
    Dp = np.zeros(n + 1, np.int64)
    np.add.at(Dp, cols + 1, 1)
    Dp = np.cumsum(Dp)
    d = fplus_amd.factor_sn(n, Dp, rows.astype(np.int64), np.ascontiguousarray(vals), relax_rows=4)
    B = sp.csc_matrix((vals, rows, Dp), shape=(n, n))
    B.sum_duplicates()
    B.sort_indices()
    r = O.factor_sn(n, B.indptr.astype(np.int64), B.indices.astype(np.int64), np.ascontiguousarray(B.data), relax_rows=4)
    _same_lu(d, r, "factor_sn with duplicates")


def test_relax_auto_against_reference_fixture():
    """relax_rows = -1 (the automatic rule, with the small-sparse bounds: n = 512) against the reference build's factors"""
    z = np.load(os.path.join(GOLDEN, "ref_convdiff3d_8_auto.npz"))
    n = int(z["n"])
    assert int(z["relax"]) == -1
    sym = fplus_amd.symbolic_from_orders(n, z["Ap"], z["Ai"], np.arange(n), np.arange(n))
    num = fplus_amd.Numeric(sym, z["Ap"], z["Ai"], z["Ax"], fplus_amd.default_options(relax_rows=-1))
    assert np.array_equal(num.Dr, z["Dr"]) and np.array_equal(num.Dc, z["Dc"])
    f = num.factors()
    d = dict(p=f["p"], nsn=f["nsupernodes"], Lp=f["Lp"], Li=f["Li"], Lx=f["Lx"], Up=f["Up"], Ui=f["Ui"], Ux=f["Ux"])
    r = dict(p=z["p"], nsn=int(z["nsn"]), Lp=z["Lp"], Li=z["Li"], Lx=z["Lx"], Up=z["Up"], Ui=z["Ui"], Ux=z["Ux"])
    _same_lu(d, r, "relax -1")


@pytest.mark.parametrize("case", ["shiftinvert-4x4x32", "convdiff3d-10"])
def test_one_workgroup_panel_lu(case):
    """the single-workgroup panel LU (the fallback when the cooperative grid cannot be resident) against the oracle"""
    if case.startswith("shift"):
        n, Ap, Ai, Ax = gen.shift_invert(4, 4, 32)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(10)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    ref = O.Numeric(n, Ap, Ai, Ax, sym.row_perm, sym.col_perm, relax_rows=32)
    fplus_amd.tune("panel_lu_max_workgroups", 0)
    try:
        num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=32))
    finally:
        fplus_amd.tune("panel_lu_max_workgroups", 512)
    f = num.factors()
    d = dict(p=f["p"], nsn=f["nsupernodes"], Lp=f["Lp"], Li=f["Li"], Lx=f["Lx"], Up=f["Up"], Ui=f["Ui"], Ux=f["Ux"])
    _same_lu(d, ref.lu, f"{case} one-workgroup panel LU")
    assert np.array_equal(num.p, ref.p)


def test_midsize_shiftinvert_against_reference_digests():
    """n = 16384 (16 modes x 4 fields x 256 coefficients), the orders stored with the digests"""
    z = np.load(os.path.join(GOLDEN, "ref_shiftinvert_16x4x256.npz"))
    n, Ap, Ai, Ax = gen.shift_invert(16, 4, 256)
    assert gen.digest(n, Ap, Ai, Ax) == str(z["digest"])
    rp, cp = z["row_perm"], z["col_perm"]
    sym = fplus_amd.symbolic_from_orders(n, Ap, Ai, rp, cp)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, fplus_amd.default_options(relax_rows=32))
    st = fplus_amd.stats(num)
    print(f"mid-size: panels {st['panels']}, levels {st['levels']}, shared levels {st['shared_levels']}, widest block "
          f"{st['max_block_w']}, supernodes {st['nsupernodes']}")
    assert st["panels"] == 64 and st["shared_levels"] > 0 and st["max_block_w"] == 256
    assert np.array_equal(num.Dr[rp], z["Dr"]) and np.array_equal(num.Dc[cp], z["Dc"]), "Ruiz differs"
    assert np.array_equal(num.p, rp[z["p"]]), "pivot order differs"
    f = num.factors()
    assert f["nsupernodes"] == int(z["nsn"])
    assert np.array_equal(f["Lp"], z["Lp"]) and np.array_equal(f["Up"], z["Up"]), "column counts differ"

    def colsum(p, x):
        return np.add.reduceat(x, p[:-1]), np.add.reduceat(np.abs(x), p[:-1])

    for k, (p, x) in (("L", (f["Lp"], f["Lx"])), ("U", (f["Up"], f["Ux"]))):
        s, a = colsum(p, x)
        ds = np.max(np.abs(s - z[k + "sum"]) / z[k + "abs"])
        da = np.max(np.abs(a - z[k + "abs"]) / z[k + "abs"])
        print(f"mid-size {k}: column sums differ by {ds:.2e}, sums of moduli by {da:.2e} (relative to the sums of moduli)")
        assert ds < 1e-9 and da < 1e-9, k
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = z["b"]
    x = num.solve(b)
    m = acc.all_measures(A, x, b, num.Dr, num.Dc)
    mo = acc.all_measures(A, z["xfull"], b, num.Dr, num.Dc)
    d = np.linalg.norm(x - z["xfull"]) / np.linalg.norm(z["xfull"])
    print(f"mid-size solve: {m}; the reference's {mo}; |x - x_ref| / |x_ref| = {d:.2e}")
    assert m["equilibrated"] < 1e-14 and m["componentwise"] < max(100 * mo["componentwise"], 1e-12)
    assert d < 1e-8


def test_factor_twice_same_structure_and_values():
    """two factorizations of one matrix (with shared dependency levels): the same pivots and patterns, and values equal
    to round-off -- not bit for bit: on a level whose update tasks meet in a row the subtractions are atomic, so their
    order (and the last bits of the sums) follows the order the tiles run in, and the pivot growth carries those
    differences to about 5e-12 normwise on this n = 4096 pencil (bound: 1e-10, the per-column parity bound)"""
    n, Ap, Ai, Ax = gen.shift_invert(8, 4, 128, seed=4)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    outs = []
    for _ in range(2):
        num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
        f = num.factors()
        outs.append((num.p.copy(), f, fplus_amd.stats(num)))
    (p1, f1, st), (p2, f2, _) = outs
    assert st["shared_levels"] > 0
    assert np.array_equal(p1, p2) and np.array_equal(f1["Li"], f2["Li"]) and np.array_equal(f1["Ui"], f2["Ui"])
    nL = int(np.sum(f1["Lx"].view(np.float64) != f2["Lx"].view(np.float64)))
    nU = int(np.sum(f1["Ux"].view(np.float64) != f2["Ux"].view(np.float64)))
    dl = np.linalg.norm(f1["Lx"] - f2["Lx"]) / np.linalg.norm(f1["Lx"])
    du = np.linalg.norm(f1["Ux"] - f2["Ux"]) / np.linalg.norm(f1["Ux"])
    print(f"run to run: {nL} L and {nU} U doubles differ; normwise {dl:.2e} (L), {du:.2e} (U)")
    assert dl < 1e-10 and du < 1e-10


def test_deterministic_updates_are_bit_reproducible():
    """with the tuning knob deterministic_updates every dependency level is split so that no two of its update tasks
    write one workspace row: two factorizations of the pencil above, and two refactorizations replaying the recorded
    schedule, give the same factors bit for bit (the knob's levels hold no row written twice, so no atomics run), and
    the factors lie within the run-to-run bound of the default form's"""
    n, Ap, Ai, Ax = gen.shift_invert(8, 4, 128, seed=4)
    _, _, _, Ax2 = gen.shift_invert(8, 4, 128, seed=4, sigma=0.3 + 0.71j)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    base = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    fb = base.factors()
    lev0 = fplus_amd.stats(base)["levels"]
    try:
        fplus_amd.tune("deterministic_updates", 1)
        outs, refs = [], []
        for _ in range(2):
            num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(refactor=1))
            outs.append((num.p.copy(), num.factors(), fplus_amd.stats(num)))
            num.refactor(Ax2)
            refs.append(num.factors())
    finally:
        fplus_amd.tune("deterministic_updates", 0)
    (p1, f1, st), (p2, f2, _) = outs
    print(f"levels: {lev0} by default, {st['levels']} split")
    assert st["levels"] > lev0
    assert np.array_equal(p1, p2)
    for key in ("Lp", "Li", "Up", "Ui"):
        assert np.array_equal(f1[key], f2[key]), key
        assert np.array_equal(refs[0][key], refs[1][key]), key
    for key in ("Lx", "Ux"):
        assert np.array_equal(f1[key].view(np.float64), f2[key].view(np.float64)), key
        assert np.array_equal(refs[0][key].view(np.float64), refs[1][key].view(np.float64)), key
    assert np.array_equal(p1, base.p)
    assert np.linalg.norm(f1["Lx"] - fb["Lx"]) / np.linalg.norm(fb["Lx"]) < 1e-10
    assert np.linalg.norm(f1["Ux"] - fb["Ux"]) / np.linalg.norm(fb["Ux"]) < 1e-10


def test_pivot_ties_are_counted():
    """the panel LU counts the columns whose largest candidate modulus two rows hold exactly (fplus_amd_stats.pivot_ties:
    there the device's rule, the lowest workspace position, and the reference's, the first candidate of the column's
    search order, can choose differently): k blocks [[1, 1], [1, -1]] (Ruiz leaves them as they are) have one such
    column each, in the cooperative kernel and in the one-workgroup kernel; a matrix of random values has none"""
    k = 300
    A = sp.kron(sp.identity(k), sp.csc_matrix(np.array([[1.0, 1.0], [1.0, -1.0]]))).tocsc()
    A.sort_indices()
    n = A.shape[0]
    b = np.arange(1.0, n + 1)
    try:
        for cap in (512, 0):
            fplus_amd.tune("panel_lu_max_workgroups", cap)
            sym = fplus_amd.Symbolic(n, A.indptr, A.indices)
            num = fplus_amd.Numeric(sym, A.indptr, A.indices, A.data)
            assert fplus_amd.stats(num)["pivot_ties"] == k, cap
            x = num.solve(b)
            assert np.abs(A @ x - b).max() < 1e-12 * np.abs(b).max()
    finally:
        fplus_amd.tune("panel_lu_max_workgroups", 512)
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 64, seed=3)
    num = fplus_amd.Numeric(fplus_amd.Symbolic(n, Ap, Ai), Ap, Ai, Ax)
    assert fplus_amd.stats(num)["pivot_ties"] == 0


def test_sparselu_default_arguments_against_oracle():
    """SparseLU(A) with the reference's defaults: identity orders, relax 0, Ruiz to 1e-2 in at most 20 iterations"""
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 32, seed=2)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    lu = fplus_amd.SparseLU(A)
    ref = O.Numeric(n, Ap, Ai, Ax, np.arange(n), np.arange(n), relax_rows=0)
    assert np.array_equal(lu.perm_r, ref.p) and np.array_equal(lu.perm_c, np.arange(n))
    assert np.array_equal(lu.Dr, ref.Dr) and np.array_equal(lu.Dc, ref.Dc)
    f = lu._num.factors()
    d = dict(p=f["p"], nsn=f["nsupernodes"], Lp=f["Lp"], Li=f["Li"], Lx=f["Lx"], Up=f["Up"], Ui=f["Ui"], Ux=f["Ux"])
    _same_lu(d, ref.lu, "SparseLU defaults")
    b = gen.rhs(n, 2, True)
    for op in ("N", "T", "H"):
        m = acc.all_measures(A, lu.solve(b, op), b, lu.Dr, lu.Dc, op)
        assert m["equilibrated"] < 1e-14 and m["componentwise"] < 1e-8, (op, m)
    L, U, p, q, Dr, Dc = fplus_amd.sparse_lu_full_pivot(A, _return_csc=True)
    assert fplus_amd.lu_residual(A, L, U, p, q, Dr, Dc) < 1e-13
    x = fplus_amd.lu_solve(L, U, p, q, b, Dr, Dc)
    assert acc.equilibrated(A, x, b, Dr, Dc) < 1e-14
