"""The oracle against the golden fixtures written from the reference's own numeric sources (tests/golden/ref_*.npz,
made by tests/golden/make_golden.py): Ruiz, the factorization and the sweeps bit-identical."""
import glob
import os

import numpy as np
import pytest

import oracle as O

FIXTURES = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "ref_*.npz")))


def _colsum(Lp, Lx):
    n = len(Lp) - 1
    s = np.add.reduceat(Lx, Lp[:-1]) if len(Lx) else np.zeros(n, Lx.dtype)
    a = np.add.reduceat(np.abs(Lx), Lp[:-1]) if len(Lx) else np.zeros(n)
    empty = np.diff(Lp) == 0
    s[empty] = 0
    a[empty] = 0
    return s, a


# the fixtures that hold their matrix (the mid-size one is regenerated from its seed, and checked by the device tests)
FULL = [p for p in FIXTURES if "Ap" in np.load(p).files]


@pytest.mark.parametrize("path", FULL, ids=lambda p: os.path.basename(p)[4:-4])
def test_oracle_matches_reference(path):
    z = np.load(path)
    n, Ap, Ai, Ax = int(z["n"]), z["Ap"], z["Ai"], z["Ax"]
    x, Dr, Dc, _ = O.ruiz(n, Ap, Ai, Ax)
    assert np.array_equal(Dr, z["Dr"]) and np.array_equal(Dc, z["Dc"])
    if "xscaled" in z:
        assert np.array_equal(np.asarray(x).view(np.float64), z["xscaled"].view(np.float64))
    f = O.factor_sn(n, Ap, Ai, x, relax_rows=int(z["relax"]))
    assert f["nsn"] == int(z["nsn"])
    for k in ("p", "Lp", "Up", "Ui"):
        assert np.array_equal(f[k], z[k]), k
    if "Lx" in z:
        assert np.array_equal(f["Li"], z["Li"])
        assert np.array_equal(f["Lx"].view(np.float64), z["Lx"].view(np.float64))
        assert np.array_equal(f["Ux"].view(np.float64), z["Ux"].view(np.float64))
    else:
        for (s, a), k in ((_colsum(f["Lp"], f["Lx"]), "L"), (_colsum(f["Up"], f["Ux"]), "U")):
            assert np.array_equal(s, z[k + "sum"]) and np.array_equal(a, z[k + "abs"]), k
    cplx = np.iscomplexobj(Ax)
    for tr in ((0, 1, 2) if cplx else (0, 1)):
        xs = np.array(z["b"], copy=True)
        rc = O.lib().orc_tri_solve(n, O._p64(f["Lp"]), O._p64(f["Li"]), O._pd(O._f64view(f["Lx"])), O._p64(f["Up"]),
                                   O._p64(f["Ui"]), O._pd(O._f64view(f["Ux"])), O._pd(xs.view(np.float64)), tr, int(cplx))
        assert rc == 0
        assert np.array_equal(xs.view(np.float64), z[f"x_trans{tr}"].view(np.float64)), f"trans {tr}"


def test_reference_build_agrees_when_present():
    """when oracle/_ref is built here, the oracle and the reference agree on a fresh seeded matrix too"""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built")
    from fplus_amd import gen
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 24, seed=5)
    x1, Dr1, Dc1, _ = O.ruiz(n, Ap, Ai, Ax)
    x2, Dr2, Dc2 = O.ref_ruiz(n, Ap, Ai, Ax)
    assert np.array_equal(Dr1, Dr2) and np.array_equal(Dc1, Dc2)
    f1 = O.factor_sn(n, Ap, Ai, x1, relax_rows=8)
    f2 = O.ref_factor_sn(n, Ap, Ai, x2, relax_rows=8)
    for k in ("p", "Lp", "Li", "Up", "Ui"):
        assert np.array_equal(f1[k], f2[k]), k
    assert np.array_equal(f1["Lx"].view(np.float64), f2["Lx"].view(np.float64))


def test_midsize_fixture_inputs():
    """the mid-size fixture's matrix is the seeded generator's (same SHA-256), its orders are permutations, and its
    solution of the original system solves it (componentwise backward error at the oracle's level)"""
    import scipy.sparse as sp
    from fplus_amd import accuracy as acc
    from fplus_amd import gen
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ref_shiftinvert_16x4x256.npz"))
    n, Ap, Ai, Ax = gen.shift_invert(16, 4, 256)
    assert gen.digest(n, Ap, Ai, Ax) == str(z["digest"])
    assert np.array_equal(np.sort(z["row_perm"]), np.arange(n)) and np.array_equal(np.sort(z["col_perm"]), np.arange(n))
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    Dr = np.empty(n)
    Dc = np.empty(n)
    Dr[z["row_perm"]] = z["Dr"]
    Dc[z["col_perm"]] = z["Dc"]
    m = acc.all_measures(A, z["xfull"], z["b"], Dr, Dc)
    assert m["equilibrated"] < 1e-15 and m["componentwise"] < 1e-9, m
