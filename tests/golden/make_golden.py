"""Writes the golden fixtures of tests/golden/ from the reference's own numeric sources.

The reference's Ruiz, supernodal kernel and sweeps (src/equil.c, src/factor_sn.c, src/solve.c, src/solve_plan.c) are
compiled where they lie under /root/reference by oracle/ref.mk into oracle/_ref/libfplus_ref.so; this script runs them
on small seeded inputs and stores inputs and outputs as .npz data.  Run in the development container only (the
reference tree is not on the GPU box):

    make -C oracle ref && python tests/golden/make_golden.py

Large outputs are stored as digests (the pivot order, the partition, the patterns' sizes and per-column checksums)
rather than in full; the mid-size case stores no matrix either (it is regenerated from its seed, and its SHA-256 is
kept), only the orders it was factored with.  Name cases on the command line to write only those.
"""
import os
import sys

import numpy as np
import scipy.sparse as sp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from fplus_amd import gen  # noqa: E402


def seeded_perm(n, seed):
    return np.random.default_rng(seed).permutation(n).astype(np.int64)


def permuted(n, Ap, Ai, Ax, rp, cp):
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))[rp, :][:, cp].tocsc()
    A.sort_indices()
    return A.indptr.astype(np.int64), A.indices.astype(np.int64), np.ascontiguousarray(A.data)


def col_checksums(Lp, Lx):
    """per column: the sum of the values and the sum of their moduli"""
    n = len(Lp) - 1
    s = np.add.reduceat(Lx, Lp[:-1]) if len(Lx) else np.zeros(n, Lx.dtype)
    a = np.add.reduceat(np.abs(Lx), Lp[:-1]) if len(Lx) else np.zeros(n)
    empty = np.diff(Lp) == 0
    s[empty] = 0
    a[empty] = 0
    return s, a


def cases():
    out = []
    for cplx in (1, 0):
        n, Ap, Ai, Ax = gen.reference_test_matrix(300, bool(cplx), 6.0)
        out.append((f"reftest_{'z' if cplx else 'd'}", n, Ap, Ai, Ax, 4, "full"))
    n, Ap, Ai, Ax = gen.convdiff2d(24)
    rp, cp = seeded_perm(n, 11), seeded_perm(n, 12)
    out.append(("convdiff2d_24_perm", n, *permuted(n, Ap, Ai, Ax, rp, cp), 32, "full"))
    out.append(("convdiff2d_24_relax0", n, Ap, Ai, Ax, 0, "full"))
    n, Ap, Ai, Ax = gen.shift_invert(4, 4, 32)
    out.append(("shiftinvert_4x4x32", n, Ap, Ai, Ax, 32, "full"))
    n, Ap, Ai, Ax = gen.convdiff3d(8)
    out.append(("convdiff3d_8_auto", n, Ap, Ai, Ax, -1, "full"))
    # a mid-size member of the complex shift-invert family (n = 16384, 64 panels of 256 columns on the device), with the
    # orders of this build's own analysis stored beside the digests (the test factors with exactly these orders); the
    # matrix itself is regenerated from the seeded generator and identified by its SHA-256
    n, Ap, Ai, Ax = gen.shift_invert(16, 4, 256)
    out.append(("shiftinvert_16x4x256", n, Ap, Ai, Ax, 32, "orders"))
    for name in ("relaxed_panel_case3", "relaxed_panel_case50"):
        z = np.load(os.path.join(HERE, name + ".npz"))
        A = sp.csr_matrix((z["data"], z["indices"], z["indptr"]), shape=(1200, 1200)).tocsc()
        A.sort_indices()
        out.append((name + "_r2", 1200, A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data, 2, "digest"))
    return out


def main(only=None):
    if not O.ref_available():
        sys.exit("oracle/_ref/libfplus_ref.so is not built: make -C oracle ref")
    for name, n, Ap, Ai, Ax, relax, kind in cases():
        if only and name not in only:
            continue
        A0 = (Ap, Ai, Ax)
        if kind == "orders":
            import fplus_amd
            sym = fplus_amd.Symbolic(n, Ap, Ai)
            rp, cp = sym.row_perm.copy(), sym.col_perm.copy()
            Ap, Ai, Ax = permuted(n, Ap, Ai, Ax, rp, cp)
        x, Dr, Dc = O.ref_ruiz(n, Ap, Ai, Ax)
        f = O.ref_factor_sn(n, Ap, Ai, x, relax_rows=relax)
        rng = np.random.default_rng(3)
        cplx = np.iscomplexobj(Ax)
        b = rng.standard_normal(n) + (1j * rng.standard_normal(n) if cplx else 0)
        out = dict(n=n, relax=relax, Dr=Dr, Dc=Dc, p=f["p"], nsn=f["nsn"], Lp=f["Lp"], Up=f["Up"], b=b)
        if kind == "orders":   # the matrix is regenerated by the test: its digest, and the orders used
            out.update(digest=gen.digest(n, *A0), row_perm=rp, col_perm=cp)
        else:
            out.update(Ap=Ap, Ai=Ai, Ax=Ax)
        for tr in ((0, 1, 2) if cplx else (0, 1)):
            xs = np.array(b, copy=True)
            rc = O.ref().fplus_tri_solve(n, O._p64(f["Lp"]), O._p64(f["Li"]), f["Lx"].ctypes.data, O._p64(f["Up"]),
                                        O._p64(f["Ui"]), f["Ux"].ctypes.data, xs.ctypes.data, tr, int(cplx))
            assert rc == 0
            out[f"x_trans{tr}"] = xs
        if kind == "full":
            out.update(xscaled=x, Li=f["Li"], Lx=f["Lx"], Ui=f["Ui"], Ux=f["Ux"])
        elif kind == "orders":
            # the solution of the original system A x = b through the reference's composition (src/numeric.c:88-131):
            # c = Dr b[row_perm] in pivot order, z = U^-1 L^-1 c, x[col_perm] = Dc z
            c = (Dr * b[rp])[f["p"]]
            rc = O.ref().fplus_tri_solve(n, O._p64(f["Lp"]), O._p64(f["Li"]), f["Lx"].ctypes.data, O._p64(f["Up"]),
                                        O._p64(f["Ui"]), f["Ux"].ctypes.data, c.ctypes.data, 0, int(cplx))
            assert rc == 0
            xf = np.empty_like(c)
            xf[cp] = Dc * c
            out.update(xfull=xf)
            out.update(Lsum=col_checksums(f["Lp"], f["Lx"])[0], Labs=col_checksums(f["Lp"], f["Lx"])[1],
                       Usum=col_checksums(f["Up"], f["Ux"])[0], Uabs=col_checksums(f["Up"], f["Ux"])[1])
        else:
            out.update(Lsum=col_checksums(f["Lp"], f["Lx"])[0], Labs=col_checksums(f["Lp"], f["Lx"])[1],
                       Usum=col_checksums(f["Up"], f["Ux"])[0], Uabs=col_checksums(f["Up"], f["Ux"])[1],
                       Ui=f["Ui"])
        path = os.path.join(HERE, f"ref_{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: n={n} nnzL={f['Lp'][-1]} nnzU={f['Up'][-1]} nsn={f['nsn']}")


if __name__ == "__main__":
    main(sys.argv[1:] or None)   # names of cases to write (default: all)
