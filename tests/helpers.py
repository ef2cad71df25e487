"""shared helpers of the tests: matrices, comparisons of device factors against the oracle"""
import os

import numpy as np
import scipy.sparse as sp

from fplus_amd import gen

REF_TEST_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sorted_csc(Lp, Li, Lx):
    """per column: rows ascending, values along (on copies: scipy keeps the caller's value array as its own data but may
    convert the indices, so sorting in place would permute the caller's values against its indices)"""
    n = len(Lp) - 1
    A = sp.csc_matrix((np.array(Lx, copy=True), np.array(Li, copy=True), np.array(Lp, copy=True)), shape=(n, n))
    A.sort_indices()
    return A


def same_pattern(A, B):
    return np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)


def rel_diff(A, B):
    """max over columns of ||a_j - b_j|| / ||b_j|| (entries on a common pattern)"""
    d = (A - B).tocsc()
    num = np.sqrt(np.asarray(abs(d).power(2).sum(axis=0))).ravel()
    den = np.sqrt(np.asarray(abs(B).power(2).sum(axis=0))).ravel()
    den[den == 0] = 1.0
    return float(np.max(num / den)) if len(num) else 0.0


def norm_diff(A, B):
    """||A - B||_F / ||B||_F (entries on a common pattern)"""
    nb = np.sqrt(float(abs(B).power(2).sum()))
    return float(np.sqrt(float(abs(A - B).power(2).sum())) / (nb if nb > 0 else 1.0))


def parity_matrices(small=True):
    """(name, n, Ap, Ai, Ax, relax) cases the device path is checked on"""
    out = []
    for cplx in (True, False):
        out.append((f"reftest-{'z' if cplx else 'd'}", *gen.reference_test_matrix(300, cplx, 6.0), 4))
    out.append(("convdiff2d-24", *gen.convdiff2d(24), 32))
    out.append(("convdiff2d-24-relax0", *gen.convdiff2d(24), 0))
    out.append(("convdiff3d-10", *gen.convdiff3d(10), 32))
    out.append(("shiftinvert-4x4x32", *gen.shift_invert(4, 4, 32), 32))
    if not small:
        out.append(("convdiff2d-96", *gen.convdiff2d(96), 32))
        out.append(("shiftinvert-8x4x64", *gen.shift_invert(8, 4, 64), 32))
    return out
