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


def csr_arrays(A):
    """(n, indptr, indices, data) of a canonical CSR matrix"""
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A.shape[0], A.indptr.astype(np.int64), A.indices.astype(np.int64), A.data


def row_kernel_cases():
    """(name, n, Rp, Ri, Rx) CSR cases of the row kernel, equilibrated by the oracle's Ruiz (CSR route) as its pipeline
    would (src/pipeline.c:186-200)"""
    import oracle
    out = []
    for cplx in (False, True):
        n, Ap, Ai, Ax = gen.reference_test_matrix(120, cplx, 2.0)
        # the CSC arrays of A are the CSR arrays of A^T, as the reference's own test feeds them (test_fplus.c:236)
        out.append((f"reftest120-{'z' if cplx else 'd'}", n, Ap, Ai, Ax))
    n, Ap, Ai, Ax = gen.convdiff2d(12)
    out.append(("convdiff2d-12", n, Ap, Ai, Ax))
    rng = np.random.default_rng(5)
    A = sp.random(90, 90, density=0.05, random_state=rng, format="csr") + sp.eye(90) * 0.1
    out.append(("random90", *csr_arrays(A)))
    return [(name, n, Rp, Ri, oracle.ruiz(n, Rp, Ri, Rx, is_csc=False)[0]) for name, n, Rp, Ri, Rx in out]
