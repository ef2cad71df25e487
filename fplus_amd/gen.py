"""Synthetic matrices of the benchmark and parity workloads, canonical int64 CSC, seeded numpy draws.

Every generator returns ``(n, Ap, Ai, Ax)`` with rows ascending in each column and no duplicates.  The draw order of
each generator is fixed and written in its docstring, so the same seed gives the same matrix on every machine.
"""
import hashlib

import numpy as np
import scipy.sparse as sp


def _csc_from_coo(n, rows, cols, vals):
    A = sp.csc_matrix((vals, (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return (n, A.indptr.astype(np.int64), A.indices.astype(np.int64), np.ascontiguousarray(A.data))


def digest(n, Ap, Ai, Ax):
    """SHA-256 of (Ap, Ai, Ax) as stored: the identity of a generated matrix"""
    h = hashlib.sha256()
    for a in (Ap, Ai, Ax):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def convdiff2d(g=96, seed=0, scale_exp=6.0):
    """Nonsymmetric 2D convection-diffusion, 5-point stencil on a g x g grid (n = g^2).

    Draws, in this order, from ``default_rng(seed)``: one U(0,1) per stored entry in canonical CSC order (the diagonal
    becomes 4 + 0.1 U, an off-diagonal -1 + 0.3 (2U - 1)); then the row scales 10^U(-s, s) (n draws) and the column
    scales 10^U(-s, s) (n draws); A <- diag(R) A diag(C).  Returns (n, Ap, Ai, Ax)."""
    n = g * g
    idx = np.arange(n).reshape(g, g)
    rows, cols = [idx.ravel()], [idx.ravel()]
    for di, dj in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        a = idx[max(di, 0):g + min(di, 0), max(dj, 0):g + min(dj, 0)]
        b = idx[max(-di, 0):g + min(-di, 0), max(-dj, 0):g + min(-dj, 0)]
        rows.append(a.ravel())
        cols.append(b.ravel())
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    _, Ap, Ai, _ = _csc_from_coo(n, r, c, np.ones(len(r)))
    rng = np.random.default_rng(seed)
    u = rng.random(len(Ai))
    colidx = np.repeat(np.arange(n), np.diff(Ap))
    Ax = np.where(Ai == colidx, 4.0 + 0.1 * u, -1.0 + 0.3 * (2.0 * u - 1.0))
    R = 10.0 ** rng.uniform(-scale_exp, scale_exp, n)
    Cs = 10.0 ** rng.uniform(-scale_exp, scale_exp, n)
    Ax = Ax * R[Ai] * Cs[colidx]
    return n, Ap, Ai, np.ascontiguousarray(Ax)


def convdiff3d(g=64, seed=0, peclet=10.0, scale_exp=4.0):
    """Upwind 3D convection-diffusion, 7-point stencil on a g^3 grid (n = g^3), real.

    Unit diffusion plus first-order upwind convection along (1, 1, 1) at cell Peclet number ``peclet``: the upwind
    neighbour of each axis carries -(1 + peclet), the downwind one -1, the diagonal 6 + 3 peclet.  Draws from
    ``default_rng(seed)``: the row scales 10^U(-s, s) (n draws), then the column scales (n draws)."""
    n = g ** 3
    idx = np.arange(n).reshape(g, g, g)
    rows, cols, vals = [idx.ravel()], [idx.ravel()], [np.full(n, 6.0 + 3.0 * peclet)]
    for ax in range(3):
        for step, v in ((-1, -(1.0 + peclet)), (1, -1.0)):
            sl_a = [slice(None)] * 3
            sl_b = [slice(None)] * 3
            if step < 0:
                sl_a[ax] = slice(1, g)
                sl_b[ax] = slice(0, g - 1)
            else:
                sl_a[ax] = slice(0, g - 1)
                sl_b[ax] = slice(1, g)
            a = idx[tuple(sl_a)].ravel()
            b = idx[tuple(sl_b)].ravel()
            rows.append(a)
            cols.append(b)
            vals.append(np.full(len(a), v))
    n_, Ap, Ai, Ax = _csc_from_coo(n, np.concatenate(rows), np.concatenate(cols), np.concatenate(vals))
    rng = np.random.default_rng(seed)
    R = 10.0 ** rng.uniform(-scale_exp, scale_exp, n)
    Cs = 10.0 ** rng.uniform(-scale_exp, scale_exp, n)
    colidx = np.repeat(np.arange(n), np.diff(Ap))
    return n, Ap, Ai, np.ascontiguousarray(Ax * R[Ai] * Cs[colidx])


def shift_invert(modes=64, nfields=4, ncoef=1024, seed=0, sigma=0.3 + 0.7j, scale_exp=8.0, scale_hi=None, hb=3):
    """Complex shift-invert pencil A = L - sigma M of a spectral problem (n = modes * nfields * ncoef).

    Unknown (m, f, k) = mode m, field f, coefficient k sits at index (m * nfields + f) * ncoef + k.  Inside a mode each
    field block is banded, |k - k'| <= hb, and the fields are coupled coefficient to coefficient (k' = k); neighbouring
    modes couple field to itself by a band |k - k'| <= 1.  Entries N(0,1) (1 + k)^2 (k the row coefficient), a quarter
    of that across modes.  The last two rows of
    every (mode, field) block are tau rows: dense over that field's ncoef coefficients in the mode, entries N(0,1)
    (they replace the band rows).  M is diagonal U(0.5, 1.5), zero on the tau rows.  Row and column scales
    10^U(-s, s_hi) (s_hi = s unless given).

    Draws from ``default_rng(seed)``, in order: the band entries of L (one N(0,1) per stored entry in canonical CSC
    order of the band pattern), the tau entries (one N(0,1) per tau entry, tau rows in index order, columns ascending),
    the diagonal of M (n draws of U(0.5, 1.5)), the row scales (n), the column scales (n)."""
    hi = scale_exp if scale_hi is None else scale_hi
    n = modes * nfields * ncoef
    blk = nfields * ncoef
    k = np.arange(ncoef)
    rr, cc = [], []
    # in-mode field coupling bands
    for m in range(modes):
        for f in range(nfields):
            for g in range(nfields):
                for d in (range(-hb, hb + 1) if g == f else (0,)):
                    kk = k[(k + d >= 0) & (k + d < ncoef) & (k < ncoef - 2)]
                    rr.append(m * blk + f * ncoef + kk)
                    cc.append(m * blk + g * ncoef + kk + d)
        for mm in (m - 1, m + 1):
            if mm < 0 or mm >= modes:
                continue
            for f in range(nfields):
                for d in (-1, 0, 1):
                    kk = k[(k + d >= 0) & (k + d < ncoef) & (k < ncoef - 2)]
                    rr.append(m * blk + f * ncoef + kk)
                    cc.append(mm * blk + f * ncoef + kk + d)
    r = np.concatenate(rr)
    c = np.concatenate(cc)
    B = sp.csc_matrix((np.ones(len(r)), (r, c)), shape=(n, n))
    B.sum_duplicates()
    B.sort_indices()
    rng = np.random.default_rng(seed)
    bi = B.indices.astype(np.int64)
    bj = np.repeat(np.arange(n), np.diff(B.indptr))
    coefrow = bi % ncoef
    same_mode = (bi // blk) == (bj // blk)
    vals = rng.standard_normal(len(bi)) * (1.0 + coefrow) ** 2
    vals = np.where(same_mode, vals, 0.25 * vals)
    # tau rows
    tr, tc = [], []
    for m in range(modes):
        for f in range(nfields):
            for t in (ncoef - 2, ncoef - 1):
                tr.append(np.full(ncoef, m * blk + f * ncoef + t))
                tc.append(m * blk + f * ncoef + k)
    tr = np.concatenate(tr)
    tc = np.concatenate(tc)
    tv = rng.standard_normal(len(tr))
    Md = rng.uniform(0.5, 1.5, n)
    tau_row = (np.arange(n) % ncoef) >= ncoef - 2
    Md[tau_row] = 0.0
    R = 10.0 ** rng.uniform(-scale_exp, hi, n)
    Cs = 10.0 ** rng.uniform(-scale_exp, hi, n)
    rows = np.concatenate([bi, tr, np.arange(n)])
    cols = np.concatenate([bj, tc, np.arange(n)])
    v = np.concatenate([vals.astype(np.complex128), tv.astype(np.complex128), -sigma * Md])
    n_, Ap, Ai, Ax = _csc_from_coo(n, rows, cols, v)
    colidx = np.repeat(np.arange(n), np.diff(Ap))
    Ax = Ax * R[Ai] * Cs[colidx]
    return n, Ap, Ai, np.ascontiguousarray(Ax)


def reference_test_matrix(n=300, cplx=True, s=6.0, seed=88172645463325252):
    """The matrix family of the reference's C tests (tests/c/test_fplus.c:22-49): a band of half-width 2, three random
    entries per column, dense rows 1 and n/2, dense column n/3, row scales 10^[-s, s].  Drawn with the same xorshift
    generator and draw order, so the same seed reproduces the same matrices."""
    state = [seed & 0xFFFFFFFFFFFFFFFF]
    M = 0xFFFFFFFFFFFFFFFF

    def urand():
        x = state[0]
        x ^= (x << 13) & M
        x ^= x >> 7
        x ^= (x << 17) & M
        state[0] = x
        return (x >> 11) / 9007199254740992.0

    dense = np.zeros((n, n), dtype=bool)          # dense[j, i]: column j holds row i
    for j in range(n):
        for d in range(-2, 3):
            if 0 <= j + d < n:
                dense[j, j + d] = True
        for _ in range(3):
            dense[j, int(urand() * n)] = True
    dense[:, 1] = True
    dense[:, n // 2] = True
    dense[n // 3, :] = True
    rs = np.array([10.0 ** (s * (2 * urand() - 1)) for _ in range(n)])
    Ap = [0]
    Ai, Ax = [], []
    for j in range(n):
        for i in range(n):
            if dense[j, i]:
                re = 2 * urand() - 1 + (0.5 if i == j else 0.0)
                im = 2 * urand() - 1
                Ai.append(i)
                Ax.append(rs[i] * (re + 1j * im) if cplx else rs[i] * re)
        Ap.append(len(Ai))
    Ax = np.array(Ax, dtype=np.complex128 if cplx else np.float64)
    return n, np.array(Ap, np.int64), np.array(Ai, np.int64), Ax


def rhs(n, nrhs, cplx, seed=1):
    rng = np.random.default_rng(seed)
    b = rng.standard_normal((n, nrhs))
    if cplx:
        b = b + 1j * rng.standard_normal((n, nrhs))
    return np.asfortranarray(b)
