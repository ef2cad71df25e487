"""fplus_amd_solve_device_noalloc (the device solve with its scratch in the library's fixed per-GPU array) against
fplus_amd_solve_device, through the C interface on device buffers.

The two run the same kernels on the same factorization; the no-allocation form may take the right-hand sides in
smaller passes, and every right-hand side's arithmetic is independent of the others and of the pass width, so the
results agree bit for bit.  Two calls in a row, the second with a smaller input (another factorization, fewer
right-hand sides), catch stale contents of the shared array; the transposed solves use the array's other layout; a
limit below one right-hand side's scratch is refused with FPLUS_ERR_BADARG and leaves b untouched.
"""
import ctypes as C

import numpy as np
import pytest

import fplus_amd
from fplus_amd import gen
from fplus_amd._lib import lib

pytestmark = pytest.mark.gpu


def _dev(a):
    p = C.c_void_p()
    assert lib().fplus_amd_dev_alloc(a.nbytes, C.byref(p)) == 0
    assert lib().fplus_amd_dev_copy(p, C.c_void_p(a.ctypes.data), a.nbytes, 0) == 0
    return p


def _host(p, like):
    out = np.empty_like(like)
    assert lib().fplus_amd_dev_copy(C.c_void_p(out.ctypes.data), p, out.nbytes, 1) == 0
    return out


def _solve_both(num, b, trans=0):
    """(existing entry point's x, no-allocation entry point's x) for b (n x k, Fortran order)"""
    n, k = b.shape
    err = C.create_string_buffer(512)
    d1, d2 = _dev(b), _dev(b)
    try:
        rc = lib().fplus_amd_solve_device(num.ptr, d1, k, n, trans, err, 512)
        assert rc == 0, err.value
        rc = lib().fplus_amd_solve_device_noalloc(num.ptr, d2, k, n, trans, err, 512)
        assert rc == 0, err.value
        return _host(d1, b), _host(d2, b)
    finally:
        lib().fplus_amd_dev_free(d1)
        lib().fplus_amd_dev_free(d2)


def _factor(case):
    if case == "complex":
        n, Ap, Ai, Ax = gen.shift_invert(8, 4, 256, seed=3)
    else:
        n, Ap, Ai, Ax = gen.convdiff2d(48, seed=2)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    return fplus_amd.Numeric(sym, Ap, Ai, Ax), n, np.iscomplexobj(Ax)


def test_noalloc_solve_matches_solve_device_twice():
    big, n1, c1 = _factor("complex")
    small, n2, c2 = _factor("real")
    assert n2 < n1
    b1 = np.asfortranarray(gen.rhs(n1, 20, c1, seed=5))     # more than one pass of 16
    b2 = np.asfortranarray(gen.rhs(n2, 3, c2, seed=6))
    for num, b in ((big, b1), (small, b2)):
        x_ref, x_na = _solve_both(num, b)
        print(f"n {b.shape[0]}, nrhs {b.shape[1]}: max |x_noalloc - x| = {np.abs(x_na - x_ref).max():.3e}")
        assert np.array_equal(x_na, x_ref)
        assert not np.array_equal(x_ref, b)
    # the transposed solves, on the big factorization again
    for trans in (1, 2):
        x_ref, x_na = _solve_both(big, b1[:, :4].copy(order="F"), trans)
        assert np.array_equal(x_na, x_ref)


def test_noalloc_limit_is_enforced():
    num, n, cplx = _factor("real")
    b = np.asfortranarray(gen.rhs(n, 2, cplx, seed=7))
    err = C.create_string_buffer(512)
    d = _dev(b)
    try:
        assert lib().fplus_amd_tune(b"solve_noalloc_limit_bytes", 4096) == 0
        rc = lib().fplus_amd_solve_device_noalloc(num.ptr, d, 2, n, 0, err, 512)
        print(err.value.decode())
        assert rc == -3 and b"scratch" in err.value
        assert np.array_equal(_host(d, b), b)
    finally:
        assert lib().fplus_amd_tune(b"solve_noalloc_limit_bytes", 0) == 0
        lib().fplus_amd_dev_free(d)
    assert lib().fplus_amd_tune(b"solve_noalloc_limit_bytes", 1 << 40) != 0   # above the array
