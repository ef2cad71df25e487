"""Same-pattern refactorization, FP32 block updates and the refined solves (GMRES-IR) on the device.

- A refactorization with the values the factorization was made with replays it: the same pivot order and patterns, L
  and U to 1e-10 normwise -- the bound two factorizations of one matrix meet (the atomic updates of shared levels sum in
  the order the tiles run; test_gpu_paths measures 5e-12 on the n = 4096 pencil) -- while the FP64 path those values
  came from is itself pinned to the oracle.
- A refactorization with new values (another shift of the pencil) keeps the recorded pivot order and factors the new
  matrix exactly: ||Dr A[p, q] Dc - L U|| / ||Dr A[p, q] Dc|| at round-off (lu_residual), and the refined solve reaches
  an equilibrated backward error below 1e-14.
- FP32 updates (pivots, panel LU, Ruiz and solves in FP64): L and U differ from the FP64 factors by FP32 round-off
  (1e-10 to 1e-3 normwise: not zero, so the FP32 products ran, and not more), and GMRES-IR brings every right-hand side
  to an equilibrated backward error below 1e-14 (and a componentwise one below 1e-6, the direct FP64 solves' bound).
  The same for the BF16 hi/lo split products, whose factors lie 1e-8 to 1e-1 from the FP64 ones.
- Static pivoting: a refactorization whose recorded pivot is exactly zero for the new values fails with
  FPLUS_ERR_SINGULAR, or, with static_pivots = 1, replaces it, counts it, and the refined solve still converges.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import accuracy as acc
from fplus_amd import gen
from helpers import norm_diff, same_pattern, sorted_csc

pytestmark = pytest.mark.gpu


def _lu(num):
    f = num.factors()
    return sorted_csc(f["Lp"], f["Li"], f["Lx"]), sorted_csc(f["Up"], f["Ui"], f["Ux"]), f


def _residual(A, num):
    """lu_residual of the device factors, in the original coordinates"""
    f = num.factors()
    return fplus_amd.lu_residual(A, f["L"], f["U"], num.p, num.q, num.Dr, num.Dc)


@pytest.mark.parametrize("case", ["shiftinvert-8x4x64", "convdiff3d-12"])
def test_refactor_same_values_replays_the_factorization(case):
    if case.startswith("shift"):
        n, Ap, Ai, Ax = gen.shift_invert(8, 4, 64)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(12)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(refactor=1))
    L0, U0, f0 = _lu(num)
    p0 = num.p.copy()
    st = fplus_amd.stats(num)
    assert st["replay_bytes"] > 0
    num.refactor(Ax)
    L1, U1, f1 = _lu(num)
    assert np.array_equal(num.p, p0) and same_pattern(L0, L1) and same_pattern(U0, U1)
    dl, du = norm_diff(L1, L0), norm_diff(U1, U0)
    print(f"{case}: refactorization vs factorization: L {dl:.2e}, U {du:.2e}")
    assert dl < 1e-10 and du < 1e-10
    b = gen.rhs(n, 3, np.iscomplexobj(Ax))
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    m = acc.all_measures(A, num.solve(b), b, num.Dr, num.Dc)
    assert m["equilibrated"] < 1e-14, m


def test_refactor_new_shift_and_refined_solve():
    n, Ap, Ai, Ax0 = gen.shift_invert(8, 4, 64, sigma=0.3 + 0.7j)
    _, _, _, Ax1 = gen.shift_invert(8, 4, 64, sigma=0.3 + 0.5j)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax0, amd_options=fplus_amd.amd_options(refactor=1))
    p0 = num.p.copy()
    num.refactor(Ax1, static_pivots=1)
    assert np.array_equal(num.p, p0)
    st = fplus_amd.stats(num)
    A1 = sp.csc_matrix((Ax1, Ai, Ap), shape=(n, n))
    res = _residual(A1, num)
    print(f"new shift: lu_residual {res:.2e}, replaced pivots {st['refactor_replaced']}, min |pivot|/colmax "
          f"{st['refactor_min_ratio']:.2e}")
    assert res < 1e-12
    b = gen.rhs(n, 4, True)
    x, info = num.solve_refined(b)
    m = acc.all_measures(A1, x, b, num.Dr, num.Dc)
    print(f"new shift refined: {info}, {m}")
    assert info["converged"] == 1 and m["equilibrated"] < 1e-14, (info, m)


# the low-precision update forms: FP32 (precision 1) and the BF16 hi/lo split products (precision 4, about 16 bits per
# operand); the factors' distance from the FP64 ones lies in these bands
LOWP = {1: (1e-10, 1e-3), 4: (1e-8, 1e-1)}


@pytest.mark.parametrize("prec", [1, 4], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("case", ["shiftinvert-8x4x64", "convdiff3d-16"])
def test_fp32_updates_with_gmres_ir(case, prec):
    if case.startswith("shift"):
        n, Ap, Ai, Ax = gen.shift_invert(8, 4, 64)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(16)
    cplx = np.iscomplexobj(Ax)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num64 = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    num32 = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(precision=prec))
    assert fplus_amd.stats(num32)["precision"] == prec
    L64, U64, _ = _lu(num64)
    L32, U32, _ = _lu(num32)
    same = np.array_equal(num32.p, num64.p)
    if same:
        dl, du = norm_diff(L32, L64), norm_diff(U32, U64)
        print(f"{case}: precision {prec} updates vs FP64: L {dl:.2e}, U {du:.2e}")
        lo, hi = LOWP[prec]
        assert lo < max(dl, du) < hi, (dl, du)
    b = gen.rhs(n, 8, cplx)
    x0 = num32.solve(b)
    m0 = acc.all_measures(A, x0, b, num32.Dr, num32.Dc)
    x, info = num32.solve_refined(b)
    m = acc.all_measures(A, x, b, num32.Dr, num32.Dc)
    print(f"{case}: same pivots as FP64: {same}; direct solve {m0}; refined {info}, {m}")
    assert info["converged"] == 1 and info["outer"] >= 1
    assert m["equilibrated"] < 1e-14 and m["componentwise"] < 1e-6, m
    # the refined solve of an FP64 factorization takes no step (the first solve already meets the target) or one
    _, info64 = num64.solve_refined(b)
    assert info64["converged"] == 1 and info64["outer"] <= 1


def test_static_pivots():
    """a recorded pivot that the new values make exactly zero"""
    n, Ap, Ai, Ax = gen.convdiff2d(12, scale_exp=0.0)
    sym = fplus_amd.symbolic_from_orders(n, Ap, Ai, np.arange(n), np.arange(n))
    opts = fplus_amd.default_options(relax_rows=0, equilibrate=0)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, opts, amd_options=fplus_amd.amd_options(refactor=1))
    # column 0's recorded pivot row: its entry becomes zero (the pattern stays)
    r0 = int(num.p[0])
    Ax2 = Ax.copy()
    col0 = slice(Ap[0], Ap[1])
    Ax2[col0] = np.where(Ai[col0] == r0, 0.0, Ax2[col0])
    with pytest.raises(fplus_amd.FplusError) as e:
        num.refactor(Ax2, static_pivots=0)
    assert e.value.code == -2   # FPLUS_ERR_SINGULAR
    with pytest.raises(fplus_amd.FplusError):
        num.solve(np.ones(n))        # unusable after the failure
    num.refactor(Ax2, static_pivots=1)
    st = fplus_amd.stats(num)
    assert st["refactor_replaced"] >= 1
    A2 = sp.csc_matrix((Ax2, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 2, False)
    x, info = num.solve_refined(b, max_outer=20)
    m = acc.all_measures(A2, x, b, num.Dr, num.Dc)
    print(f"static pivots: replaced {st['refactor_replaced']}, refined {info}, {m}")
    assert info["converged"] == 1 and m["equilibrated"] < 1e-14


def test_pipelined_replay_matches_one_panel_after_another():
    """the pipelined replay (tuning knob refactor_pipeline: a panel's update levels that read no block of the previous
    panel run on a second stream while that panel finishes) gives the factors of the replay one panel after another --
    the same pivots and patterns, values to the atomic updates' round-off -- on a pencil of 32 panels"""
    n, Ap, Ai, Ax = gen.shift_invert(8, 4, 256, sigma=0.3 + 0.7j)
    _, _, _, Ax2 = gen.shift_invert(8, 4, 256, sigma=0.3 + 0.71j)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(refactor=1))
    A2 = sp.csc_matrix((Ax2, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 2, True)
    try:
        out = {}
        for pipe in (0, 1, 1):
            fplus_amd.tune("refactor_pipeline", pipe)
            num.refactor(Ax2)
            L, U, _ = _lu(num)
            out.setdefault(pipe, []).append((L, U, num.p.copy()))
            x, info = num.solve_refined(b)
            assert acc.equilibrated(A2, x, b, num.Dr, num.Dc) < 1e-14
        (L0, U0, p0), = out[0]
        for L1, U1, p1 in out[1]:
            assert np.array_equal(p1, p0) and same_pattern(L0, L1) and same_pattern(U0, U1)
            dl, du = norm_diff(L1, L0), norm_diff(U1, U0)
            print(f"pipelined vs one panel after another: L {dl:.2e}, U {du:.2e}")
            assert dl < 1e-11 and du < 1e-11
    finally:
        fplus_amd.tune("refactor_pipeline", 1)


def test_pipelined_replay_of_a_large_pencil():
    """a pencil of 131072 unknowns (512 panels) refactored, pipelined, in the pivot order of its neighbouring shift: the
    panel LU without row exchanges writes its factored pivot block back only after every workgroup has read it, also
    when the second stream's work delays some of its workgroups (a refactorization that read half-factored pivot blocks
    fails the pivot check or the accuracy bar)"""
    n, Ap, Ai, Ax = gen.shift_invert(32, 4, 1024, sigma=0.3 + 0.7j)
    _, _, _, Ax2 = gen.shift_invert(32, 4, 1024, sigma=0.3 + 0.69j)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax2, amd_options=fplus_amd.amd_options(refactor=1))
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 1, True)
    fplus_amd.tune("refactor_pipeline", 1)
    for _ in range(2):
        num.refactor(Ax)
        x, info = num.solve_refined(b)
        assert acc.equilibrated(A, x, b, num.Dr, num.Dc) < 1e-14
