"""Same-pattern shift sweeps (fplus_amd_refactor_sweep) and the refactorization replayed through a captured HIP graph.

- A sweep over several shifts of the pencil, in the pivot order of one pivoting factorization, solves every shift's
  system: each shift's solutions meet an equilibrated backward error below 1e-14 and agree with those of a pivoting
  factorization made for that shift alone to 1e-8 relative.
- With the "refactor_graph" knob on, the first refactorization captures the replay and every later one launches the
  graph: the graph is replayed several times with the sweep's own eager work (the refined solves, downloads) between the
  replays, and every replay's results agree with the sweep launched kernel by kernel to 1e-10 (the update atomics of
  shared levels sum in launch order, the only difference between two runs).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import accuracy as acc
from fplus_amd import gen

pytestmark = pytest.mark.gpu

SHIFTS = [0.3 + 1j * (0.5 + k / 32) for k in range(5)]


def _sweep_case():
    n, Ap, Ai, Ax0 = gen.shift_invert(4, 4, 256, sigma=0.3 + 0.45j)
    axs = [gen.shift_invert(4, 4, 256, sigma=s)[3] for s in SHIFTS]
    bs = [gen.rhs(n, 2, True, seed=10 + k) for k in range(len(SHIFTS))]
    return n, Ap, Ai, Ax0, axs, bs


def test_sweep_matches_pivoting_factorizations():
    n, Ap, Ai, Ax0, axs, bs = _sweep_case()
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax0, amd_options=fplus_amd.amd_options(refactor=1))
    xs, infos = num.refactor_sweep(axs, bs, precision=0)
    for k, (ax, b, x, info) in enumerate(zip(axs, bs, xs, infos)):
        A = sp.csc_matrix((ax, Ai, Ap), shape=(n, n))
        ref = fplus_amd.Numeric(sym, Ap, Ai, ax)
        x_ref = ref.solve(b)
        m = acc.all_measures(A, x, b, ref.Dr, ref.Dc)
        d = np.abs(x - x_ref).max() / np.abs(x_ref).max()
        print(f"shift {SHIFTS[k]}: {info}, {m}, |x - x_pivoting| / |x| {d:.2e}")
        assert info["converged"] == 1 and info["fallback"] == 0
        assert m["equilibrated"] < 1e-14 and d < 1e-8


def test_sweep_through_captured_graph():
    n, Ap, Ai, Ax0, axs, bs = _sweep_case()
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax0, amd_options=fplus_amd.amd_options(refactor=1))
    x_eager, _ = num.refactor_sweep(axs, bs, precision=0)
    fplus_amd.tune("refactor_graph", 1)
    try:
        x_graph, infos = num.refactor_sweep(axs, bs, precision=0)
        # the graph once more, after other eager work of the program, on the first shift again
        x_again, _ = num.refactor_sweep(axs[:1], bs[:1], precision=0)
    finally:
        fplus_amd.tune("refactor_graph", 0)
    for k, (xe, xg) in enumerate(zip(x_eager, x_graph)):
        d = np.abs(xg - xe).max() / np.abs(xe).max()
        print(f"shift {k}: graph vs launched {d:.2e}, {infos[k]}")
        assert d < 1e-10 and infos[k]["converged"] == 1
    assert np.abs(x_again[0] - x_eager[0]).max() / np.abs(x_eager[0]).max() < 1e-10
