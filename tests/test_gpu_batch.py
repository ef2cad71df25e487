"""Independent systems side by side: factorizations and solves on several host threads, each sending its device work to
a stream of its own (fplus_amd_thread_stream) with a smaller host pool per factorization (tuning knob "host_threads"),
against the same systems one after another on the process's stream: the same pivots and supernodes, solutions to
1e-10 (the factorization's shared-level sums can round differently from run to run), equilibrated backward errors
below 1e-14."""
import threading

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import _lib
from fplus_amd import accuracy as acc
from fplus_amd import gen

pytestmark = pytest.mark.gpu


def _systems():
    out = []
    for i, (m, c) in enumerate(((2, 256), (4, 256), (1, 1024), (3, 512), (2, 512), (4, 128))):
        n, Ap, Ai, Ax = gen.shift_invert(m, 4, c, seed=70 + i)
        out.append((n, Ap, Ai, Ax, gen.rhs(n, 3, True, seed=80 + i)))
    return out


def _solve(case):
    n, Ap, Ai, Ax, b = case
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    return num.p.copy(), num.factors()["nsupernodes"], num.solve(b), num.Dr.copy(), num.Dc.copy()


def test_threads_with_own_streams():
    cases = _systems()
    seq = [_solve(c) for c in cases]
    res = [None] * len(cases)
    errs = []

    def worker(k):
        try:
            _lib.check(_lib.lib().fplus_amd_thread_stream(1), None)
            for i in range(k, len(cases), 3):
                res[i] = _solve(cases[i])
        except Exception as e:   # noqa: BLE001 -- reported below
            errs.append(e)
        finally:
            _lib.lib().fplus_amd_thread_stream(0)

    fplus_amd.tune("host_threads", 4)
    try:
        th = [threading.Thread(target=worker, args=(k,)) for k in range(3)]
        for t in th:
            t.start()
        for t in th:
            t.join()
    finally:
        fplus_amd.tune("host_threads", 0)
    assert not errs, errs
    for i, (c, a, s) in enumerate(zip(cases, res, seq)):
        n, Ap, Ai, Ax, b = c
        assert np.array_equal(a[0], s[0]) and a[1] == s[1], f"system {i}: pivots or supernodes differ"
        d = np.max(np.linalg.norm(a[2] - s[2], axis=0) / np.linalg.norm(s[2], axis=0))
        be = acc.equilibrated(sp.csc_matrix((Ax, Ai, Ap), shape=(n, n)), a[2], b, a[3], a[4])
        print(f"system {i}: n {n}, side by side vs one after another {d:.1e}, equilibrated backward error {be:.1e}")
        assert d < 1e-10 and be < 1e-14, i
