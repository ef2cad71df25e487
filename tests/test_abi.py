"""The C interface without a GPU: the library loads and exports every symbol include/*.h declares, the option
defaults and struct layouts of the reference interface, the host symbolic analysis, and the argument checks that
reject bad input before any device work (reference tests/c/test_fplus.c:268-279)."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import _lib, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"\b(fplus_\w+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 35
    missing = [nm for nm in names if not hasattr(L, nm)]
    assert not missing, f"not exported: {missing}"
    assert set(names) == set(_lib.SIGNATURES), "the ctypes table and the headers differ"


def test_options_defaults_and_layout():
    o = fplus_amd.default_options()
    assert (o.pin_dense, o.tau_density, o.tau_min_nnz, o.postorder) == (1, 0.5, 10, 1)
    assert (o.equilibrate, o.ruiz_max_iter, o.ruiz_tol) == (1, 20, 1e-2)
    assert (o.pivot_rel, o.pivot_strategy, o.drop_tol) == (1.4901161193847656e-08, 1, 0.0)
    assert (o.relax_rows, o.panel, o.gpu_min_flops, o.verbose) == (32, 512, 5e7, 0)



def _c_layout(include_dir, tmp_path):
    """sizeof and field offsets of fplus_options / fplus_factors as gcc lays them out from a header"""
    import subprocess
    fields = {"fplus_options": [f for f, _ in _lib.Options._fields_], "fplus_factors": [f for f, _ in _lib.Factors._fields_]}
    src = ["#include <stdio.h>", "#include <stddef.h>", '#include "fplus.h"', "int main(void){"]
    for st, fs in fields.items():
        src.append(f'printf("%zu\\n", sizeof({st}));')
        for f in fs:
            src.append(f'printf("%zu\\n", offsetof({st}, {f}));')
    src.append("return 0;}")
    c = tmp_path / "layout.c"
    c.write_text("\n".join(src))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", include_dir, str(c), "-o", str(exe)], check=True)
    return [int(v) for v in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()]


def test_struct_layouts(tmp_path):
    """the ctypes mirrors and include/fplus.h agree with gcc's layout, and with the reference header when present"""
    mine = _c_layout(os.path.join(ROOT, "include"), tmp_path)
    ct = [C.sizeof(_lib.Options)] + [getattr(_lib.Options, f).offset for f, _ in _lib.Options._fields_]
    ct += [C.sizeof(_lib.Factors)] + [getattr(_lib.Factors, f).offset for f, _ in _lib.Factors._fields_]
    assert mine == ct
    if os.path.isdir("/root/reference/include"):
        assert _c_layout("/root/reference/include", tmp_path) == mine


def test_analyze_pins_dense_lines_and_orders():
    n, Ap, Ai, Ax = gen.reference_test_matrix(300, True, 6.0)
    s = fplus_amd.Symbolic(n, Ap, Ai)
    assert s.n_dense_rows >= 2 and s.n_dense_cols >= 1        # tests/c/test_fplus.c:100-101
    assert sorted(s.row_perm) == list(range(n)) and sorted(s.col_perm) == list(range(n))
    counts = np.bincount(Ai, minlength=n)
    dense_rows = set(np.nonzero(counts > 150)[0])
    assert set(s.row_perm[n - len(dense_rows):]) == dense_rows, "dense rows go last"
    dense_cols = set(np.nonzero(np.diff(Ap) > 150)[0])
    assert set(s.col_perm[n - len(dense_cols):]) == dense_cols, "dense columns go last"


def test_coletree_postorder_is_a_postorder():
    n, Ap, Ai, Ax = gen.convdiff2d(12)
    post = np.empty(n, np.int64)
    rc = _lib.lib().fplus_coletree_postorder(n, n, Ap.ctypes.data_as(_lib.I64P), Ai.ctypes.data_as(_lib.I64P),
                                             post.ctypes.data_as(_lib.I64P))
    assert rc == 0 and sorted(post) == list(range(n))
    # the column etree of A (of A^T A): parent[j] = min over rows r of col j of the next column sharing r in A^T A
    A = sp.csc_matrix((np.ones(len(Ai)), Ai, Ap), shape=(n, n))
    M = (A.T @ A).tocsc()
    parent = np.full(n, -1)
    L = np.linalg.cholesky(M.toarray() + n * np.eye(n))    # symbolic structure of the Cholesky factor of A^T A
    for j in range(n):
        below = np.nonzero(np.abs(L[j + 1:, j]) > 1e-12)[0]
        parent[j] = j + 1 + below[0] if len(below) else -1
    pos = np.empty(n, np.int64)
    pos[post] = np.arange(n)
    for j in range(n):
        if parent[j] >= 0:
            assert pos[j] < pos[parent[j]], "a column before its parent"


def test_min_degree_reduces_fill():
    """the ordering beats the natural order on a 2D grid: nnz(L + U) of the oracle's factorization with each order"""
    import oracle as O
    n, Ap, Ai, Ax = gen.convdiff2d(60)
    s = fplus_amd.Symbolic(n, Ap, Ai)

    def fill(rp, cp):
        num = O.Numeric(n, Ap, Ai, Ax, rp, cp, relax_rows=0)
        return num.lu["Lp"][-1] + num.lu["Up"][-1]

    ident = np.arange(n)
    assert fill(s.row_perm, s.col_perm) < 0.7 * fill(ident, ident)


def _err():
    return C.create_string_buffer(256)


def test_validation_without_device():
    """bad input is rejected before any device work (tests/c/test_fplus.c:268-279)"""
    L = _lib.lib()
    p = np.array([0, 2, 3], np.int64)
    i_bad = np.array([0, 5, 1], np.int64)
    i_unsorted = np.array([1, 0, 1], np.int64)
    x = np.array([1, 2, 3], np.complex128)
    f = _lib.Factors()
    e = _err()
    P = _lib.I64P
    rc = L.fplus_factor_sn(2, p.ctypes.data_as(P), i_bad.ctypes.data_as(P), x.ctypes.data, 1, 0, 0, 1, 0, 5e7, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"out of range" in e.value
    sym = C.c_void_p()
    rc = L.fplus_analyze(2, p.ctypes.data_as(P), i_unsorted.ctypes.data_as(P), 1, 0.5, 10, None, None, 1, C.byref(sym), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG
    rc = L.fplus_factor_sn(-1, p.ctypes.data_as(P), i_bad.ctypes.data_as(P), x.ctypes.data, 1, 0, 0, 1, 0, 5e7, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG
    xn = np.array([1, np.nan, 3], np.complex128)
    i_ok = np.array([0, 1, 1], np.int64)
    rc = L.fplus_factor_sn(2, p.ctypes.data_as(P), i_ok.ctypes.data_as(P), xn.ctypes.data, 1, 0, 0, 1, 0, 5e7, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"non-finite" in e.value
    rp = np.array([0, 0], np.int64)
    rc = L.fplus_factor_sn_perm(2, p.ctypes.data_as(P), i_ok.ctypes.data_as(P), x.ctypes.data, 1, rp.ctypes.data_as(P),
                                rp.ctypes.data_as(P), 1, 20, 1e-2, 0, 32, 512, 1, 5e7, 1e-8, C.byref(f), None, None, e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"permutation" in e.value


def test_row_kernel_validation_without_device():
    """the row kernel's argument checks come before any device work, with the reference's messages
    (src/factor_row.c:1200-1210, src/numeric.c:51-57)"""
    L = _lib.lib()
    P = _lib.I64P
    p = np.array([0, 2, 3], np.int64)
    i_ok = np.array([0, 1, 1], np.int64)
    i_unsorted = np.array([1, 0, 1], np.int64)
    x = np.array([1.0, 2.0, 3.0])
    f = _lib.Factors()
    e = _err()
    args = (p.ctypes.data_as(P), i_ok.ctypes.data_as(P), x.ctypes.data)
    rc = L.fplus_factor_row(2, *args, 0.0, 0, 0, 3, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"pivot_strategy must be 0 (full), 1 (partial), or 2 (rook); got 3" in e.value
    rc = L.fplus_factor_row(2, *args, -1.0, 0, 0, 2, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"drop_tol must be >= 0" in e.value
    rc = L.fplus_factor_row(2, *args, 0.0, 0, 0, 2, -1.0, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"pivot_rel must be >= 0" in e.value
    rc = L.fplus_factor_row(2, p.ctypes.data_as(P), i_unsorted.ctypes.data_as(P), x.ctypes.data, 0.0, 0, 0, 2, 1e-8, C.byref(f), e, 256)
    assert rc == _lib.FPLUS_ERR_BADARG and b"strictly increasing" in e.value
    n, Ap, Ai, Ax = gen.convdiff2d(4)
    s = fplus_amd.Symbolic(n, Ap, Ai)
    for kw in (dict(pivot_strategy=7), dict(drop_tol=-1.0)):
        with pytest.raises(fplus_amd.FplusError) as er:
            fplus_amd.Numeric(s, Ap, Ai, Ax, fplus_amd.default_options(**kw))
        assert er.value.code == _lib.FPLUS_ERR_BADARG


def test_generators_are_deterministic():
    assert gen.digest(*gen.convdiff2d(96)) == gen.digest(*gen.convdiff2d(96))
    n, Ap, Ai, Ax = gen.convdiff2d(96)
    assert n == 9216 and len(Ai) == 45696
    n, Ap, Ai, Ax = gen.convdiff3d(8)
    assert len(Ai) == 7 * 512 - 6 * 64


def test_tuning_knobs_accept_their_range_and_reject_the_rest():
    """fplus_amd_tune (include/fplus_amd.h): every on/off knob takes 0 and 1 and returns FPLUS_ERR_BADARG for 2, the
    numeric ones their documented ranges, unknown names FPLUS_ERR_BADARG; the defaults are restored afterwards"""
    L = _lib.lib()
    defaults = {"refactor_graph": 0, "refactor_pipeline": 1, "deterministic_updates": 0, "update_atomic_rows": 1,
                "solve_dataflow": 0, "xcd_tile_order": 1}
    try:
        for name in defaults:
            for v in (0, 1):
                assert L.fplus_amd_tune(name.encode(), v) == 0, (name, v)
            assert L.fplus_amd_tune(name.encode(), 2) == _lib.FPLUS_ERR_BADARG, name
        assert L.fplus_amd_tune(b"panel_lu_max_workgroups", 0) == 0
        assert L.fplus_amd_tune(b"panel_lu_max_workgroups", 513) == _lib.FPLUS_ERR_BADARG
        for p in range(5):
            assert L.fplus_amd_tune(b"update_tile_policy", p) == 0
        assert L.fplus_amd_tune(b"update_tile_policy", 5) == _lib.FPLUS_ERR_BADARG
        assert L.fplus_amd_tune(b"no_such_knob", 0) == _lib.FPLUS_ERR_BADARG
        assert L.fplus_amd_tune(None, 0) == _lib.FPLUS_ERR_BADARG
    finally:
        for name, v in defaults.items():
            L.fplus_amd_tune(name.encode(), v)
        L.fplus_amd_tune(b"panel_lu_max_workgroups", 512)
        L.fplus_amd_tune(b"update_tile_policy", 0)
