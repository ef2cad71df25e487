"""The update-tile autotuner (fplus_amd.autotune): every tile policy gives the same factorization -- pivot order,
partition and patterns equal, values within the 1e-10 run-to-run bound of the update sums -- and the tuner leaves the
fastest policy in force."""
import numpy as np
import pytest

import fplus_amd
from fplus_amd import autotune as at
from fplus_amd import gen
from helpers import norm_diff, same_pattern, sorted_csc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["shiftinvert-8x4x128", "convdiff3d-16"])
def test_tile_policies_agree_and_autotune(case):
    if case.startswith("shift"):
        n, Ap, Ai, Ax = gen.shift_invert(8, 4, 128)
    else:
        n, Ap, Ai, Ax = gen.convdiff3d(16)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    try:
        ref = None
        for p in sorted(at.POLICIES):
            fplus_amd.tune("update_tile_policy", p)
            num = fplus_amd.Numeric(sym, Ap, Ai, Ax)
            f = num.factors()
            L, U = sorted_csc(f["Lp"], f["Li"], f["Lx"]), sorted_csc(f["Up"], f["Ui"], f["Ux"])
            if ref is None:
                ref = (num.p.copy(), L, U)
                continue
            assert np.array_equal(num.p, ref[0]) and same_pattern(L, ref[1]) and same_pattern(U, ref[2])
            dl, du = norm_diff(L, ref[1]), norm_diff(U, ref[2])
            print(f"{case} policy {p}: L {dl:.1e}, U {du:.1e} from policy 0")
            assert dl < 1e-10 and du < 1e-10
        rep = at.autotune(sym, Ap, Ai, Ax)
        print(f"{case}: {rep}")
        assert rep["best"] in at.POLICIES and rep["times"][rep["best"]] == min(rep["times"].values())
    finally:
        fplus_amd.tune("update_tile_policy", 0)
    with pytest.raises(fplus_amd.FplusError):
        fplus_amd.tune("update_tile_policy", 5)
