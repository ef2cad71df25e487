"""The learned precision selector on the device: its training logs, training, inference and the mixed factorization.

- Logs (precision 3): per block 12 finite features, the FP32-vs-FP64 difference of its update products (non-zero: the
  FP32 products ran) within FP32 round-off of the products' size (|P32 - P64| <= 1e-4 |P64|), and factors equal to the
  FP64 factorization's (the FP64 products are the ones applied) to the 1e-10 bound of two runs of one matrix.
- Training on logs of seeded matrices of the shift-invert family with 17 orders of scaling: the device-trained MLP
  separates "FP32 safe" from not (labels by a threshold on the logged difference) on its training data and on a held-out
  seed, and its device predictions match a NumPy forward pass of the same weights to FP32 round-off.
- Inference inside the factorization (precision 2): threshold 0 marks every block FP32, threshold 1 none; a trained
  selector marks a share in between; each mixed refactorization with its refined solve reaches an equilibrated backward
  error below 1e-14; and refactor_solve falls back to FP64 updates when the refinement is given no room to converge.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import fplus_amd
from fplus_amd import accuracy as acc
from fplus_amd import gen
from helpers import norm_diff, sorted_csc

pytestmark = pytest.mark.gpu


def _family(seed, modes=4, ncoef=256):
    return gen.shift_invert(modes, 4, ncoef, seed=seed, scale_exp=8.0, scale_hi=9.0)


def _logs(seeds):
    F, D, V = [], [], []
    for s in seeds:
        n, Ap, Ai, Ax = _family(s)
        sym = fplus_amd.Symbolic(n, Ap, Ai)
        num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(precision=3))
        f, d, v = fplus_amd.log_blocks(num)
        F.append(f), D.append(d), V.append(v)
    return np.concatenate(F), np.concatenate(D), np.concatenate(V)


def _mlp_numpy(w, F):
    """the MLP forward pass (weights as fplus_amd_selector_get_weights lays them out), P(class 1)"""
    nf, h = 12, 64
    o = 0
    W1 = w[o:o + h * nf].reshape(h, nf); o += h * nf
    b1 = w[o:o + h]; o += h
    W2 = w[o:o + h * h].reshape(h, h); o += h * h
    b2 = w[o:o + h]; o += h
    W3 = w[o:o + 2 * h].reshape(2, h); o += 2 * h
    b3 = w[o:o + 2]; o += 2
    mu = w[o:o + nf]; o += nf
    sd = w[o:o + nf]
    x = (F.astype(np.float64) - mu) / sd
    a1 = np.maximum(x @ W1.T + b1, 0)
    a2 = np.maximum(a1 @ W2.T + b2, 0)
    z = a2 @ W3.T + b3
    z -= z.max(1, keepdims=True)
    e = np.exp(z)
    return e[:, 1] / e.sum(1)


def test_logs_and_logged_factorization_is_fp64():
    n, Ap, Ai, Ax = _family(0)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num3 = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(precision=3))
    num0 = fplus_amd.Numeric(sym, Ap, Ai, Ax)
    f, d, v = fplus_amd.log_blocks(num3)
    assert f.shape == (fplus_amd.stats(num3)["nblocks"], 12) and np.isfinite(f).all()
    upd = v > 0
    print(f"blocks {len(d)}, updating {upd.sum()}, diff median {np.median(d[upd]):.2e}, max ratio {(d[upd] / v[upd]).max():.2e}")
    assert upd.sum() > 0.5 * len(d)
    assert (d >= 0).all() and (d[upd] > 0).mean() > 0.9
    assert (d[upd] <= 1e-4 * v[upd]).all()
    assert np.array_equal(num3.p, num0.p)
    f3, f0 = num3.factors(), num0.factors()
    L3, L0 = sorted_csc(f3["Lp"], f3["Li"], f3["Lx"]), sorted_csc(f0["Lp"], f0["Li"], f0["Lx"])
    U3, U0 = sorted_csc(f3["Up"], f3["Ui"], f3["Ux"]), sorted_csc(f0["Up"], f0["Ui"], f0["Ux"])
    assert norm_diff(L3, L0) < 1e-10 and norm_diff(U3, U0) < 1e-10
    with pytest.raises(fplus_amd.FplusError):
        fplus_amd.log_blocks(num0)     # no logs without precision 3


def test_train_predict_and_mixed_factorization():
    F, D, V = _logs(range(4))
    Fh, Dh, _ = _logs([100])
    tau = float(np.quantile(D, 0.5))
    y = (D <= tau).astype(np.float32)
    sel = fplus_amd.Selector(seed=3)
    loss = sel.train(F, y, epochs=60, lr=0.05, seed=1)
    p = sel.predict(F)
    ph = sel.predict(Fh)
    acc_tr = ((p > 0.5) == (y > 0.5)).mean()
    acc_ho = ((ph > 0.5) == (Dh <= tau)).mean()
    print(f"samples {len(y)}, loss {loss:.3f}, train accuracy {acc_tr:.3f}, held-out accuracy {acc_ho:.3f}")
    assert np.isfinite(loss) and loss < 0.6
    assert acc_tr > 0.75 and acc_ho > 0.7
    # the device forward pass against NumPy's on the trained weights
    assert np.allclose(p, _mlp_numpy(sel.weights.astype(np.float64), F), atol=1e-4)
    # weights round trip
    sel2 = fplus_amd.Selector(seed=9, weights=sel.weights)
    assert np.array_equal(sel2.predict(F), p)

    n, Ap, Ai, Ax = _family(100)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 4, True)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    shares = {}
    for thr in (0.0, 0.5, 1.0):
        sel.threshold = thr
        num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(precision=2, refactor=1, selector=sel))
        st = fplus_amd.stats(num)
        shares[thr] = st["f32_flop_share"]
        x, info = num.solve_refined(b)
        m = acc.all_measures(A, x, b, num.Dr, num.Dc)
        # the same selection replayed in a refactorization
        x2, info2 = num.refactor_solve(Ax, b, precision=2)
        m2 = acc.all_measures(A, x2, b, num.Dr, num.Dc)
        st2 = fplus_amd.stats(num)
        print(f"threshold {thr}: FP32 blocks {st['f32_blocks']} of {st['nblocks']}, flop share {st['f32_flop_share']:.3f}; "
              f"refined {info} {m}; refactor_solve {info2} {m2}")
        assert info["converged"] == 1 and m["equilibrated"] < 1e-14
        assert info2["converged"] == 1 and info2["fallback"] == 0 and m2["equilibrated"] < 1e-14
        assert st2["f32_blocks"] == st["f32_blocks"] or thr == 0.5
    assert shares[0.0] == 1.0 and shares[1.0] == 0.0
    assert 0.0 < shares[0.5] < 1.0


def test_refactor_solve_falls_back_to_fp64():
    n, Ap, Ai, Ax = _family(7)
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    b = gen.rhs(n, 2, True)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(refactor=1))
    # FP32 updates with no refinement step allowed: the first solve misses 1e-14, so the FP64 refactorization runs
    x, info = num.refactor_solve(Ax, b, precision=1, max_outer=0)
    m = acc.all_measures(A, x, b, num.Dr, num.Dc)
    print(f"fallback: {info} {m}")
    assert info["fallback"] == 1 and info["bwd_first"] > 1e-14
    assert m["equilibrated"] < 1e-14
    assert fplus_amd.stats(num)["precision"] == 0
