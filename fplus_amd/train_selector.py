"""Train the learned precision selector on logs of the shift-invert family with 17 orders of scaling.

    python -m fplus_amd.train_selector [--matrices 200] [--tau 1e-7] [--out fplus_amd/data/selector_shiftinvert.npz]

Each training matrix is a member of the shift-invert pencil family (4 fields x 1024 coefficients per mode, row and
column scales 10^U(-8, 9)) with its own seed and a mode count drawn so that n is log-uniform in [8192, 65536].  Each
is factored once with precision 3 (a member whose pivot fails the threshold test is skipped and the next seed drawn): every block's update products run in FP64 (applied) and in FP32 beside them, and
the factorization logs per block its 12 features and the largest |P32 - P64| (in the equilibrated scale).  A block is
labelled FP32-safe when that difference is at most tau.  The MLP (12 -> 64 -> 64 -> 2, FP32) is trained on the device
on all blocks of the training matrices and scored on held-out seeds.  The decision threshold is then chosen on two more held-out members at the largest training
size by the time of a refactorization + refined solve (calibrate_threshold).  The weights and the threshold go to
--out, the scores to stdout as one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

import fplus_amd
from fplus_amd import gen

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "data", "selector_shiftinvert.npz")


def family_member(seed, modes):
    return gen.shift_invert(modes, 4, 1024, seed=seed, scale_exp=8.0, scale_hi=9.0)


def draw_modes(rng):
    """modes so that n = 4096 modes is log-uniform in [8192, 65536]"""
    return int(np.clip(np.rint(2.0 ** rng.uniform(1.0, 4.0)), 2, 16))


def log_matrix(seed, modes):
    n, Ap, Ai, Ax = family_member(seed, modes)
    sym = fplus_amd.Symbolic(n, Ap, Ai)
    num = fplus_amd.Numeric(sym, Ap, Ai, Ax, amd_options=fplus_amd.amd_options(precision=3))
    f, d, v = fplus_amd.log_blocks(num)
    return n, f, d, v


def calibrate_threshold(sel, thresholds=(0.5, 0.9, 0.99, 0.999, 1.0), seeds=(910000, 910001), modes=16, nrhs=16, tol=1e-13):
    """the decision threshold by the time of a refactorization + refined solve (tol, nrhs right-hand sides) on held-out
    members of the family at the largest training size: for each seed, a pivoting factorization at a neighbouring
    shift is replayed at the member's own values under each threshold; the threshold with the least total time wins
    (the lowest on ties).  A threshold of 1.0 marks no block (a probability never exceeds it): every update in FP64,
    the refined solve then usually done at once.  Returns (best, {threshold: seconds})."""
    tot = {t: 0.0 for t in thresholds}
    for seed in seeds:
        n, Ap, Ai, Ax = gen.shift_invert(modes, 4, 1024, seed=seed, scale_exp=8.0, scale_hi=9.0)
        _, _, _, Axp = gen.shift_invert(modes, 4, 1024, seed=seed, scale_exp=8.0, scale_hi=9.0, sigma=0.3 + 0.69j)
        sym = fplus_amd.Symbolic(n, Ap, Ai)
        try:
            num = fplus_amd.Numeric(sym, Ap, Ai, Axp, amd_options=fplus_amd.amd_options(refactor=1, selector=sel))
        except fplus_amd.FplusError:
            continue
        b = gen.rhs(n, nrhs, True, seed=seed)
        num.refactor_solve(Ax, b, precision=0, tol=tol, inner_tol=1e-2)   # warm the buffers
        for t in thresholds:
            sel.threshold = t
            t0 = time.perf_counter()
            num.refactor_solve(Ax, b, precision=2, tol=tol, inner_tol=1e-2)
            tot[t] += time.perf_counter() - t0
        del num
    best = min(thresholds, key=lambda t: (round(tot[t], 3), t))
    sel.threshold = best
    return best, tot


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--matrices", type=int, default=200)
    ap.add_argument("--heldout", type=int, default=8)
    ap.add_argument("--tau", type=float, default=1e-7)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--lr", type=float, default=0.05)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--calibrate-only", action="store_true",
                    help="keep the weights in --out and only choose the decision threshold (calibrate_threshold)")
    args = ap.parse_args(argv)
    if args.calibrate_only:
        sel = fplus_amd.Selector.load(args.out)
        best, tot = calibrate_threshold(sel)
        sel.save(args.out)
        print(json.dumps({"threshold": best, "refactor_solve_s": {str(k): v for k, v in tot.items()}}))
        return
    rng = np.random.default_rng(args.seed)
    t0 = time.perf_counter()
    F, D, sizes = [], [], []
    seed, skipped = 1000, []
    while len(D) < args.matrices:
        try:
            n, f, d, _ = log_matrix(seed, draw_modes(rng))
        except fplus_amd.FplusError as e:   # a member whose pivot fails the threshold: not a training sample
            skipped.append(seed)
            print(f"seed {seed} skipped: {e}", file=sys.stderr, flush=True)
            seed += 1
            continue
        seed += 1
        F.append(f), D.append(d), sizes.append(n)
        if len(D) % 10 == 0:
            print(f"logged {len(D)}/{args.matrices} matrices, {sum(len(x) for x in D)} blocks, {time.perf_counter() - t0:.0f} s",
                  file=sys.stderr, flush=True)
    F, D = np.concatenate(F), np.concatenate(D)
    y = (D <= args.tau).astype(np.float32)
    t1 = time.perf_counter()
    sel = fplus_amd.Selector(seed=args.seed)
    loss = sel.train(F, y, epochs=args.epochs, lr=args.lr, seed=args.seed)
    t2 = time.perf_counter()
    acc_tr = float(((sel.predict(F) > 0.5) == (y > 0.5)).mean())
    Fh, Dh = [], []
    seed = 900000
    while len(Dh) < args.heldout:
        try:
            _, f, d, _ = log_matrix(seed, draw_modes(rng))
            Fh.append(f), Dh.append(d)
        except fplus_amd.FplusError:
            skipped.append(seed)
        seed += 1
    Fh, Dh = np.concatenate(Fh), np.concatenate(Dh)
    yh = Dh <= args.tau
    ph = sel.predict(Fh) > 0.5
    rep = {"matrices": args.matrices, "skipped_singular": skipped, "n_range": [min(sizes), max(sizes)], "blocks": int(len(y)), "tau": args.tau,
           "label_fp32_share": float(y.mean()), "loss": loss, "train_accuracy": acc_tr,
           "heldout_matrices": args.heldout, "heldout_blocks": int(len(yh)), "heldout_accuracy": float((ph == yh).mean()),
           "heldout_false_fp32": float((ph & ~yh).mean()), "log_s": t1 - t0, "train_s": t2 - t1}
    best, tot = calibrate_threshold(sel)
    rep["threshold"] = best
    rep["calibration_refactor_solve_s"] = {str(k): v for k, v in tot.items()}
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    sel.save(args.out)
    print(json.dumps(rep))


if __name__ == "__main__":
    main()
