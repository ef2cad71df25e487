"""The learned per-block precision selector (fplus_amd_selector_*, include/fplus_amd.h).

A block is a supernode's columns within one panel; its updates of later columns are products with its inverse form
[L11^-1; L21 L11^-1].  A factorization with ``amd_options(precision=3)`` computes every such product in FP64 (applied)
and in FP32 beside it and logs, per block, the relative difference of the two and 12 features of the block
(``log_blocks``).  ``Selector`` is an MLP 12 -> 64 -> 64 -> 2 in FP32, trained on those logs with labels "the FP32
products stay within tau" and evaluated on the device as each block is made, in a factorization or refactorization with
``precision=2``: the blocks it marks update in FP32, the rest in FP64.  A refined solve (GMRES-IR) wins back what the
FP32 updates lose; ``Numeric.refactor_solve`` falls back to FP64 updates when it does not converge.
"""
import ctypes as C

import numpy as np

from ._lib import FplusError, check, lib

NFEAT = 12
# the features k_features (selector_dev.hip) computes per block, in order
FEATURE_NAMES = ("log2_width", "log2_rows_below_plus_1", "log2_rows", "log10_max_inverse_form", "log10_max_L21",
                 "log10_pivot_spread", "log10_min_pivot_over_colmax", "log10_col_scale_spread", "log10_row_scale_spread",
                 "log10_max_pivot", "position", "log10_max_L11inv")

_FP = C.POINTER(C.c_float)


def _fp(a):
    return a.ctypes.data_as(_FP)


class Selector:
    """fplus_amd_selector: create (seeded weights), train, predict, weights, threshold"""

    def __init__(self, seed=0, weights=None, threshold=0.5):
        self.ptr = C.c_void_p()
        check(lib().fplus_amd_selector_create(int(seed), C.byref(self.ptr)), None)
        if weights is not None:
            self.weights = weights
        self.threshold = threshold

    @property
    def nweights(self):
        return 5122 + 2 * NFEAT

    @property
    def weights(self):
        w = np.empty(self.nweights, np.float32)
        check(lib().fplus_amd_selector_get_weights(self.ptr, _fp(w), w.size), None)
        return w

    @weights.setter
    def weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        check(lib().fplus_amd_selector_set_weights(self.ptr, _fp(w), w.size), None)

    @property
    def threshold(self):
        return self._threshold

    @threshold.setter
    def threshold(self, t):
        check(lib().fplus_amd_selector_set_threshold(self.ptr, float(t)), None)
        self._threshold = float(t)

    def train(self, features, labels, epochs=60, lr=0.05, seed=0):
        """SGD with momentum on (features, labels 0/1) on the device; returns the last epoch's mean cross-entropy"""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, NFEAT)
        y = np.ascontiguousarray(labels, np.float32).reshape(-1)
        if len(y) != len(f):
            raise ValueError("one label per feature row")
        loss = C.c_float(0.0)
        check(lib().fplus_amd_selector_train(self.ptr, _fp(f), _fp(y), len(y), int(epochs), float(lr), int(seed),
                                             C.byref(loss)), None)
        return float(loss.value)

    def predict(self, features):
        """P(FP32 safe) per feature row"""
        f = np.ascontiguousarray(features, np.float32).reshape(-1, NFEAT)
        p = np.empty(len(f), np.float32)
        check(lib().fplus_amd_selector_predict(self.ptr, _fp(f), len(f), _fp(p)), None)
        return p

    def save(self, path):
        np.savez(path, weights=self.weights, threshold=np.float32(self.threshold))

    @classmethod
    def load(cls, path):
        with np.load(path, allow_pickle=False) as z:
            return cls(weights=z["weights"], threshold=float(z["threshold"]))

    def __del__(self):
        if getattr(self, "ptr", None) and lib is not None:
            lib().fplus_amd_selector_free(self.ptr)
            self.ptr = None


def log_blocks(num):
    """(features [nblocks, 12], diff [nblocks], pmax [nblocks]) of a factorization or refactorization made with precision
    3: per block the largest |P32 - P64| over its update products and the largest |P64|, in the equilibrated scale"""
    nb = C.c_int64(0)
    rc = lib().fplus_amd_selector_log(num.ptr, None, None, None, 0, C.byref(nb))
    if rc != 0 and nb.value == 0:
        raise FplusError(rc, "no blocks")
    f = np.empty((nb.value, NFEAT), np.float32)
    d = np.empty(nb.value, np.float64)
    v = np.empty(nb.value, np.float64)
    check(lib().fplus_amd_selector_log(num.ptr, _fp(f), d.ctypes.data_as(C.POINTER(C.c_double)),
                                       v.ctypes.data_as(C.POINTER(C.c_double)), nb.value, C.byref(nb)), None)
    return f, d, v
