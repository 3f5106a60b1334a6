"""Learned error correction for the Poisson band (host side of the
msbg_learned_* entry points).  No reference counterpart: the reference has no
learned component.

The network maps a 16^3 block with a 2-cell halo (scaled residual, phase
field; 20^3 x 2 bf16, channels-last) to a 16^3 error estimate through four
3^3 convolutions (2 -> 16 -> 16 -> 16 -> 1, ReLU between), run and trained on
the GPU as BF16 MFMA implicit GEMMs with FP32 accumulation.
"""
import ctypes as C
import os

import numpy as np

from . import lib as _lib

LAYER_PARAMS = 27 * 16 * 16 + 16
CIN = (2, 16, 16, 16)
COUT = (16, 16, 16, 1)


def init_params(seed=0):
    """He-normal weights (fan-in 27 * c_in), zero biases, padding zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(4 * LAYER_PARAMS, np.float32)
    for l in range(4):
        w = np.zeros((27, 16, 16), np.float32)
        std = np.sqrt(2.0 / (27 * CIN[l])) * (0.5 if l == 3 else 1.0)
        w[:, : CIN[l], : COUT[l]] = rng.normal(scale=std, size=(27, CIN[l], COUT[l]))
        p[l * LAYER_PARAMS: l * LAYER_PARAMS + 27 * 256] = w.ravel()
    return p


class CorrectorNet:
    def __init__(self, lib=None, params=None, out_scale=1.0):
        self.lib = lib or _lib.load()
        h = C.c_void_p()
        self.lib.call("msbg_learned_new", C.byref(h))
        self.h = h.value
        n = C.c_int()
        self.lib.call("msbg_learned_param_count", C.byref(n))
        self.n_params = n.value
        if params is not None:
            self.set_params(params, out_scale)

    def set_params(self, params, out_scale=1.0):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.n_params
        self.lib.call("msbg_learned_set_params", self.h, p.ctypes.data, float(out_scale))

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        s = C.c_float()
        self.lib.call("msbg_learned_get_params", self.h, p.ctypes.data, C.byref(s))
        return p, s.value

    def forward(self, x0, n):
        """x0: device bf16 bits (n x 20^3 x 2) -> device float32 (n x 4096)."""
        y = self.lib.empty((n, 4096), np.float32)
        self.lib.call("msbg_learned_forward", self.h, x0.ptr, n, y.ptr)
        return y

    def loss_grad(self, x0, target, n):
        loss = C.c_double()
        g = np.zeros(self.n_params, np.float32)
        self.lib.call("msbg_learned_loss_grad", self.h, x0.ptr, target.ptr, n, C.byref(loss), g.ctypes.data)
        return loss.value, g

    def train(self, x0, target, n_pairs, steps, batch=64, lr=1e-3, seed=0):
        """Adam on batches drawn from the n_pairs device pairs; returns the
        per-step batch losses (before each step)."""
        losses = np.zeros(steps, np.float32)
        self.lib.call("msbg_learned_train", self.h, x0.ptr, target.ptr, n_pairs, steps, batch, lr, seed,
                      losses.ctypes.data)
        return losses

    def save(self, path):
        p, s = self.get_params()
        np.savez(path, params=p, out_scale=np.float32(s))

    @classmethod
    def load(cls, path, lib=None):
        with np.load(path, allow_pickle=False) as f:
            return cls(lib, f["params"], float(f["out_scale"]))

    def close(self):
        if self.h:
            self.lib.call("msbg_learned_free", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "corrector.npz")
