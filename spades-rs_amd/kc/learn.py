"""Learned first pass of the solid-k-mer selection (no counterpart in the reference).

Mlp wraps the device network of libkc.so (include/kc.h, kc_mlp_*): a 16 -> 64 -> 64 -> 1
MLP with bf16 weights and fp32 accumulation, trained on the GPU by its own forward and
backward kernels.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import KcError, f32p, lib, ptr, u8p

N_FEATURES = 16


class Mlp:
    def __init__(self, seed=1, device=0):
        h = ctypes.c_void_p()
        st = lib().kc_mlp_create(ctypes.byref(h), device, int(seed))
        if st != _abi.KC_OK:
            raise KcError(st, "kc_mlp_create failed")
        self._h = h
        self.n_params = lib().kc_mlp_num_params()

    def close(self):
        if getattr(self, "_h", None):
            lib().kc_mlp_destroy(self._h)
            self._h = None

    __del__ = close

    def _chk(self, st):
        if st != _abi.KC_OK:
            raise KcError(st, lib().kc_mlp_last_error(self._h).decode())

    @staticmethod
    def _x(X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2 or X.shape[1] != N_FEATURES:
            raise ValueError(f"features must be (n, {N_FEATURES})")
        return X

    @property
    def params(self):
        w = np.zeros(self.n_params, dtype=np.float32)
        self._chk(lib().kc_mlp_get_params(self._h, ptr(w, f32p), w.shape[0]))
        return w

    @params.setter
    def params(self, w):
        w = np.ascontiguousarray(w, dtype=np.float32)
        self._chk(lib().kc_mlp_set_params(self._h, ptr(w, f32p), w.shape[0]))

    def predict(self, X):
        X = self._x(X)
        p = np.zeros(X.shape[0], dtype=np.float32)
        if X.shape[0]:
            self._chk(lib().kc_mlp_forward(self._h, ptr(X, f32p), X.shape[0], ptr(p, f32p)))
        return p

    def gradient(self, X, y, sample_w=None, pos_weight=1.0):
        X = self._x(X)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        sw = None if sample_w is None else np.ascontiguousarray(sample_w, dtype=np.float32)
        g = np.zeros(self.n_params, dtype=np.float32)
        loss = ctypes.c_double(0)
        self._chk(lib().kc_mlp_gradient(self._h, ptr(X, f32p), ptr(y, u8p), ptr(sw, f32p), X.shape[0],
                                        float(pos_weight), ptr(g, f32p), ctypes.byref(loss)))
        return loss.value, g

    def train(self, X, y, sample_w=None, epochs=4, batch=4096, lr=2e-3, pos_weight=1.0, seed=1):
        X = self._x(X)
        y = np.ascontiguousarray(y, dtype=np.uint8)
        sw = None if sample_w is None else np.ascontiguousarray(sample_w, dtype=np.float32)
        loss = ctypes.c_double(0)
        self._chk(lib().kc_mlp_train(self._h, ptr(X, f32p), ptr(y, u8p), ptr(sw, f32p), X.shape[0], int(epochs),
                                     int(batch), float(lr), float(pos_weight), int(seed), ctypes.byref(loss)))
        return loss.value


def count_histogram(counts, n_bins=1024):
    """hist[c] = entries with count c for c < n_bins - 1, hist[n_bins - 1] those with more."""
    c = np.minimum(np.asarray(counts, dtype=np.int64), n_bins - 1)
    return np.bincount(c, minlength=n_bins).astype(np.uint64)


def cutoff_is_fragile(counts, min_coverage=5.0, resamples=32, seed=0):
    """The library-level margin (kc_cutoff_is_fragile in libkc.so) over a count vector:
    True when the cutoff's valley search (assemble.rs:692-735) is too fragile for a
    learned selection.  SpectrumEngine.cutoff_fragile gives the same for a counted
    library without copying its counts."""
    counts = np.asarray(counts)
    hist = count_histogram(counts)
    f = ctypes.c_int32(0)
    st = lib().kc_cutoff_is_fragile(ptr(hist, _abi.u64p), hist.shape[0], counts.shape[0], float(min_coverage),
                                    int(resamples), int(seed), ctypes.byref(f))
    if st != _abi.KC_OK:
        raise KcError(st, "kc_cutoff_is_fragile failed")
    return bool(f.value)
