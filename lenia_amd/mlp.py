"""Python host for the parameter-proposal MLP (include/lenia_mlp.h): ctypes over
liblenia_hip.so.  Training, inference and the GPU parameter sampler all run in the
library's HIP kernels; this module only moves arrays and drives the loop."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

IN, INP, HID, OUT, OUTP = 81, 128, 256, 3, 64
SHAPES = [("W1", (INP, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
          ("W3", (HID, HID)), ("b3", (HID,)), ("W4", (HID, OUTP)), ("b4", (OUTP,))]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed: int) -> np.ndarray:
    """He-normal weights on the real fan-in, zero biases, zero padding."""
    rng = np.random.default_rng(seed)
    parts = []
    for name, shape in SHAPES:
        a = np.zeros(shape, np.float32)
        if name == "W1":
            a[:IN] = rng.normal(0.0, np.sqrt(2.0 / IN), (IN, HID))
        elif name == "W4":
            a[:, :OUT] = rng.normal(0.0, np.sqrt(1.0 / HID), (HID, OUT))
        elif name.startswith("W"):
            a[:] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        parts.append(a.reshape(-1))
    return np.concatenate(parts).astype(np.float32)


class MLP:
    def __init__(self, max_batch: int = 4096, device: int = 0):
        self._L = L.load()
        self.count = int(self._L.lm_param_count())
        h = C.c_void_p()
        L.check(self._L.lm_create(device, int(max_batch), C.byref(h)), "lm_create")
        self._h = h
        self.max_batch = max_batch

    def close(self):
        if self._h:
            self._L.lm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        L.check(self._L.lm_set_params(self._h, _fp(p), p.size), "lm_set_params")

    def params(self) -> np.ndarray:
        p = np.empty(self.count, np.float32)
        L.check(self._L.lm_get_params(self._h, _fp(p), p.size), "lm_get_params")
        return p

    def set_data(self, x: np.ndarray, y: np.ndarray):
        self._x = np.ascontiguousarray(x, dtype=np.float32)
        self._y = np.ascontiguousarray(y, dtype=np.float32)
        L.check(self._L.lm_set_data(self._h, _fp(self._x), _fp(self._y), self._x.shape[0]), "lm_set_data")
        self.n = self._x.shape[0]

    def train(self, epochs: int, batch: int, lr: float, seed: int = 0) -> list:
        """Shuffled epochs over the resident set; returns the mean loss per epoch."""
        rng = np.random.default_rng(seed)
        nb = self.n // batch
        hist = []
        for _ in range(epochs):
            order = np.ascontiguousarray(rng.permutation(self.n)[:nb * batch].astype(np.int32))
            loss = C.c_float()
            L.check(self._L.lm_train(self._h, order.ctypes.data_as(C.POINTER(C.c_int)), nb, batch, float(lr),
                                     C.byref(loss)), "lm_train")
            hist.append(loss.value)
        return hist

    def gradients(self, x: np.ndarray, y: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.ascontiguousarray(y, dtype=np.float32)
        g = np.empty(self.count, np.float32)
        loss = C.c_float()
        L.check(self._L.lm_gradients(self._h, _fp(x), _fp(y), x.shape[0], _fp(g), C.byref(loss)), "lm_gradients")
        return loss.value, g

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        out = np.empty((x.shape[0], OUT), np.float32)
        L.check(self._L.lm_forward(self._h, _fp(x), x.shape[0], _fp(out)), "lm_forward")
        return out

    def score_sampled(self, seed: int, n: int) -> np.ndarray:
        out = np.empty(n, np.float32)
        L.check(self._L.lm_score_sampled(self._h, C.c_uint64(seed), C.c_longlong(n), _fp(out)), "lm_score_sampled")
        return out

    def sample_features(self, seed: int, first: int, n: int) -> np.ndarray:
        out = np.empty((n, IN), np.float32)
        L.check(self._L.lm_sample_features(self._h, C.c_uint64(seed), C.c_longlong(first), n, _fp(out)),
                "lm_sample_features")
        return out
