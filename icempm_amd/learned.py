"""The learned return map: an MLP (5 -> 64 -> 64 -> 3, tanh, BF16 on MFMA)
trained on the GPU (icempm_mlp_*), with device arrays for its data."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, byref, c_float, c_void_p

import numpy as np

from . import _lib
from ._lib import IcempmError

N_PARAMS = 6288
O_W1, O_B1, O_W2, O_B2, O_W3, O_B3 = 0, 1024, 1088, 5184, 5248, 6272


class DeviceArray:
    """A float32 array in device memory."""

    def __init__(self, host: np.ndarray):
        self._H = _lib.hip()
        a = np.ascontiguousarray(host, dtype=np.float32)
        self.shape = a.shape
        self.nbytes = a.nbytes
        p = c_void_p()
        if self._H.icempm_device_alloc(self.nbytes, byref(p)) != 0:
            raise IcempmError(1, "device_alloc failed")
        self.ptr = p
        if self._H.icempm_copy_to_device(self.ptr, a.ctypes.data, self.nbytes) != 0:
            raise IcempmError(1, "copy_to_device failed")

    @classmethod
    def empty(cls, shape) -> "DeviceArray":
        return cls(np.zeros(shape, dtype=np.float32))

    def host(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float32)
        if self._H.icempm_copy_to_host(out.ctypes.data, self.ptr, self.nbytes) != 0:
            raise IcempmError(1, "copy_to_host failed")
        return out

    def free(self) -> None:
        if getattr(self, "ptr", None):
            self._H.icempm_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def init_params(seed: int) -> np.ndarray:
    """Glorot-scaled random weights, zero padding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, dtype=np.float32)
    W1 = np.zeros((64, 16), dtype=np.float32)
    W1[:, :5] = rng.normal(0, np.sqrt(2 / (5 + 64)), (64, 5))
    p[O_W1:O_B1] = W1.ravel()
    p[O_W2:O_B2] = rng.normal(0, np.sqrt(2 / 128), 4096)
    W3 = np.zeros((16, 64), dtype=np.float32)
    W3[:3] = rng.normal(0, np.sqrt(2 / 67), (3, 64))
    p[O_W3:O_B3] = W3.ravel()
    return p


class MLP:
    def __init__(self, params: np.ndarray):
        self._H = _lib.hip()
        p = np.ascontiguousarray(params, dtype=np.float32)
        m = c_void_p()
        rc = self._H.icempm_mlp_create(byref(m), p.ctypes.data_as(POINTER(c_float)))
        if rc != 0:
            raise IcempmError(rc, "icempm_mlp_create failed")
        self._m = m

    def _check(self, rc):
        if rc != 0:
            raise IcempmError(rc, (self._H.icempm_mlp_last_error(self._m) or b"").decode())

    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, dtype=np.float32)
        self._check(self._H.icempm_mlp_get_params(self._m, out.ctypes.data_as(POINTER(c_float))))
        return out

    @property
    def device_params(self) -> int:
        return self._H.icempm_mlp_device_params(self._m)

    def step(self, X: DeviceArray, T: DeviceArray, r0: int, B: int, lr: float,
             want_grad: bool = False):
        g = np.empty(N_PARAMS, dtype=np.float32) if want_grad else None
        loss = c_float(0)
        self._check(self._H.icempm_mlp_step(self._m, X.ptr, T.ptr, r0, B, lr,
                                            g.ctypes.data_as(POINTER(c_float)) if want_grad else None,
                                            byref(loss)))
        return loss.value, g

    def infer(self, X: DeviceArray, n: int) -> np.ndarray:
        Y = DeviceArray.empty((n, 3))
        self._check(self._H.icempm_mlp_infer(self._m, X.ptr, n, Y.ptr))
        out = Y.host()
        Y.free()
        return out

    def close(self):
        if getattr(self, "_m", None):
            self._H.icempm_mlp_destroy(self._m)
            self._m = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mfma_probe(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """D = A (16 x K) B (K x 16) through the MFMA tile routine, K in {16, 32, 64}."""
    H = _lib.hip()
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    if A.shape[0] != 16 or B.shape[1] != 16 or A.shape[1] != B.shape[0]:
        raise ValueError("mfma_probe: A must be 16 x K and B K x 16")
    D = np.zeros((16, 16), dtype=np.float32)
    fp = POINTER(c_float)
    rc = H.icempm_mfma_probe_k(A.ctypes.data_as(fp), B.ctypes.data_as(fp), D.ctypes.data_as(fp),
                               int(A.shape[1]))
    if rc != 0:
        raise IcempmError(rc, "mfma_probe failed")
    return D


# -- training the return map on harvested pairs --------------------------------
def train_return_map(X: np.ndarray, T: np.ndarray, steps: int = 3000, batch: int = 8192,
                     lr: float = 3e-3, seed: int = 0, log_every: int = 0):
    """Train the MLP on harvested pairs with the library's training driver
    (icempm_mlp_train: normalisation, one seeded shuffle, Glorot initial
    weights, Adam on MFMA over consecutive windows of ``batch`` rows, cosine
    decay of the learning rate to lr/20).  Returns (params, norm, losses):
    losses holds (step, window loss) every ``log_every`` steps and
    (steps, loss over all rows) last."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    n = int(X.shape[0])
    if X.shape != (n, 5) or T.shape != (n, 3):
        raise ValueError("train_return_map: X must be n x 5 and T n x 3")
    H = _lib.hip()
    fp = POINTER(c_float)
    params = np.zeros(N_PARAMS, dtype=np.float32)
    norm = np.zeros(16, dtype=np.float32)
    every = int(log_every) if log_every else 0
    trace = np.zeros(max(1, (steps + every - 1) // every) if every else 1, dtype=np.float32)
    full = c_float(0)
    rc = H.icempm_mlp_train(X.ctypes.data_as(fp), T.ctypes.data_as(fp), n, int(steps), int(batch), float(lr),
                            int(seed) & 0xFFFFFFFFFFFFFFFF, params.ctypes.data_as(fp), norm.ctypes.data_as(fp),
                            trace.ctypes.data_as(fp) if every else None, every, byref(full))
    if rc != 0:
        raise IcempmError(rc, "icempm_mlp_train failed")
    losses = [(k * every, float(v)) for k, v in enumerate(trace)] if every else []
    losses.append((steps, float(full.value)))
    return params, norm, losses
