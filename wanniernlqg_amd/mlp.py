"""ctypes binding of the learned refinement proposal network (``wnlqg_mlp_*`` in include/wnlqg.h):
an 8 -> 64 -> 64 -> 1 SiLU perceptron trained and evaluated by the library's own HIP kernels.
No CPU fallback: the library must be built."""

from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from .engine import load_library

PARAMS = 4801
HIDDEN = 64
INPUTS = 8


class MlpError(RuntimeError):
    pass


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ProposalNetwork:
    """Device-resident MLP (FP32)."""

    def __init__(self, seed: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        self._check(self.lib.wnlqg_mlp_create(C.c_uint64(seed), C.byref(h)))
        self.handle = h

    def _check(self, rc: int):
        if rc != 0:
            buf = C.create_string_buffer(1024)
            self.lib.wnlqg_mlp_last_error(buf, 1024)
            raise MlpError(f"wnlqg_mlp error {rc}: {buf.value.decode(errors='replace')}")

    def close(self):
        if getattr(self, "handle", None):
            self.lib.wnlqg_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(PARAMS, dtype=np.float32)
        self._check(self.lib.wnlqg_mlp_get_params(self.handle, _fp(p)))
        return p

    @params.setter
    def params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (PARAMS,)
        self._check(self.lib.wnlqg_mlp_set_params(self.handle, _fp(p)))

    def predict(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, INPUTS)
        y = np.zeros(x.shape[0], dtype=np.float32)
        self._check(self.lib.wnlqg_mlp_predict(self.handle, _fp(x), x.shape[0], _fp(y)))
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, INPUTS)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        g = np.zeros(PARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(self.lib.wnlqg_mlp_gradient(self.handle, _fp(x), _fp(t), x.shape[0], _fp(g), C.byref(loss)))
        return g, float(loss.value)

    def gradient_noalloc(self, x_ptr: int, t_ptr: int, n: int, grad_ptr: int, stream: int = 0):
        """wnlqg_mlp_gradient_noalloc on device buffers (addresses): x [n][INPUTS], t [n], grad [PARAMS + 1] floats;
        grad receives the gradient of the mean squared error and then the sum of squared errors.  stream 0:
        the network's own stream, synchronized before returning."""
        self._check(self.lib.wnlqg_mlp_gradient_noalloc(self.handle, C.c_void_p(x_ptr), C.c_void_p(t_ptr), int(n),
                                                        C.c_void_p(grad_ptr), C.c_void_p(stream or None)))

    def train(self, x: np.ndarray, t: np.ndarray, steps: int, batch: int = 4096, lr: float = 3e-3,
              seed: int = 0, history: bool = True) -> Optional[np.ndarray]:
        x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, INPUTS)
        t = np.ascontiguousarray(t, dtype=np.float32).reshape(-1)
        h = np.zeros(steps, dtype=np.float32) if history else None
        self._check(self.lib.wnlqg_mlp_train(self.handle, _fp(x), _fp(t), x.shape[0], steps, batch, C.c_float(lr),
                                             C.c_uint64(seed), _fp(h) if history else C.POINTER(C.c_float)()))
        return h
