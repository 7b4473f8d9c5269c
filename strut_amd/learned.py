"""Learned Newton predictor: a 4 x 256 SiLU MLP trained on the device by the
library's own bf16 matrix-core kernels (FP32 accumulate, FP32 master
weights, Adam), on (features, target) pairs that the frame transient records
while it runs.

For a frame with n free DOFs the features of step k are the last three
displacement increments, v_k and a_k at every free DOF and the ground
acceleration at t_k and t_{k+1} (5 n + 2 values); the target is u_{k+1}
minus the Newmark predictor u_k + dt v_k + dt^2 (1/2 - beta) a_k
(transient_nonlinear.mojo:553-590).  With FrameBatch.set_predictor each step
starts from the network's proposal; the FP64 residual test
(transient_nonlinear.mojo:2357-2371) is unchanged, and the step restarts
from the Newmark predictor whenever the proposal raises the residual.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_float, c_int32, c_int64, c_uint64, c_void_p
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _native
from ._native import StrutGpuError

HIDDEN = 256
LAYERS = 5


def _declare(lib):
    if getattr(lib, "_mlp_declared", False):
        return
    fp = POINTER(c_float)
    sig = {
        "sgpu_mlp_create": (ctypes.c_int, [c_void_p, c_int32, c_int32, c_uint64, POINTER(c_void_p)]),
        "sgpu_mlp_destroy": (None, [c_void_p]),
        "sgpu_mlp_param_count": (c_int64, [c_void_p]),
        "sgpu_mlp_layout": (ctypes.c_int, [c_void_p, POINTER(c_int32)]),
        "sgpu_mlp_get_params": (ctypes.c_int, [c_void_p, fp]),
        "sgpu_mlp_set_params": (ctypes.c_int, [c_void_p, fp]),
        "sgpu_mlp_set_normalization": (ctypes.c_int, [c_void_p, fp, fp, fp, fp]),
        "sgpu_mlp_get_normalization": (ctypes.c_int, [c_void_p, fp]),
        "sgpu_mlp_fit_normalization": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64]),
        "sgpu_mlp_gradients": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int32, fp, POINTER(c_double)]),
        "sgpu_mlp_train": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_int32, c_float,
                                          c_uint64, POINTER(c_double)]),
        "sgpu_mlp_predict": (ctypes.c_int, [c_void_p, c_void_p, c_int64, c_void_p]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    lib._mlp_declared = True


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_float))


def _dev(t) -> c_void_p:
    """Device address of a contiguous float32 CUDA tensor."""
    if not t.is_cuda or t.dtype.itemsize != 4 or not t.is_contiguous():
        raise ValueError("expected a contiguous float32 device tensor")
    return c_void_p(t.data_ptr())


class Mlp:
    """One predictor network on its own device context."""

    def __init__(self, n_in: int, n_out: int, seed: int = 0, device: int = 0):
        self.lib = _native.load()
        _declare(self.lib)
        self.device = device
        self.ctx = c_void_p()
        rc = self.lib.sgpu_ctx_create(device, ctypes.byref(self.ctx))
        if rc != 0:
            raise StrutGpuError(rc, f"sgpu_ctx_create(device={device}) failed")
        self.h = c_void_p()
        self._check(self.lib.sgpu_mlp_create(self.ctx, n_in, n_out, seed, ctypes.byref(self.h)))
        self.n_in, self.n_out = n_in, n_out

    def _check(self, rc: int) -> int:
        if rc < 0:
            msg = self.lib.sgpu_last_error(self.ctx)
            raise StrutGpuError(rc, msg.decode() if msg else "")
        return rc

    def layout(self) -> Dict[str, int]:
        out = np.zeros(8, dtype=np.int32)
        self._check(self.lib.sgpu_mlp_layout(self.h, out.ctypes.data_as(POINTER(c_int32))))
        keys = ("n_in", "n_out", "k0", "n5", "hidden", "layers", "params", "adam_steps")
        return {k: int(v) for k, v in zip(keys, out)}

    def params(self) -> np.ndarray:
        p = np.zeros(int(self.lib.sgpu_mlp_param_count(self.h)), dtype=np.float32)
        self._check(self.lib.sgpu_mlp_get_params(self.h, _fptr(p)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != int(self.lib.sgpu_mlp_param_count(self.h)):
            raise ValueError("parameter count mismatch")
        self._check(self.lib.sgpu_mlp_set_params(self.h, _fptr(p)))

    def layers(self, p: Optional[np.ndarray] = None) -> List[tuple]:
        """(W [N][K], b [N]) per layer from the flat (padded) parameters."""
        lay = self.layout()
        p = self.params() if p is None else p
        out, off = [], 0
        for l in range(LAYERS):
            K = lay["k0"] if l == 0 else HIDDEN
            N = lay["n5"] if l == LAYERS - 1 else HIDDEN
            W = p[off:off + N * K].reshape(N, K)
            off += N * K
            b = p[off:off + N]
            off += N
            out.append((W, b))
        return out

    def normalization(self) -> Dict[str, np.ndarray]:
        out = np.zeros(2 * self.n_in + 2 * self.n_out, dtype=np.float32)
        self._check(self.lib.sgpu_mlp_get_normalization(self.h, _fptr(out)))
        a, b = self.n_in, 2 * self.n_in
        return {"x_mean": out[:a], "x_scale": out[a:b], "y_mean": out[b:b + self.n_out],
                "y_scale": out[b + self.n_out:]}

    def set_normalization(self, x_mean, x_scale, y_mean, y_scale) -> None:
        arrs = [np.ascontiguousarray(v, dtype=np.float32) for v in (x_mean, x_scale, y_mean, y_scale)]
        self._check(self.lib.sgpu_mlp_set_normalization(self.h, *[_fptr(a) for a in arrs]))

    def fit_normalization(self, x, y) -> None:
        self._check(self.lib.sgpu_mlp_fit_normalization(self.h, _dev(x), _dev(y), x.shape[0]))

    def gradients(self, x, y):
        """(flat FP32 gradients, loss) of one forward/backward pass over the
        device samples x, y (no update)."""
        g = np.zeros(int(self.lib.sgpu_mlp_param_count(self.h)), dtype=np.float32)
        loss = c_double()
        self._check(self.lib.sgpu_mlp_gradients(self.h, _dev(x), _dev(y), x.shape[0], _fptr(g),
                                                ctypes.byref(loss)))
        return g, loss.value

    def train(self, x, y, batch: int = 512, epochs: int = 20, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        losses = np.zeros(epochs)
        self._check(self.lib.sgpu_mlp_train(self.h, _dev(x), _dev(y), x.shape[0], batch, epochs, lr, seed,
                                            losses.ctypes.data_as(POINTER(c_double))))
        return losses

    def predict(self, x):
        import torch
        y = torch.empty((x.shape[0], self.n_out), dtype=torch.float32, device=x.device)
        self._check(self.lib.sgpu_mlp_predict(self.h, _dev(x), x.shape[0], _dev(y)))
        return y

    def close(self) -> None:
        if self.h:
            self.lib.sgpu_mlp_destroy(self.h)
            self.h = c_void_p()
        if self.ctx:
            self.lib.sgpu_ctx_destroy(self.ctx)
            self.ctx = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def feature_count(n_free: int) -> int:
    return 5 * n_free + 2


def train_predictor(x, y, seed: int = 0, epochs: int = 30, batch: int = 512, lr: float = 1e-3,
                    device: int = 0) -> tuple:
    """Fit the normalisation on (x, y) and train a new network on them;
    returns (network, per-epoch losses)."""
    net = Mlp(x.shape[1], y.shape[1], seed=seed, device=device)
    net.fit_normalization(x, y)
    losses = net.train(x, y, batch=batch, epochs=epochs, lr=lr, seed=seed + 1)
    return net, losses
