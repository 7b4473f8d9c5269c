"""CPU oracle of the learned predictor's network (test infrastructure only).

Restates, in numpy float64, the forward and backward pass that
strut_amd/csrc/predictor.hip runs on the matrix cores: five linear layers
(4 x 256 SiLU hidden, linear output), mean squared error of the normalised
outputs.  With ``emulate=True`` every value is rounded where the device
rounds it: features, hidden activations and pre-activation gradients to
bf16 (round to nearest even), products summed exactly and rounded to FP32,
biases added in FP32.  With ``emulate=False`` everything stays float64, which
the CPU tests pin against finite differences and against torch autograd.

The predictor has no counterpart in the reference (its Newmark predictor,
solver/run_case/analysis/transient_nonlinear.mojo:553-590, is the plain
method); the features and target follow transient_nonlinear.mojo:496-590.
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np

HIDDEN = 256
LAYERS = 5


def bf16(a: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def f32(a) -> np.ndarray:
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def silu(z: np.ndarray) -> np.ndarray:
    return z / (1.0 + np.exp(-z))


def silu_grad(z: np.ndarray) -> np.ndarray:
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unflatten(p: np.ndarray, k0: int, n5: int) -> List[Tuple[np.ndarray, np.ndarray]]:
    """(W [N][K], b [N]) per layer from the flat parameter layout of
    sgpu_mlp_get_params: per layer W row-major then b."""
    out, off = [], 0
    for l in range(LAYERS):
        K = k0 if l == 0 else HIDDEN
        N = n5 if l == LAYERS - 1 else HIDDEN
        W = np.asarray(p[off:off + N * K], dtype=np.float64).reshape(N, K)
        off += N * K
        b = np.asarray(p[off:off + N], dtype=np.float64)
        off += N
        out.append((W, b))
    return out


def flatten(grads: List[Tuple[np.ndarray, np.ndarray]]) -> np.ndarray:
    return np.concatenate([np.concatenate([g.reshape(-1), gb]) for g, gb in grads])


def normalise(x, y, norm, k0: int, n5: int, emulate: bool):
    """Padded, normalised features [rows][k0] and targets [rows][n5]."""
    rows, n_in = x.shape
    n_out = y.shape[1]
    if emulate:
        xn = f32(f32(f32(x) - f32(norm["x_mean"])) * f32(norm["x_scale"]))
        tn = f32(f32(f32(y) - f32(norm["y_mean"])) / f32(norm["y_scale"]))
        xn = bf16(xn)
    else:
        xn = (x - norm["x_mean"]) * norm["x_scale"]
        tn = (y - norm["y_mean"]) / norm["y_scale"]
    X = np.zeros((rows, k0))
    X[:, :n_in] = xn
    T = np.zeros((rows, n5))
    T[:, :n_out] = tn
    return X, T


def forward_backward(layers, X: np.ndarray, T: np.ndarray, n_out: int, emulate: bool):
    """Loss, de-normalisation-free outputs and per-layer (dW, db) of the mean
    squared error over the first n_out outputs."""
    rnd = bf16 if emulate else (lambda a: a)
    acc = f32 if emulate else (lambda a: a)
    Ws = [rnd(W) for W, _ in layers]
    bs = [acc(b) for _, b in layers]
    rows = X.shape[0]
    H = [X]
    Z = []
    a = X
    for l in range(LAYERS - 1):
        z = acc(acc(a @ Ws[l].T) + bs[l])
        Z.append(z)
        a = rnd(acc(silu(z)))
        H.append(a)
    y = acc(acc(a @ Ws[-1].T) + bs[-1])
    e = np.zeros_like(y)
    e[:, :n_out] = acc(y[:, :n_out] - T[:, :n_out])
    loss = float(np.sum(e * e)) / (rows * n_out)
    scale = 2.0 / (rows * n_out)
    if emulate:
        scale = float(np.float32(scale))
    dz = rnd(acc(scale * e))
    grads = [None] * LAYERS
    for l in range(LAYERS - 1, -1, -1):
        dW = acc(dz.T @ H[l])
        db = acc(dz.sum(axis=0))
        grads[l] = (dW, db)
        if l > 0:
            dh = acc(dz @ Ws[l])
            dz = rnd(acc(dh * silu_grad(Z[l - 1])))
    return loss, y, grads
