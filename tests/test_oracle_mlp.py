"""CPU pins of the predictor network's oracle (oracle/mlp.py): the float64
backward pass against central finite differences and against torch
autograd, the forward pass against torch, and the bf16 rounding against
torch's bfloat16 conversion."""

import numpy as np
import pytest
import torch

from oracle import mlp as om


def _net(n_in=7, n_out=5, seed=0):
    rng = np.random.default_rng(seed)
    k0, n5 = (n_in + 15) // 16 * 16, (n_out + 15) // 16 * 16
    layers = []
    for l in range(om.LAYERS):
        K = k0 if l == 0 else om.HIDDEN
        N = n5 if l == om.LAYERS - 1 else om.HIDDEN
        kv = n_in if l == 0 else om.HIDDEN
        nv = n_out if l == om.LAYERS - 1 else om.HIDDEN
        W = np.zeros((N, K))
        W[:nv, :kv] = rng.uniform(-1, 1, (nv, kv)) / np.sqrt(kv)
        b = np.zeros(N)
        b[:nv] = rng.uniform(-0.2, 0.2, nv)
        layers.append((W, b))
    x = rng.normal(size=(24, n_in))
    y = rng.normal(size=(24, n_out))
    norm = {"x_mean": rng.normal(size=n_in) * 0.1, "x_scale": rng.uniform(0.5, 2.0, n_in),
            "y_mean": rng.normal(size=n_out) * 0.1, "y_scale": rng.uniform(0.5, 2.0, n_out)}
    X, T = om.normalise(x, y, norm, k0, n5, emulate=False)
    return layers, X, T, n_out


def test_oracle_gradients_match_finite_differences():
    layers, X, T, n_out = _net()
    loss, _, grads = om.forward_backward(layers, X, T, n_out, emulate=False)
    rng = np.random.default_rng(1)
    h = 1e-6
    for l in range(om.LAYERS):
        W, b = layers[l]
        nv = n_out if l == om.LAYERS - 1 else om.HIDDEN
        kv = 7 if l == 0 else om.HIDDEN
        for _ in range(4):
            i, j = int(rng.integers(nv)), int(rng.integers(kv))
            Wp, Wm = W.copy(), W.copy()
            Wp[i, j] += h
            Wm[i, j] -= h
            lp = om.forward_backward([*layers[:l], (Wp, b), *layers[l + 1:]], X, T, n_out, False)[0]
            lm = om.forward_backward([*layers[:l], (Wm, b), *layers[l + 1:]], X, T, n_out, False)[0]
            fd = (lp - lm) / (2 * h)
            assert abs(fd - grads[l][0][i, j]) <= 1e-6 * max(1e-3, abs(fd)), (l, i, j, fd, grads[l][0][i, j])
        i = int(rng.integers(nv))
        bp, bm = b.copy(), b.copy()
        bp[i] += h
        bm[i] -= h
        lp = om.forward_backward([*layers[:l], (W, bp), *layers[l + 1:]], X, T, n_out, False)[0]
        lm = om.forward_backward([*layers[:l], (W, bm), *layers[l + 1:]], X, T, n_out, False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grads[l][1][i]) <= 1e-6 * max(1e-3, abs(fd))
    assert loss > 0


def test_oracle_matches_torch_autograd():
    layers, X, T, n_out = _net(n_in=19, n_out=11, seed=3)
    loss, y, grads = om.forward_backward(layers, X, T, n_out, emulate=False)
    tW = [torch.tensor(W, requires_grad=True) for W, _ in layers]
    tb = [torch.tensor(b, requires_grad=True) for _, b in layers]
    a = torch.tensor(X)
    for l in range(om.LAYERS - 1):
        a = torch.nn.functional.silu(torch.nn.functional.linear(a, tW[l], tb[l]))
    ty = torch.nn.functional.linear(a, tW[-1], tb[-1])
    tl = torch.mean((ty[:, :n_out] - torch.tensor(T[:, :n_out])) ** 2)
    tl.backward()
    assert abs(tl.item() - loss) <= 1e-12 * loss
    np.testing.assert_allclose(ty.detach().numpy(), y, rtol=0, atol=1e-12)
    for l in range(om.LAYERS):
        np.testing.assert_allclose(grads[l][0], tW[l].grad.numpy(), rtol=1e-9, atol=1e-15)
        np.testing.assert_allclose(grads[l][1], tb[l].grad.numpy(), rtol=1e-9, atol=1e-15)


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(size=5000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 5000),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0, 65504.0],
                                 dtype=np.float32)])
    ref = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(om.bf16(v), ref)
    assert om.bf16(np.float32(1.0 + 2 ** -8)) == 1.0          # tie to even (down)
    assert om.bf16(np.float32(1.0 + 3 * 2 ** -8)) == 1.0 + 2 ** -6  # tie to even (up)


def test_flatten_layout_round_trip():
    layers, *_ = _net(n_in=33, n_out=20)
    flat = om.flatten(layers)
    back = om.unflatten(flat, 48, 32)
    for (W, b), (W2, b2) in zip(layers, back):
        np.testing.assert_array_equal(W, W2)
        np.testing.assert_array_equal(b, b2)
