"""GPU checks of the learned predictor's network (strut_amd/csrc/predictor.hip)
through the C ABI: forward outputs and gradients against the CPU oracle
(oracle/mlp.py, with bf16 rounding at the device's rounding points), padding
that stays zero, and a training run that lowers the loss.

Tolerances: the device sums bf16 products in FP32 inside the matrix cores and
the oracle sums them exactly; the two round to bf16 at the same points, so
they differ where a value sits within an FP32 ulp of a bf16 rounding
boundary.  Gradients agree to 2e-3 relative (Frobenius, per layer) and
outputs to 1e-3."""

import numpy as np
import pytest
import torch

from oracle import mlp as om
from strut_amd.learned import Mlp

pytestmark = pytest.mark.gpu


def _setup(n_in, n_out, rows, seed):
    net = Mlp(n_in, n_out, seed=seed)
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(rows, n_in)).astype(np.float32) * rng.uniform(0.5, 3.0, n_in).astype(np.float32)
    y = (np.tanh(x[:, :1]) * rng.normal(size=(1, n_out)) + 0.1 * rng.normal(size=(rows, n_out))).astype(np.float32)
    norm = {"x_mean": rng.normal(size=n_in).astype(np.float32) * 0.1,
            "x_scale": rng.uniform(0.3, 1.5, n_in).astype(np.float32),
            "y_mean": rng.normal(size=n_out).astype(np.float32) * 0.1,
            "y_scale": rng.uniform(0.5, 2.0, n_out).astype(np.float32)}
    net.set_normalization(norm["x_mean"], norm["x_scale"], norm["y_mean"], norm["y_scale"])
    return net, x, y, norm


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("n_in,n_out,rows", [(242, 48, 96), (37, 5, 40), (602, 120, 64)])
def test_gpu_gradients_match_oracle(n_in, n_out, rows):
    print(f"case n_in={n_in} n_out={n_out} rows={rows}", flush=True)
    net, x, y, norm = _setup(n_in, n_out, rows, seed=n_in)
    lay = net.layout()
    dev = torch.device("cuda", 0)
    g, loss = net.gradients(torch.tensor(x, device=dev), torch.tensor(y, device=dev))
    layers = om.unflatten(net.params(), lay["k0"], lay["n5"])
    X, T = om.normalise(x.astype(np.float64), y.astype(np.float64), norm, lay["k0"], lay["n5"], emulate=True)
    ref_loss, _, ref = om.forward_backward(layers, X, T, n_out, emulate=True)
    assert abs(loss - ref_loss) <= 1e-4 * ref_loss
    got = om.unflatten(g, lay["k0"], lay["n5"])
    for l in range(om.LAYERS):
        ew, eb = _rel(got[l][0], ref[l][0]), _rel(got[l][1], ref[l][1])
        print(f"layer {l}: dW rel {ew:.2e} db rel {eb:.2e}")
        assert ew <= 2e-3 and eb <= 2e-3
    # padding rows / columns receive no gradient
    assert not np.any(got[0][0][:, n_in:]) and not np.any(got[-1][0][n_out:]) and not np.any(got[-1][1][n_out:])
    net.close()


def test_gpu_predict_matches_oracle_forward():
    n_in, n_out, rows = 1262, 252, 1100  # the 12-story frame's sizes, two inference chunks
    net, x, y, norm = _setup(n_in, n_out, rows, seed=9)
    lay = net.layout()
    dev = torch.device("cuda", 0)
    out = net.predict(torch.tensor(x, device=dev)).cpu().numpy().astype(np.float64)
    layers = om.unflatten(net.params(), lay["k0"], lay["n5"])
    X, T = om.normalise(x.astype(np.float64), y.astype(np.float64), norm, lay["k0"], lay["n5"], emulate=True)
    _, yn, _ = om.forward_backward(layers, X, T, n_out, emulate=True)
    ref = yn[:, :n_out] * norm["y_scale"].astype(np.float64) + norm["y_mean"].astype(np.float64)
    e = _rel(out, ref)
    print("forward rel", e)
    assert e <= 1e-3
    net.close()


def test_gpu_training_lowers_loss_and_keeps_padding_zero():
    rng = np.random.default_rng(11)
    n_in, n_out, rows = 50, 9, 4096
    x = rng.normal(size=(rows, n_in)).astype(np.float32)
    A = rng.normal(size=(n_in, n_out)) / np.sqrt(n_in)
    y = (np.sin(x @ A) + 0.3 * (x @ A) ** 2).astype(np.float32)
    dev = torch.device("cuda", 0)
    xt, yt = torch.tensor(x, device=dev), torch.tensor(y, device=dev)
    net = Mlp(n_in, n_out, seed=4)
    net.fit_normalization(xt, yt)
    nrm = net.normalization()
    np.testing.assert_allclose(nrm["x_mean"], x.mean(0), atol=1e-5)
    np.testing.assert_allclose(1.0 / nrm["x_scale"], x.std(0), rtol=1e-4)
    losses = net.train(xt, yt, batch=256, epochs=15, lr=2e-3, seed=1)
    print("losses", np.round(losses, 5))
    assert np.all(np.isfinite(losses)) and losses[-1] < 0.1 * losses[0]
    lay = net.layout()
    assert lay["adam_steps"] == 15 * (rows // 256)
    layers = om.unflatten(net.params(), lay["k0"], lay["n5"])
    assert not np.any(layers[0][0][:, n_in:]) and not np.any(layers[-1][0][n_out:])
    pred = net.predict(xt).cpu().numpy()
    assert _rel(pred, y) < 0.35
    net.close()
