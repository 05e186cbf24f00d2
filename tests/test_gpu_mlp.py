"""Warm-start MLP on the GPU: BF16 MFMA forward/backward against the FP64
oracle (oracle/mlp_oracle.py, pinned by finite differences), and training
that reduces the loss."""
import numpy as np
import pytest

import sankhya_gpu as sg
import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,B", [((16, 64, 100), 32), ((16, 512, 700), 64), ((5, 40, 3000), 70),
                                     ((16, 512, 1000), 200), ((16, 128, 20000), 64)])
def test_gradients_match_oracle(shape, B):
    n_in, hid, n_out = shape
    rng = np.random.default_rng(1)
    net = sg.Mlp(n_in, hid, n_out, seed=3)
    p = net.params()
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    T = rng.standard_normal((B, n_out)).astype(np.float32)
    im, isd = X.mean(0), X.std(0) + 0.1
    om, osd = T.mean(0), T.std(0) + 0.1
    net.set_scaling(im, isd, om, osd)
    loss, g = net.gradients(X, T)
    ref_loss, ref_g, _ = mo.forward_backward(p, X, T, n_in, hid, n_out, im.astype(np.float64),
                                             isd.astype(np.float64), om.astype(np.float64),
                                             osd.astype(np.float64), bf16=True)
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss)
    # blockwise relative error of the gradient (W1, b1, W2, b2, W3, b3)
    sizes = [n_in * hid, hid, hid * hid, hid, hid * n_out, n_out]
    o = 0
    for s in sizes:
        a, b = g[o:o + s].astype(np.float64), ref_g[o:o + s]
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-12, (s, np.linalg.norm(a - b) / np.linalg.norm(b))
        o += s


def test_predict_matches_oracle_forward():
    rng = np.random.default_rng(2)
    net = sg.Mlp(16, 128, 333, seed=5)
    X = rng.standard_normal((50, 16)).astype(np.float32)
    om = rng.standard_normal(333).astype(np.float32)
    osd = (rng.random(333) + 0.5).astype(np.float32)
    net.set_scaling(np.zeros(16), np.ones(16), om, osd)
    Y = net.predict(X)
    _, _, ref = mo.forward_backward(net.params(), X, np.zeros((50, 333)), 16, 128, 333, np.zeros(16), np.ones(16),
                                    om.astype(np.float64), osd.astype(np.float64), bf16=True)
    assert np.max(np.abs(Y - ref)) <= 1e-3 * (np.max(np.abs(ref)) + 1)


def test_training_reduces_loss_on_a_smooth_map():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((2048, 16)).astype(np.float32)
    M = rng.standard_normal((16, 200)).astype(np.float32) * 0.3
    Y = np.tanh(X @ M) + 0.1 * (X @ M) ** 2
    net = sg.Mlp(16, 256, 200, seed=9)
    net.set_scaling(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    hist = net.train(X, Y, epochs=30, batch=128, lr=2e-3)
    assert hist[-1] < 0.2 * hist[0], hist


def test_adam_step_and_both_weight_copies():
    """One training step (batch = the whole set): the parameters move by
    Adam's first step, p - lr * g / (|g| + eps) elementwise (m/(1-b1) = g,
    v/(1-b2) = g^2), and the BF16 weight copies the kernels read -- row-major
    for the backward products (refreshed by Adam), transposed for the
    forward ones -- are those of the new parameters: prediction and
    gradients afterwards match the oracle at the new point."""
    rng = np.random.default_rng(6)
    n_in, hid, n_out, B = 16, 192, 700, 96
    net = sg.Mlp(n_in, hid, n_out, seed=11)
    X = rng.standard_normal((B, n_in)).astype(np.float32)
    T = rng.standard_normal((B, n_out)).astype(np.float32)
    net.set_scaling(X.mean(0), X.std(0) + 0.1, T.mean(0), T.std(0) + 0.1)
    p0 = net.params().astype(np.float64)
    _, g = net.gradients(X, T)
    lr = 1e-3
    net.train(X, T, epochs=1, batch=B, lr=lr)
    p1 = net.params().astype(np.float64)
    g = g.astype(np.float64)
    step = p0 - lr * g / (np.abs(g) + 1e-8)
    big = np.abs(g) > 1e-5          # the batch order changes g's last bits
    assert big.mean() > 0.5
    assert np.max(np.abs(p1 - step)[big]) <= 1e-3 * lr
    assert np.max(np.abs(p1 - p0)) <= 1.01 * lr
    # the copies: forward (transposed copies) and backward (row-major copies)
    im, isd, om, osd = (a.astype(np.float64) for a in (X.mean(0), X.std(0) + 0.1, T.mean(0), T.std(0) + 0.1))
    Y = net.predict(X)
    _, ref_g, ref_y = mo.forward_backward(net.params(), X, T, n_in, hid, n_out, im, isd, om, osd, bf16=True)
    assert np.max(np.abs(Y - ref_y)) <= 1e-3 * (np.max(np.abs(ref_y)) + 1)
    _, g1 = net.gradients(X, T)
    assert np.linalg.norm(g1 - ref_g) <= 2e-2 * np.linalg.norm(ref_g)
