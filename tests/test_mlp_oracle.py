"""The MLP oracle itself, pinned by central finite differences (FP64)."""
import numpy as np

import mlp_oracle as mo


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, hid, n_out, B = 3, 5, 4, 6
    npar = n_in * hid + hid + hid * hid + hid + hid * n_out + n_out
    p = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((B, n_in))
    T = rng.standard_normal((B, n_out))
    args = (n_in, hid, n_out, np.zeros(n_in), np.ones(n_in), np.zeros(n_out), np.ones(n_out))
    loss, g, _ = mo.forward_backward(p, X, T, *args)
    h = 1e-6
    fd = np.zeros(npar)
    for k in range(npar):
        e = np.zeros(npar); e[k] = h
        fd[k] = (mo.forward_backward(p + e, X, T, *args)[0] - mo.forward_backward(p - e, X, T, *args)[0]) / (2 * h)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-8)


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], np.float32)
    r = mo.bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6
    assert r[3] == -2.5 and abs(r[4] - 3.0e-3) < 3.0e-3 * 2 ** -8
