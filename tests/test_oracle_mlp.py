"""Pins the MLP oracle (oracle/mlp.py) by finite differences and closed forms."""
import numpy as np

from oracle import mlp


def test_gelu_derivative_by_fd():
    z = np.linspace(-5, 5, 101)
    h = 1e-6
    fd = (mlp.gelu(z + h) - mlp.gelu(z - h)) / (2 * h)
    assert np.allclose(fd, mlp.gelu_grad(z), atol=1e-8)
    assert mlp.gelu(np.array([0.0]))[0] == 0.0


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init(1)
    x = rng.standard_normal((5, mlp.IN))
    t = rng.standard_normal((5, mlp.OUT))
    y, cache = mlp.forward(p, x)
    loss, dy = mlp.mse(y, t)
    g = mlp.backward(p, cache, dy)
    assert g.shape == (mlp.N_PARAMS,)
    idx = rng.choice(mlp.N_PARAMS, 60, replace=False)
    h = 1e-6
    for k in idx:
        e = np.zeros(mlp.N_PARAMS)
        e[k] = h
        lp = mlp.mse(mlp.forward(p + e, x)[0], t)[0]
        lm = mlp.mse(mlp.forward(p - e, x)[0], t)[0]
        assert abs((lp - lm) / (2 * h) - g[k]) <= 1e-6 * max(1.0, abs(g[k]))


def test_adam_first_step_is_sign_times_lr():
    g = np.array([0.3, -2.0, 1e-3])
    p, m, v = mlp.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    assert np.allclose(p, -0.01 * np.sign(g))


def test_parameter_count():
    assert mlp.N_PARAMS == 144899
