"""Pins of the surrogate-network oracle (oracle/mlp_oracle.py) on the CPU: the analytic
gradient against central finite differences, a closed form, SiLU's derivative, BF16 rounding,
the Adam step and the pair features."""
import numpy as np

import mlp_oracle as M


def test_layout_size():
    assert M.NW == 33793  # 4*128+128 + 2*(128*128+128) + 128 + 1


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    theta = M.glorot(1) + rng.normal(0, 0.05, M.NW)  # nonzero biases too
    X = rng.normal(0, 1, (40, 4))
    y = rng.normal(0, 1, 40)
    _, g = M.loss_grad(theta, X, y)
    idx = np.concatenate([rng.choice(np.arange(o, o + int(np.prod(s))), 6, replace=False)
                          for o, s in M.OFF.values() if int(np.prod(s)) > 1] + [[M.OFF["b4"][0]]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_grad(tp, X, y)[0] - M.loss_grad(tm, X, y)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_constant_network_closed_form():
    """All weights zero: f = b4, loss = mean((b4 - y)^2)/2, dL/db4 = mean(b4 - y), and the
    only other nonzero gradient is w4's (SiLU(0) = 0 makes even that vanish)."""
    theta = np.zeros(M.NW)
    theta[M.OFF["b4"][0]] = 0.7
    y = np.array([0.1, 0.5, 2.0, -1.0])
    X = np.ones((4, 4))
    loss, g = M.loss_grad(theta, X, y)
    assert np.allclose(M.forward(theta, X), 0.7)
    assert abs(loss - 0.5 * np.mean((0.7 - y) ** 2)) < 1e-15
    assert abs(g[M.OFF["b4"][0]] - np.mean(0.7 - y)) < 1e-15
    assert np.count_nonzero(g) == 1


def test_silu_derivative():
    z = np.linspace(-6, 6, 101)
    fd = (M.silu(z + 1e-6) - M.silu(z - 1e-6)) / 2e-6
    assert np.allclose(M.silu_prime(z), fd, atol=1e-8)


def test_bf16_rounding():
    assert M.bf16(1.0) == 1.0
    assert M.bf16(1.0 + 2 ** -8) == 1.0          # halfway: ties to even
    assert M.bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6
    assert M.bf16(-3.14159) == -3.140625
    assert np.all(np.abs(M.bf16(np.linspace(-5, 5, 1001)) - np.linspace(-5, 5, 1001)) <= 2 ** -8 * 5)


def test_adam_first_step_is_signed_lr():
    """With bias correction the first Adam step is lr * g/(|g| + eps) = lr * sign(g)."""
    g = np.array([3.0, -0.5, 1e-3])
    th, m, v = M.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(th, -0.01 * np.sign(g), rtol=1e-4)


def test_pair_features_body_frame():
    """Identity frame: the features are the distance and the lab direction cosines; a rotation
    of frame and partner together leaves them unchanged."""
    f = M.pair_features(np.eye(3), 1.0, [1, 2, 3], [4, 6, 3])
    assert np.allclose(f, [5.0, 0.6, 0.8, 0.0])
    c, s = np.cos(0.7), np.sin(0.7)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    g = M.pair_features(np.eye(3) @ R.T, 1.0, [0, 0, 0], R @ np.array([3.0, 4.0, 0.0]))
    assert np.allclose(g, [5.0, 0.6, 0.8, 0.0])
