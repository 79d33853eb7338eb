"""Pins of the learned preconditioner's oracle (oracle/mlp.py): bfloat16
rounding on hand-worked values, backward against central finite
differences, the features on a hand-checked vector, Adam's first step."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    # bfloat16 keeps 8 significand bits: 1 + 2^-7 is representable,
    # 1 + 2^-8 is a tie (rounds to even: 1), 1 + 3*2^-8 is a tie (to 1 + 2^-6)
    x = np.array([1.0, 1 + 2.0 ** -7, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -8 + 2.0 ** -12,
                  -2.5, 3.0e-3], dtype=np.float32)
    want = [1.0, 1 + 2.0 ** -7, 1.0, 1 + 2.0 ** -6, 1 + 2.0 ** -7, -2.5]
    got = mlp.bf(x)
    assert list(got[:6]) == want
    # 3e-3 = 1.536 * 2^-9: nearest bfloat16 is (1 + 69/128) * 2^-9
    assert got[6] == (1 + 69 / 128) * 2.0 ** -9


def test_silu_derivative():
    z = np.linspace(-6, 6, 25)
    h = 1e-6
    fd = (mlp.silu(z + h) - mlp.silu(z - h)) / (2 * h)
    assert np.abs(fd - mlp.dsilu(z)).max() < 1e-9


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(4)
    p = mlp.init_params(3).astype(np.float64)
    p[17152:17280] = rng.uniform(-0.3, 0.3, 128)      # non-zero output layer
    p[512:640] = rng.uniform(-0.1, 0.1, 128)
    n = 40
    x = rng.standard_normal((n, 4))
    ut = rng.standard_normal(n)
    dt = rng.standard_normal(n)
    # float64 throughout for the difference quotients
    def L(q):
        y, _ = mlp.forward(q, x.astype(np.float32), rounding=False)
        e = (1 + y) * ut.astype(np.float32) + dt.astype(np.float32)
        return np.mean(e * e)
    _, g = mlp.loss_grad(p, x, ut, dt, rounding=False)
    idx = list(rng.choice(mlp.P, 60, replace=False)) + [mlp.OFF['b3'], mlp.OFF['w3'], mlp.OFF['W1']]
    h = 1e-6
    for i in idx:
        q1, q2 = p.copy(), p.copy()
        q1[i] += h
        q2[i] -= h
        fd = (L(q1) - L(q2)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_untrained_network_is_the_plain_preconditioner():
    p = mlp.init_params(0)
    x = np.random.default_rng(1).standard_normal((10, 4)).astype(np.float32)
    y, _ = mlp.forward(p, x)
    assert np.all(y == 0.0)


def test_features_hand_checked():
    diag = np.array([4.0, 2.0, 1e-12])
    t = np.array([0.0, 0.5, 0.0])
    rest = np.array([1.0, -1.0, 2e-12])
    x, u, ru = mlp.features(diag, rest, t, it=16)
    assert np.all(mlp.features(diag, rest, t, it=70)[0][:, 3] == 1.0)    # clamped at 32
    # g = D t + rest = [1, 0, 2e-12]; u = g / D' = [0.25, 0, 0.02]
    assert np.allclose(u, [0.25, 0.0, 0.02])
    assert np.isclose(ru, np.sqrt((0.25 ** 2 + 0.02 ** 2) / 3))
    assert np.allclose(x[:, 0], [0.0, -0.5, -1.0])
    assert np.allclose(x[:, 1], np.arcsinh(np.array([0.25, 0.0, 0.02]) / ru), rtol=1e-6)
    rt = np.sqrt(0.25 / 3)
    assert np.allclose(x[:, 2], np.arcsinh(np.array([0.0, 0.5, 0.0]) / rt), rtol=1e-6)
    assert np.all(x[:, 3] == 0.5)


def test_adam_first_step():
    p = np.array([1.0, -2.0])
    g = np.array([0.5, -4.0])
    q, m, v = mlp.adam(p, np.zeros(2), np.zeros(2), g, step=1, lr=0.1)
    # bias-corrected first step moves every parameter by lr * sign(g)
    assert np.allclose(q, p - 0.1 * np.sign(g), atol=1e-7)
