"""Pins the FP64 MLP oracle: gradients against central finite differences of
its own forward pass (no bf16 rounding: the mathematics fixes the gradient),
and the bf16 rounding helper against exact bit patterns."""
import numpy as np

from oracle import mlp


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], np.float32)
    r = mlp.bf16(x)
    # ties to even: 1 + 2^-8 -> 1; 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert np.all(np.asarray(r, np.float32).view(np.uint32) & 0xFFFF == 0)


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    p = mlp.init(1)
    p[mlp.OFF["b1"]:mlp.OFF["b1"] + 256] = rng.normal(0, 0.1, 256)
    x = rng.normal(0, 1, (5, 12))
    t = rng.normal(0, 1, 5)
    y, cache = mlp.forward(p, x, emulate_bf16=False)
    loss, dy = mlp.mse_loss_grad(y, t)
    g = mlp.backward(p, cache, dy, emulate_bf16=False)
    idx = np.concatenate([rng.choice(np.arange(o, o + n), 6, replace=False) for o, n in (
        (mlp.OFF["W1"], 256 * 32), (mlp.OFF["b1"], 256), (mlp.OFF["W2"], 65536), (mlp.OFF["b2"], 256),
        (mlp.OFF["W3"], 65536), (mlp.OFF["W4"], 65536), (mlp.OFF["b4"], 256), (mlp.OFF["w5"], 256))]
                         + [[mlp.OFF["b5"]]])
    idx = [i for i in idx if not (i < 8192 and (i % 32) >= 12)]
    eps = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += eps
        lp = mlp.mse_loss_grad(mlp.forward(q, x, emulate_bf16=False)[0], t)[0]
        q[i] -= 2 * eps
        lm = mlp.mse_loss_grad(mlp.forward(q, x, emulate_bf16=False)[0], t)[0]
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # padded input columns carry no gradient
    assert np.all(g[:8192].reshape(256, 32)[:, 12:] == 0)


def test_silu_derivative_closed_form():
    z = np.linspace(-8, 8, 101)
    eps = 1e-6
    fd = (mlp.silu(z + eps) - mlp.silu(z - eps)) / (2 * eps)
    assert np.allclose(mlp.dsilu(z), fd, atol=1e-8)
    assert mlp.silu(0.0) == 0.0 and abs(mlp.dsilu(0.0) - 0.5) < 1e-15
