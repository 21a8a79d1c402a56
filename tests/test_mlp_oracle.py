"""Pins the FP64 probe-network oracle (oracle/mlp.py) by central finite
differences of its own loss (CPU only)."""
import numpy as np

from oracle import mlp


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    din, hid, dout, n = 5, 7, 6, 4
    W1 = rng.normal(0, 0.5, (din, hid))
    b1 = rng.normal(0, 0.1, hid)
    W2 = rng.normal(0, 0.5, (hid, dout))
    b2 = rng.normal(0, 0.1, dout)
    X = rng.normal(0, 1, (n, din))
    T = rng.random((n, dout))
    T /= T.sum(axis=1, keepdims=True)
    g = mlp.loss_grad(W1, b1, W2, b2, X, T, round_like_gpu=False)
    params = {"gW1": W1, "gb1": b1, "gW2": W2, "gb2": b2}
    eps = 1e-6
    for name, P in params.items():
        flat = P.reshape(-1)
        for i in range(0, flat.size, max(1, flat.size // 9)):
            old = flat[i]
            flat[i] = old + eps
            up = mlp.mean_loss(W1, b1, W2, b2, X, T)
            flat[i] = old - eps
            dn = mlp.mean_loss(W1, b1, W2, b2, X, T)
            flat[i] = old
            fd = (up - dn) / (2 * eps)
            assert abs(fd - g[name].reshape(-1)[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i)


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.5, 65504.0], np.float32)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa (1.0) wins
    assert mlp.bf16(x).tolist() == [1.0, 1.0, 1.015625, -3.5, 65536.0]
