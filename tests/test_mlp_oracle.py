"""Pins the MLP oracle (oracle/mlp.py): its exact FP64 gradient against central finite
differences, and the BF16-operand variant against the exact one."""
import numpy as np

from oracle import mlp


def _data(seed, S=24):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(S, 5)).astype(np.float32) * 2 + 1
    y = rng.normal(size=S).astype(np.float32)
    p = rng.normal(size=mlp.P_COUNT) * 0.2
    p[mlp.P_W1 + 5 * mlp.H:mlp.P_B1] = 0.0   # unused input rows
    norm = (np.full(5, 1.0), np.full(5, 2.0), 0.5, 1.5)
    return p, X, y, norm


def test_exact_gradient_matches_finite_differences():
    p, X, y, norm = _data(0)
    loss, g, _ = mlp.forward_backward(p, X, y, *norm, bf16_ops=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(mlp.P_W1, mlp.P_W1 + 5 * mlp.H), 12),
                          rng.choice(np.arange(mlp.P_B1, mlp.P_COUNT), 30)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp = mlp.forward_backward(pp, X, y, *norm, bf16_ops=False)[0]
        lm = mlp.forward_backward(pm, X, y, *norm, bf16_ops=False)[0]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_variant_close_to_exact():
    p, X, y, norm = _data(2)
    l0, g0, _ = mlp.forward_backward(p, X, y, *norm, bf16_ops=False)
    l1, g1, _ = mlp.forward_backward(p, X, y, *norm, bf16_ops=True)
    assert abs(l0 - l1) <= 0.02 * l0
    assert np.linalg.norm(g1 - g0) <= 0.05 * np.linalg.norm(g0)


def test_bf16_rounding_known_values():
    # bf16 spacing at 1.0 is 2^-7: 1 + 2^-8 is a tie (to even: 1.0), 1 + 3*2^-8 a tie that
    # rounds up to the even 1 + 2^-6, 1 + 2^-7 + 2^-9 rounds down
    vals = np.array([1.0, 1.00390625, 1.01171875, 1.009765625, -3.14159, 0.0], np.float32)
    r = mlp.bf16(vals)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625
    assert r[3] == 1.0078125 and r[5] == 0.0
    assert abs(r[4] + 3.140625) < 1e-12
