"""Pins the oracle's warm-start network (oracle/mlp.py): the analytic gradient against
central finite differences of the loss, the BF16 rounding against exact cases, and the
features against a direct evaluation of their definition."""
import math

import numpy as np

from oracle import mlp


def _params(seed):
    rng = np.random.default_rng(seed)
    p = rng.standard_normal(mlp.P) * 0.15
    return mlp.bf16_round(p.astype(np.float32))  # weights already on the BF16 grid


def test_bf16_rounding_exact_cases():
    a = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 0.0], dtype=np.float32)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    np.testing.assert_array_equal(mlp.bf16_round(a), [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5, 0.0])


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    norm = np.concatenate([rng.standard_normal(mlp.F) * 0.1, 1 + rng.uniform(size=mlp.F), [0.2, -1.0], [1.5, 0.7]])
    X = rng.standard_normal((40, mlp.F))
    Y = rng.standard_normal((40, 2))
    L, g = mlp.loss_and_grad(p, norm, X, Y)
    # biases are continuous; weights are differentiated as continuous values at their
    # BF16 points, so perturb a copy whose weights are not re-rounded
    def loss_cont(q):
        d = mlp.unpack(q, round_weights=False)
        x = (X - norm[:mlp.F]) / norm[mlp.F:2 * mlp.F]
        s = lambda z: z / (1 + np.exp(-z))  # noqa: E731
        a3 = s(s(s(x @ d["W1"] + d["b1"]) @ d["W2"] + d["b2"]) @ d["W3"] + d["b3"])
        out = a3 @ d["W4"] + d["b4"]
        yt = (Y - norm[2 * mlp.F:2 * mlp.F + 2]) / norm[2 * mlp.F + 2:]
        return 0.5 * np.sum((out - yt) ** 2) / X.shape[0]
    assert abs(loss_cont(p) - L) < 1e-12
    idx = rng.choice(mlp.P, size=60, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros(mlp.P)
        e[i] = h
        fd = (loss_cont(p + e) - loss_cont(p - e)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_features_follow_definition():
    rng = np.random.default_rng(3)
    t = np.arange(9.0)
    r = rng.uniform(size=(9, 5, 2))
    y = rng.poisson(2.0, size=(9, 5)).astype(float)
    X = mlp.features(t, r, y, 10.0, 0.2, 1.0, 2.0)
    k, i, b = 4, 2, 1
    num = den = 0.0
    for q in range(max(0, k - 6), min(8, k + 6) + 1):
        for j in range(5):
            w = math.exp(-0.5 * (t[q] - t[k]) ** 2 / mlp.HT[b] ** 2
                         - 0.5 * np.sum((r[q, j] - r[k, i]) ** 2) / mlp.HS[b] ** 2)
            num += w * y[q, j]
            den += w
    assert abs(X[k, i, 1 + b] - math.log((num + 1) / (den + 1) / 2.0)) < 1e-12
    assert abs(X[k, i, 0] - math.log1p(y[k, i])) < 1e-15
    np.testing.assert_allclose(X[k, i, 4:], np.log([10.0, 0.2, 1.0, 2.0]))
