"""The FP64 restatement of the learned-closure network (oracle/mlp_ref.py),
pinned by central finite differences of its own loss: its analytic
gradient is what the GPU backward kernels are checked against."""

import numpy as np

from mlp_ref import NPARAMS, OFF, bf16_round, init_theta, loss_and_gradient


def _batch(n=24, seed=1):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 64))
    X[:, :55] = rng.standard_normal((n, 55))
    T = np.zeros((n, 6))
    T[:, 0] = 0.3 * rng.standard_normal(n)
    T[:, 1] = 0.3 * rng.standard_normal(n)
    T[:, 2] = rng.standard_normal(n) * 1e-3
    T[:, 3] = rng.standard_normal(n) * 1e-3
    T[:, 4] = 1.0 + rng.random(n)
    return X, T


def test_gradient_matches_central_differences():
    X, T = _batch()
    theta = init_theta(3)
    theta[OFF["b3"]] = 0.3
    lossnorm, kref = 0.7, 150.0
    _, g = loss_and_gradient(theta, X, T, lossnorm, kref, bf16=False)
    rng = np.random.default_rng(5)
    picks = np.concatenate([rng.choice(np.arange(OFF["W1"], OFF["b1"]), 12, replace=False),
                            rng.choice(np.arange(OFF["b1"], OFF["W2"]), 6, replace=False),
                            rng.choice(np.arange(OFF["W2"], OFF["b2"]), 12, replace=False),
                            rng.choice(np.arange(OFF["b2"], OFF["w3"]), 6, replace=False),
                            rng.choice(np.arange(OFF["w3"], OFF["b3"]), 6, replace=False), [OFF["b3"]]])
    for p in picks:
        h = 1e-6 * max(1.0, abs(theta[p]))
        tp, tm = theta.copy(), theta.copy()
        tp[p] += h
        tm[p] -= h
        fd = (loss_and_gradient(tp, X, T, lossnorm, kref, bf16=False)[0] -
              loss_and_gradient(tm, X, T, lossnorm, kref, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(g).max()), (p, fd, g[p])


def test_bf16_rounding_is_round_to_nearest_even():
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: even mantissa wins (1.0);
    # 1 + 3*2^-8 ties between 1 + 2^-7 and 1 + 2^-6: rounds to 1 + 2^-6
    v = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 3.0e-3])
    r = bf16_round(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2.0 ** -6 and r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8
    assert NPARAMS == 64 * 64 * 2 + 64 * 3 + 1
