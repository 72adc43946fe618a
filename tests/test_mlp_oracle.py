"""Pins of the CPU reference of the surrogate network (oracle/mlp_ref.py):
its backward pass against central finite differences in float64, and its
bfloat16 rounding against hand-computed bit patterns."""
import numpy as np
import pytest

from oracle import mlp_ref as R


def _problem(seed=3, n=7):
    rng = np.random.default_rng(seed)
    p = np.concatenate([np.concatenate([rng.normal(0, 1 / np.sqrt(i), o * i), rng.normal(0, 0.1, o)])
                        for i, o in zip(R.FAN_IN, R.FAN_OUT)])
    xn = rng.uniform(-1, 1, (n, 8))
    t = rng.normal(0, 1, (n, 3))
    w = (rng.uniform(size=n) > 0.2).astype(np.float64)
    return p, xn, t, w


def test_parameter_count():
    assert R.NPARAM == 13251


def test_gradient_matches_central_differences():
    p, xn, t, w = _problem()
    mean, sd = np.array([0.1, -0.2, 0.3]), np.array([1.5, 0.7, 2.0])
    loss, g = R.loss_grad(p, xn, t, w, mean, sd)
    rng = np.random.default_rng(11)
    idx = np.concatenate([rng.choice(R.NPARAM, 60, replace=False), [0, 511, 512, 575, 13055, 13250]])
    h = 1e-6
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (R.loss_grad(pp, xn, t, w, mean, sd)[0] - R.loss_grad(pm, xn, t, w, mean, sd)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 * max(1.0, abs(g[k])) + 1e-9, (k, fd, g[k])
    assert loss > 0


def test_gradient_of_a_planted_error_fails():
    # a dropped activation-derivative factor in the backward pass is caught by the check above
    p, xn, t, w = _problem(seed=5)
    mean, sd = np.zeros(3), np.ones(3)
    _, g = R.loss_grad(p, xn, t, w, mean, sd)
    layers = R.unpack(p)
    y, cache = R.forward(p, xn)
    delta = 2.0 * w[:, None] * (y - t) / xn.shape[0]
    h, Wq, z = cache[3]
    wrong = (delta @ cache[4][1]).T @ h  # no act'(z3)
    k0 = sum(i * o + o for i, o in zip(R.FAN_IN[:3], R.FAN_OUT[:3]))
    right = g[k0:k0 + 64 * 64].reshape(64, 64)
    assert np.max(np.abs(wrong - right)) > 1e-3 * np.max(np.abs(right))
    del layers


def test_bf16_rounding_bit_patterns():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -2.5, 3.0e-39], np.float32)
    # 1+2^-8 is a tie -> even (1.0); 1+3*2^-9 rounds up to 1+2^-7; 1+2^-7+2^-8 is a
    # tie with an odd last bit -> 1+2^-6
    want = [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5]
    got = R.bf16_round(vals)
    assert list(got[:5]) == want
    assert got[5] == pytest.approx(3.0e-39, rel=2 ** -7)


def test_quantised_forward_differs_only_by_rounding():
    p, xn, _, _ = _problem(n=64)
    y0, _ = R.forward(p, xn)
    y1, _ = R.forward(p, xn, quant=True)
    d = np.abs(y1 - y0)
    assert 0 < d.max() < 0.1
