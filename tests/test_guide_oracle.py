"""Pins of the guiding-network oracle (oracle/guide.py): finite differences, closed forms."""
import numpy as np
import pytest

from oracle import guide as G


def _problem(m=64, seed=0):
    rng = np.random.default_rng(seed)
    params = np.zeros(G.PARAMS)
    for l in range(5):
        lim = np.sqrt(6.0 / G.K[l])
        params[G.W_OFF[l]:G.W_OFF[l] + G.K[l] * G.N[l]] = rng.uniform(-lim, lim, G.K[l] * G.N[l])
        params[G.B_OFF[l]:G.B_OFF[l] + G.N[l]] = rng.uniform(-0.1, 0.1, G.N[l])
    p01 = rng.random((m, 3))
    n = rng.standard_normal((m, 3))
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    wo = rng.standard_normal((m, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    feat = G.encode(p01, n, wo)
    d = rng.standard_normal((m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, m)
    return params, feat, d, w


def test_gradient_matches_finite_differences():
    params, feat, d, w = _problem()
    g = G.grad(params, feat, d, w, bf=False)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(np.arange(G.W_OFF[l], G.W_OFF[l + 1]), 6, replace=False)
                          for l in range(5)] + [rng.choice(np.arange(G.B_OFF[0], G.PARAMS), 10)])
    h = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (G.loss(p1, feat, d, w) - G.loss(p2, feat, d, w)) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-4, abs=1e-8), i


def test_vmf_mixture_is_a_density():
    rng = np.random.default_rng(2)
    o = rng.standard_normal((1, 48))
    o[0, 1:40:5] = rng.uniform(-1, 4, 8)
    # Monte Carlo over the sphere: integral of q = 1
    n = 400_000
    w = rng.standard_normal((n, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    q = np.exp(-G.nll(np.repeat(o, n, 0), w))
    assert q.mean() * 4 * np.pi == pytest.approx(1.0, abs=0.02)


def test_nll_grad_matches_finite_differences():
    rng = np.random.default_rng(3)
    o = rng.standard_normal((5, 48))
    o[:, 1:40:5] = rng.uniform(-2, 5, (5, 8))
    w = rng.standard_normal((5, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    g = G.nll_grad(o, w, np.ones(5))
    h = 1e-6
    for j in range(40):
        o1, o2 = o.copy(), o.copy()
        o1[:, j] += h
        o2[:, j] -= h
        fd = (G.nll(o1, w) - G.nll(o2, w)) / (2 * h)
        np.testing.assert_allclose(g[:, j], fd, rtol=1e-4, atol=1e-7)


def test_bf16_rounding():
    # representable values are unchanged
    x = np.array([1.0, 1.0078125, -2.5, 0.0], np.float32)
    np.testing.assert_array_equal(G.bf16(x), x.astype(np.float64))
    # 1 + 2^-8 is half way between bf16 1.0 and 1 + 2^-7: ties to even -> 1.0
    assert G.bf16(np.float32([1.0 + 2 ** -8]))[0] == 1.0
    assert G.bf16(np.float32([1.0 + 3 * 2 ** -9]))[0] == 1.0 + 2 ** -7
    # 1 + 2^-7 + 2^-8 is half way between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
    assert G.bf16(np.float32([1.0 + 2 ** -7 + 2 ** -8]))[0] == 1.0 + 2 ** -6
