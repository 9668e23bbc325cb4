"""Pins of the MLP oracle (oracle/hmlp_oracle.py): the analytic gradient
against central finite differences, closed forms of the network, the Adam
step's first move, and the minibatch hash against hand-computed values."""
import numpy as np

from oracle import hmlp_oracle as M


def random_params(rng, scale=0.5):
    p = rng.normal(0, scale, M.NP)
    p[M.NW:M.NW + 4] = rng.normal(0, 1, 4)
    p[M.NW + 4:] = rng.uniform(0.5, 2.0, 4)
    return p


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = random_params(rng)
    x = rng.normal(0, 1.5, (37, 4))
    t = rng.normal(0, 0.3, 37)
    g, _ = M.grad(p, x, t)
    eps = 1e-6
    fd = np.zeros(M.NW)
    for k in range(M.NW):
        dp = np.zeros(M.NP)
        dp[k] = eps
        fd[k] = (M.loss(p + dp, x, t) - M.loss(p - dp, x, t)) / (2 * eps)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))
    # a dropped term would show: every block carries gradient here
    for a, b in ((M.OW1, M.OB1), (M.OB1, M.OW2), (M.OW2, M.OB2), (M.OB2, M.OW3),
                 (M.OW3, M.OB3)):
        assert np.max(np.abs(g[a:b])) > 1e-4


def test_closed_forms():
    rng = np.random.default_rng(2)
    p = random_params(rng)
    x = rng.normal(0, 1, (10, 4))
    # W3 = 0: the output is b3 and only W3, b3 receive gradient
    q = p.copy()
    q[M.OW3:M.OB3] = 0
    y, (xh, a1, a2) = M.forward(q, x)
    assert np.allclose(y, q[M.OB3])
    t = np.zeros(10)
    g, l = M.grad(q, x, t)
    assert np.all(g[:M.OB2 + M.H] == 0)
    assert np.isclose(g[M.OB3], 2 * q[M.OB3])
    assert np.isclose(l, q[M.OB3] ** 2)
    # standardisation: x = mean gives x^ = 0, so a1 = tanh(b1)
    x0 = np.tile(p[M.NW:M.NW + 4], (3, 1))
    _, (xh0, a10, _) = M.forward(p, x0)
    assert np.allclose(xh0, 0) and np.allclose(a10, np.tanh(p[M.OB1:M.OW2]))


def test_adam_first_step():
    # first step: m^ = g and v^ = g^2, so each parameter moves by lr*g/(|g|+eps)
    rng = np.random.default_rng(3)
    p = random_params(rng)
    g = rng.normal(0, 1, M.NW)
    p1, m, v = M.adam(p, g, np.zeros(M.NW), np.zeros(M.NW), 1, 1e-2)
    assert np.allclose(p - p1, np.r_[1e-2 * g / (np.abs(g) + 1e-8), np.zeros(8)], atol=1e-12)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_minibatch_hash():
    # hash3 restated step by step for (1, 0, 0)
    M32 = 0xFFFFFFFF
    h = (0x9E3779B1 ^ ((0x7F4A7C15 * 0x85EBCA77) & M32) ^ ((0x165667B1 * 0xC2B2AE3D) & M32))
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & M32
    h ^= h >> 12
    h = (h * 0x297A2D39) & M32
    h ^= h >> 15
    assert M.hash3(1, 0, 0) == h
    d = M.draw(7, 3, 1000, 97)
    assert d.min() >= 0 and d.max() < 97 and len(np.unique(d)) > 80


def test_features():
    f = M.features([4.0, 0.16], [2.0, 3.0], [0.5, -1.0], 0.1)
    assert np.allclose(f[:, 0], [0.0, np.log(0.2)])
    assert np.allclose(f[:, 2], [0.05, -0.1]) and np.allclose(f[:, 3], np.log(0.1))
