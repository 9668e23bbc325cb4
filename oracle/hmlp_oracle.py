"""CPU restatement (numpy, FP64) of the h-proposal MLP: forward pass,
gradient of the mean squared error, the Adam step and the minibatch draw.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
MLP kernels.  The product never imports it.

The network has no counterpart in the reference (it seeds smSmooth's search
radius, smooth.c:1439); what this file restates is the network's own
mathematics, and it is pinned by central finite differences and closed forms
in tests/test_oracle_hmlp.py.

Parameter vector (NP floats): W1[32][4], b1[32], W2[32][32], b2[32], W3[32],
b3, input mean[4], input scale[4].
"""
import numpy as np

H = 32
OW1, OB1 = 0, 4 * H
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H
NW = OB3 + 1          # trained parameters
NP = NW + 8           # + mean, scale


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OW1:OB1].reshape(H, 4), b1=p[OB1:OW2], W2=p[OW2:OB2].reshape(H, H),
                b2=p[OB2:OW3], W3=p[OW3:OB3], b3=p[OB3], mu=p[NW:NW + 4], sd=p[NW + 4:NW + 8])


def forward(p, x):
    """x (n, 4) -> y (n,), with the activations (xh, a1, a2)."""
    q = unpack(p)
    xh = (np.asarray(x, np.float64) - q["mu"]) / q["sd"]
    a1 = np.tanh(xh @ q["W1"].T + q["b1"])
    a2 = np.tanh(a1 @ q["W2"].T + q["b2"])
    y = a2 @ q["W3"] + q["b3"]
    return y, (xh, a1, a2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def grad(p, x, t):
    """Gradient of mean (y - t)^2 with respect to the NW trained parameters."""
    q = unpack(p)
    y, (xh, a1, a2) = forward(p, x)
    n = len(y)
    d3 = 2.0 * (y - np.asarray(t, np.float64)) / n          # dL/dy
    d2 = d3[:, None] * q["W3"][None, :] * (1 - a2 ** 2)      # dL/dz2
    d1 = (d2 @ q["W2"]) * (1 - a1 ** 2)                      # dL/dz1
    g = np.zeros(NW)
    g[OW1:OB1] = (d1.T @ xh).ravel()
    g[OB1:OW2] = d1.sum(0)
    g[OW2:OB2] = (d2.T @ a1).ravel()
    g[OB2:OW3] = d2.sum(0)
    g[OW3:OB3] = a2.T @ d3
    g[OB3] = d3.sum()
    return g, float(np.mean((y - t) ** 2))


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015) on the NW trained parameters;
    step counts from 1.  Returns new (p, m, v)."""
    p = np.array(p, np.float64)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    p[:NW] -= lr * (m / c1) / (np.sqrt(v / c2) + eps)
    return p, m, v


def hash3(a, b, c):
    """The library's minibatch hash (uint32 arithmetic)."""
    M = 0xFFFFFFFF
    a, b, c = int(a) & M, int(b) & M, int(c) & M
    h = ((a * 0x9E3779B1) & M) ^ (((b + 0x7F4A7C15) & M) * 0x85EBCA77 & M) \
        ^ (((c + 0x165667B1) & M) * 0xC2B2AE3D & M)
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & M
    h ^= h >> 12
    h = (h * 0x297A2D39) & M
    h ^= h >> 15
    return h


def draw(seed, it, nb, ns):
    """Sample indices of minibatch it: hash3(seed, it, k) mod ns, k < nb."""
    return np.array([hash3(seed, it, k) % ns for k in range(nb)], np.int64)


def features(fb2, rho, divv, dt):
    """(log h, log rho, divv dt, log dt) with h = sqrt(fBall2) / 2."""
    fb2 = np.asarray(fb2, np.float64)
    return np.stack([0.5 * np.log(fb2) - np.log(2.0), np.log(rho), np.asarray(divv) * dt,
                     np.full(len(fb2), np.log(dt))], 1)
