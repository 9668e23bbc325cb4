"""Seeded synthetic initial conditions for the force-step workloads.

Positions are FP32-representable (gasoline's Tipsy reader stores FP32,
pkd.c:482-519), held in FP64 arrays.  The periodic box draws coordinates on a
2^-24 lattice so that every FP64 separation and squared distance is exact.
"""
import numpy as np


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def plummer(n, seed=1, mfrac_cut=0.999):
    """Plummer sphere, G = M = a = 1, equal masses.

    Radii from the inverse cumulative mass profile (Aarseth, Henon & Wielen
    1974) with the mass fraction capped at mfrac_cut; isotropic directions;
    speeds by rejection from g(q) = q^2 (1 - q^2)^(7/2) times the local escape
    speed sqrt(2) (1 + r^2)^(-1/4).
    Returns (r[n,3], v[n,3], m[n]).
    """
    rng = np.random.default_rng(seed)
    x1 = rng.uniform(0.0, mfrac_cut, n)
    x1 = np.maximum(x1, 1e-10)
    rad = 1.0 / np.sqrt(x1 ** (-2.0 / 3.0) - 1.0)
    dirs = _iso(rng, n)
    pos = rad[:, None] * dirs
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0, 1, todo.size)
        yy = rng.uniform(0, 0.1, todo.size)
        ok = yy < qq * qq * (1 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0) * (1.0 + rad * rad) ** (-0.25)
    vel = (q * vesc)[:, None] * _iso(rng, n)
    m = np.full(n, 1.0 / n)
    return _f32(pos), _f32(vel), m


def _iso(rng, n):
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    s = np.sqrt(1 - z * z)
    return np.stack([s * np.cos(ph), s * np.sin(ph), z], axis=1)


def sph_box(n, seed=2, sigma_v=0.1, u=1.0):
    """Adiabatic periodic box: n gas particles uniform in [0,1)^3 on a 2^-24
    lattice, equal mass 1/n, u = 1, Gaussian velocities (sigma 0.1 per axis).
    Returns dict of r, v, m, u."""
    rng = np.random.default_rng(seed)
    pos = rng.integers(0, 1 << 24, size=(n, 3)).astype(np.float64) / float(1 << 24)
    vel = _f32(rng.normal(0.0, sigma_v, size=(n, 3)))
    return dict(r=pos, v=vel, m=np.full(n, 1.0 / n), u=np.full(n, float(u)))


def evrard(n, seed=5, u=0.05):
    """Evrard collapse: cold gas sphere rho ~ 1/r inside R = 1, M = G = 1,
    u = 0.05, at rest.  Radii r = sqrt(X) (enclosed mass ~ r^2)."""
    rng = np.random.default_rng(seed)
    rad = np.sqrt(rng.uniform(0, 1, n))
    pos = rad[:, None] * _iso(rng, n)
    return dict(r=_f32(pos), v=np.zeros((n, 3)), m=np.full(n, 1.0 / n),
                u=np.full(n, float(u)))
