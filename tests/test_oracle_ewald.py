"""Pins of the oracle's periodic direct sum (orc_direct_ewald), the force
against which the periodic tree + Ewald forces' errors are stated, by what
the mathematics fixes: lattice Madelung constants, the independence of the
real / wave-vector split, the small-r expansion, the spline inside twoh, and
the reference's own Ewald expansion (pkdBucketEwald, ewald.c:15-178, as
restated in the oracle) for a point-mass root, whose Green's function it is."""
import numpy as np
import pytest

from oracle import oracle as O

# Madelung constants: the simple cubic lattice of one unit mass per cell
# with its neutralising background (the regular part of the potential at a
# site, in units of 1/L), and rock salt (per ion, in units of 1/a, a the
# nearest-neighbour distance)
SC = 2.837297479480620
NACL = 1.747564594633182


@pytest.mark.parametrize("L", [1.0, 2.5])
@pytest.mark.parametrize("alphaL", [2.0, 3.0, 6.0])
def test_simple_cubic_madelung(L, alphaL):
    assert O.ewald_self(L, alphaL) * L == pytest.approx(SC, rel=1e-13)
    # one particle: its own images and the background give m xi and no force
    a, p = O.direct_ewald(np.array([[0.1, -0.2, 0.3]]) * L, np.array([0.7]), np.array([0.01]),
                          L, alphaL=alphaL)
    assert p[0] == pytest.approx(0.7 * SC / L, rel=1e-13)
    assert np.abs(a).max() < 1e-14


def _rock_salt(L):
    pts, q = [], []
    for x in range(2):
        for y in range(2):
            for z in range(2):
                pts.append([x * L / 2, y * L / 2, z * L / 2])
                q.append(1.0 if (x + y + z) % 2 == 0 else -1.0)
    return np.array(pts) - L / 4, np.array(q)


@pytest.mark.parametrize("alphaL", [2.0, 6.0])
def test_rock_salt_madelung(alphaL):
    # signed masses: the sum is linear in them and the cell is neutral, so
    # the background drops out; psi = -V, so an ion of sign q sits at q M / a
    L = 1.3
    r, q = _rock_salt(L)
    a, p = O.direct_ewald(r, q, np.full(8, 1e-9), L, alphaL=alphaL)
    assert np.allclose(p * (L / 2), q * NACL, rtol=1e-13, atol=0)
    assert np.abs(a).max() < 1e-13


def _clumpy(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(-0.5, 0.5, (n, 3))
    r[: n // 2] = (rng.normal(0, 0.05, (n // 2, 3)) + 0.5) % 1.0 - 0.5
    return r, rng.uniform(0.5, 1.5, n) / n, np.full(n, 0.004)


def test_split_independence():
    # the real-space and wave-vector sums trade places as alpha changes; a
    # wrong constant in either (erfc term, 2 alpha / sqrt(pi), the wave-vector
    # weight, the background) changes the total with alpha
    r, m, h = _clumpy(400, 1)
    idx = np.arange(0, 400, 13)
    res = [O.direct_ewald(r, m, h, 1.0, idx=idx, alphaL=aL) for aL in (2.0, 3.0, 6.0)]
    a0, p0 = res[0]
    for a, p in res[1:]:
        assert np.abs(a - a0).max() < 1e-12 * np.abs(a0).max()
        assert np.abs(p - p0).max() < 1e-12 * np.abs(p0).max()


def test_translation_invariance():
    # a common shift (wrapped into the box) moves no force
    r, m, h = _clumpy(300, 2)
    idx = np.arange(0, 300, 17)
    a0, p0 = O.direct_ewald(r, m, h, 1.0, idx=idx, alphaL=3.0)
    r1 = (r + np.array([0.37, -0.21, 0.44]) + 0.5) % 1.0 - 0.5
    a1, p1 = O.direct_ewald(r1, m, h, 1.0, idx=idx, alphaL=3.0)
    assert np.abs(a1 - a0).max() < 1e-11 * np.abs(a0).max()
    assert np.abs(p1 - p0).max() < 1e-11 * np.abs(p0).max()


def test_small_r_newton_plus_background():
    # a unit mass and a test mass at s: Newton plus the background's
    # -(4 pi / 3) m s / L^3 on the test mass (pulled towards the source
    # by Newton, pushed off by the missing mean), the images' tidal field
    # vanishing at the lattice site by cubic symmetry, up to O(s^3)
    rng = np.random.default_rng(4)
    for _ in range(5):
        s = rng.uniform(-0.005, 0.005, 3)
        r = np.array([[0.0, 0.0, 0.0], s])
        a, _ = O.direct_ewald(r, np.array([1.0, 1e-10]), np.full(2, 1e-9), 1.0,
                              idx=np.array([1]), alphaL=3.0)
        d = -s
        newton = d / np.linalg.norm(d) ** 3
        bg = -4 * np.pi / 3 * d
        assert np.linalg.norm(a[0] - newton - bg) < 1e-3 * np.linalg.norm(bg)


def test_spline_inside_twoh_across_the_boundary():
    # two particles 0.02 apart through the face x = +-1/2, twoh = 0.05: the
    # softened sum differs from the unsoftened one by the spline minus Newton
    # of the nearest image alone
    r = np.array([[0.49, 0.1, 0.0], [-0.49, 0.1, 0.0]])
    m = np.array([0.6, 0.4])
    a1, p1 = O.direct_ewald(r, m, np.array([0.02, 0.03]), 1.0, alphaL=3.0)
    a0, p0 = O.direct_ewald(r, m, np.array([1e-9, 1e-9]), 1.0, alphaL=3.0)
    d = 0.02
    sa, sb = O.spline(d * d, 0.05)
    # on particle 0 the nearest image of particle 1 is at +x
    assert a1[0, 0] - a0[0, 0] == pytest.approx(0.4 * d * (sb - 1 / d ** 3), rel=1e-11)
    assert a1[1, 0] - a0[1, 0] == pytest.approx(-0.6 * d * (sb - 1 / d ** 3), rel=1e-11)
    assert p1[0] - p0[0] == pytest.approx(-0.4 * (sa - 1 / d), rel=1e-11)
    assert abs(a1[0, 1] - a0[0, 1]) < 1e-12 and abs(a1[0, 2] - a0[0, 2]) < 1e-12


def test_reference_ewald_expands_the_direct_sum():
    # a unit mass and a test mass of 1e-10: the root's moments are the point
    # mass itself, so the reference's walk over the 27 replicas plus its
    # Ewald correction (fEwCut 2.6, fEwhCut 2.8, alpha = 2 / L) is this
    # Green's function up to the truncation of its own sums
    rng = np.random.default_rng(5)
    for _ in range(6):
        rel = rng.uniform(-0.5, 0.5, 3)
        r = np.array([[0.0, 0.0, 0.0], rel])
        m = np.array([1.0, 1e-10])
        h = np.array([1e-7, 1e-7])
        T = O.Tree(r, m, h, nBucket=16, dTheta=0.7, iOrder=4)
        a, pot, _, _ = T.grav_all(nReps=1, period=[1.0] * 3, bEwald=True, iEwOrder=4)
        ad, pd = O.direct_ewald(r, m, h, 1.0, idx=np.array([1]), alphaL=3.0)
        assert np.linalg.norm(a[1] - ad[0]) < 2e-8 * np.linalg.norm(ad[0])
        assert abs(pot[1] - pd[0]) < 1e-8
