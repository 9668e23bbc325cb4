"""Pins of the SPH oracle (oracle/smooth_oracle.c) that do not retype its
formulas: brute-force neighbour sets, the M4 normalisation, uniform-density
and linear-velocity-field limits, and the conservation laws of the symmetric
pressure/viscosity operator."""
import numpy as np
import pytest

from gasoline_amd import ics
from oracle import oracle as O


def brute_knn(r, k, period=None):
    d = r[:, None, :] - r[None, :, :]
    if period is not None:
        L = np.asarray(period)
        d = np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))
    d2 = (d * d).sum(-1)
    idx = np.lexsort((np.broadcast_to(np.arange(len(r)), d2.shape), d2), axis=1)[:, :k]
    return idx, np.take_along_axis(d2, idx, 1)[:, -1]


@pytest.mark.parametrize("periodic", [True, False])
def test_knn_equals_brute_force(periodic):
    box = ics.sph_box(1500, seed=21)
    r = box["r"] if periodic else ics.plummer(1500, seed=22)[0]
    per = [1, 1, 1] if periodic else None
    nbr, fb2, ties = O.knn(r, 32, period=per)
    bi, bf = brute_knn(r, 32, per)
    assert ties == 0
    assert np.array_equal(nbr, bi)
    assert np.array_equal(fb2, bf)


def test_ball_gather_contains_exactly_the_knn_sets():
    box = ics.sph_box(3000, seed=23)
    nbr, fb2, _ = O.knn(box["r"], 40, period=[1, 1, 1])
    off, idx, dxyz, d2 = O.ball_gather(box["r"], fb2, period=[1, 1, 1])
    assert np.array_equal(np.diff(off), np.full(3000, 40))
    assert np.array_equal(idx.reshape(-1, 40), nbr)
    # stored separations are the minimum images: |dx| <= 1/2 and d2 = |dx|^2
    assert np.abs(dxyz).max() <= 0.5
    assert np.array_equal(d2, dxyz[:, 0] ** 2 + dxyz[:, 1] ** 2 + dxyz[:, 2] ** 2)


def test_m4_kernel_normalisation():
    # W(u)/pi integrates to 1 over the sphere of radius 2 (in units of h)
    u = np.linspace(0, 2, 200001)
    w = np.where(u < 1, 1 - 1.5 * u ** 2 + 0.75 * u ** 3, 0.25 * (2 - u) ** 3)
    integral = np.trapezoid(4 * np.pi * u ** 2 * w / np.pi, u)
    assert integral == pytest.approx(1.0, rel=1e-8)
    # the oracle's density of one particle at the centre reproduces W(0)/pi/h^3
    r = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.75]])
    fb2 = np.array([0.0625, 0.0625])           # 2h = 0.25
    lists = O.ball_gather(r, fb2)
    out = O.dendvdx(np.array([1.0, 1e-30]), np.zeros((2, 3)), np.ones(2), fb2, lists,
                    O.default_smf())
    h = 0.125
    assert out["rho"][0] == pytest.approx(1.0 / np.pi / h ** 3, rel=1e-14)


@pytest.fixture(scope="module")
def box():
    b = ics.sph_box(16384, seed=24)
    nbr, fb2, ties = O.knn(b["r"], 64, period=[1, 1, 1])
    assert ties == 0
    b["fb2"] = fb2
    b["lists"] = O.ball_gather(b["r"], fb2, period=[1, 1, 1])
    return b


def test_density_of_uniform_lattice():
    # unit mass in the unit box on a slightly jittered cubic lattice: the
    # kernel sum must give rho = 1 (the jitter only breaks distance ties)
    g = 24
    x = (np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) / g
    r = (x + np.random.default_rng(0).normal(0, 1e-3 / g, x.shape)) % 1.0
    n = len(r)
    nbr, fb2, ties = O.knn(r, 64, period=[1, 1, 1])
    assert ties == 0
    lists = O.ball_gather(r, fb2, period=[1, 1, 1])
    out = O.dendvdx(np.full(n, 1.0 / n), np.zeros((n, 3)), np.ones(n), fb2, lists,
                    O.default_smf())
    assert np.mean(out["rho"]) == pytest.approx(1.0, abs=2e-3)
    assert np.std(out["rho"]) < 2e-3
    assert (out["divv"] == 0).all()


def test_linear_velocity_field_divergence_and_curl(box):
    A = np.array([[0.3, 0.7, -0.2], [0.1, -0.5, 0.4], [0.6, 0.2, 0.8]])
    r = box["r"]
    v = (r - 0.5) @ A.T
    out = O.dendvdx(box["m"], v, box["u"], box["fb2"], box["lists"], O.default_smf())
    inner = (np.abs(r - 0.5) < 0.4).all(1)     # no wrap of the linear field
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.median(out["divv"][inner]) == pytest.approx(np.trace(A), rel=0.03)
    assert np.allclose(np.median(out["curlv"][inner], axis=0), curl, atol=0.05)


@pytest.fixture(scope="module")
def hydro(box):
    smf = O.default_smf()
    d = O.dendvdx(box["m"], box["v"], box["u"], box["fb2"], box["lists"], smf)
    pr2, c, dt = O.gas_pressure(box["u"], d["rho"], box["fb2"], smf)
    h = O.sph_pressure_sym(box["m"], box["v"], box["u"], d["rho"], pr2, c, d["balsara"],
                           d["diff"], dt, box["fb2"], box["lists"], smf)
    return d, pr2, c, dt, h


def test_gas_pressure_algebra(box, hydro):
    d, pr2, c, dt, _ = hydro
    g = 5.0 / 3.0
    assert np.allclose(pr2, (g - 1) * box["u"] / d["rho"], rtol=1e-15)
    assert np.allclose(c, np.sqrt(g * (g - 1) * box["u"]), rtol=1e-15)
    assert np.allclose(dt, 0.5 * np.sqrt(0.25 * box["fb2"]) / (2 * c), rtol=1e-15)


def test_hydro_conserves_momentum(box, hydro):
    h = hydro[4]
    p = (box["m"][:, None] * h["acc"]).sum(0)
    scale = (box["m"][:, None] * np.abs(h["acc"])).sum()
    assert np.abs(p).max() < 1e-12 * scale


def test_hydro_conserves_energy(box, hydro):
    h = hydro[4]
    m, v = box["m"], box["v"]
    ekin = (m * (v * h["acc"]).sum(1)).sum()
    eth = (m * (h["uDotPdV"] + h["uDotAV"])).sum()
    scale = (m * (np.abs(v * h["acc"]).sum(1) + np.abs(h["uDotPdV"]) + h["uDotAV"])).sum()
    assert abs(ekin + eth) < 1e-11 * scale


def test_viscosity_only_for_approaching_pairs(box):
    smf = O.default_smf()
    for sign in (+1, -1):
        v = sign * (box["r"] - 0.5) * 0.1       # pure expansion / contraction
        d = O.dendvdx(box["m"], v, box["u"], box["fb2"], box["lists"], smf)
        pr2, c, dt = O.gas_pressure(box["u"], d["rho"], box["fb2"], smf)
        h = O.sph_pressure_sym(box["m"], v, box["u"], d["rho"], pr2, c, d["balsara"],
                               d["diff"], dt, box["fb2"], box["lists"], smf)
        inner = (np.abs(box["r"] - 0.5) < 0.3).all(1)
        if sign > 0:
            assert (h["uDotAV"][inner] == 0).all() and (h["mumax"][inner] == 0).all()
            assert (h["uDotPdV"][inner] < 0).all()   # expansion cools
        else:
            assert (h["uDotAV"][inner] > 0).all() and (h["mumax"][inner] > 0).all()
            assert (h["uDotPdV"][inner] > 0).all()


def test_density_sym_weighted_sum_equals_gather(box):
    # every pair (p, q in N(p)) adds m_p*m_q*W_p*fNorm_p/2 to both ends, so the
    # mass-weighted sum of the symmetric density equals the gather density's
    m = box["m"] * np.random.default_rng(3).uniform(0.5, 1.5, len(box["m"]))
    out = O.dendvdx(m, np.zeros_like(box["r"]), box["u"], box["fb2"], box["lists"],
                    O.default_smf())
    sym = O.density_sym(m, box["fb2"], box["lists"])
    assert np.dot(m, sym) == pytest.approx(np.dot(m, out["rho"]), rel=1e-12)
    # unequal radii make it differ particle by particle
    assert np.abs(sym / out["rho"] - 1).max() > 1e-3


def test_density_sym_partial_smooth_active(box):
    # particles that are not smooth-active keep the density they held plus
    # what smooth-active neighbours scatter onto them
    n = len(box["m"])
    rng = np.random.default_rng(4)
    sact = rng.uniform(size=n) < 0.01
    m = box["m"] * rng.uniform(0.5, 1.5, n)
    rho0 = rng.uniform(1, 2, n)
    out = O.dendvdx(m, np.zeros_like(box["r"]), box["u"], box["fb2"], box["lists"],
                    O.default_smf(), smooth_active=sact.astype(np.uint8))
    sym = O.density_sym(m, box["fb2"], box["lists"], smooth_active=sact.astype(np.uint8),
                        rho=rho0)
    base = np.where(sact, 0.0, rho0)
    assert np.dot(m, sym - base) == pytest.approx(np.dot(m[sact], out["rho"][sact]), rel=1e-12)
    off, idx = box["lists"][0], box["lists"][1]
    touched = np.zeros(n, bool)
    for p in np.flatnonzero(sact):
        touched[idx[off[p]:off[p + 1]]] = True
    untouched = ~sact & ~touched
    assert untouched.any() and np.array_equal(sym[untouched], rho0[untouched])
    # (the farthest neighbour sits at u = 2, where W = 0)
    assert (sym[~sact & touched] >= rho0[~sact & touched]).all()
    assert np.mean(sym[~sact & touched] > rho0[~sact & touched]) > 0.9


def test_density_sym_equals_gather_on_symmetric_lists():
    # perfect periodic cubic lattice, 27 neighbours (shells 0, 1, sqrt2, sqrt3
    # lattice spacings; the next shell is at 2): every list is symmetric and
    # every radius equal, so the symmetric and gather forms agree
    g = 8
    r = (np.stack(np.meshgrid(*[np.arange(g)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) / g
    n = len(r)
    nbr, fb2, ties = O.knn(r, 27, period=[1, 1, 1])
    assert ties == 0 and np.all(fb2 == fb2[0])
    lists = O.ball_gather(r, fb2, period=[1, 1, 1])
    m = np.full(n, 1.0 / n)
    out = O.dendvdx(m, np.zeros((n, 3)), np.ones(n), fb2, lists, O.default_smf())
    sym = O.density_sym(m, fb2, lists)
    assert np.allclose(sym, out["rho"], rtol=1e-14, atol=0)


def test_threaded_hydro_equals_serial_scatter(box):
    # the threaded gather form sums the serial scatter form's pair terms in
    # another order: the same to rounding, mumax and dtNew (max and min)
    # bit for bit; with a third of the particles active
    smf = O.default_smf(iViscosityLimiter=1, dThermalDiffusionCoeff=0.03)
    act = (np.random.default_rng(9).uniform(size=len(box["m"])) < 0.33).astype(np.uint8)
    v = box["v"] + 0.3 * (box["r"] - 0.5)
    u = box["u"] * (1 + 0.5 * np.sin(2 * np.pi * box["r"][:, 0]))
    d = O.dendvdx(box["m"], v, u, box["fb2"], box["lists"], smf)
    pr2, c, dt = O.gas_pressure(u, d["rho"], box["fb2"], smf)
    args = (box["m"], v, u, d["rho"], pr2, c, d["balsara"], d["diff"], dt, box["fb2"],
            box["lists"], smf)
    a = O.sph_pressure_sym(*args, active=act)
    b = O.sph_pressure_sym(*args, active=act, serial=True)
    for k in ("mumax", "dtNew"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("acc", "uDotPdV", "uDotAV", "uDotDiff"):
        scale = np.abs(b[k]).max()
        assert scale > 0 and np.abs(a[k] - b[k]).max() < 1e-13 * scale, k
    assert (a["acc"][act == 0] == 0).all()
