"""Pins of the oracle's SPH operators and gravity bookkeeping by what the
mathematics fixes, not by retyping the oracle's formulas.

SPH: a perfect cubic lattice (spacing 1/16, exact in binary) with one
smoothing radius for every particle.  Deep inside it (two smoothing radii
from the edge, or four where a neighbour's own sums enter) every
neighbourhood is complete and has cubic symmetry, so

* the SPH estimate of a linear field's gradient is exact: sum_q dx_i dx_j f(r)
  = delta_ij/3 sum_q r^2 f(r) over every full shell, and DenDVDX's
  normalisation 3/|sum r^2 dW| turns it into the gradient itself
  (smoothfcn.c:3466-3480, 3518-3540);
* kernel sums approach their continuum integrals, which fix the magnitude of
  the pressure force (-grad P / rho), of the PdV work (-(P/rho) div v) and of
  the artificial viscosity's heating (the Monaghan pair term of
  SphPressureTerms.h:102-110 integrated over the M4 kernel).

A wrong factor in the kernel-gradient normalisation, a dropped half of the
symmetric pair sum or a wrong viscosity coefficient moves one of these by far
more than the lattice's discretisation error (0.15 % at 2h = 4.4 spacings).

Gravity: dtGrav of an isolated pair and of a particle facing a distant
compact cell (grav.c:100-101, 148-149, 189-190) and the softened/Newtonian
classification of an accepted cell on both sides of each of its two
boundaries (walk.c:122-171)."""
import numpy as np
import pytest

from oracle import oracle as O

SP = 1.0 / 16.0            # lattice spacing
G = 28                     # particles per side
TWOH = 4.4 * SP            # 2h, the same for every particle
GAMMA = 5.0 / 3.0


def m4_w(q):
    """M4 kernel in units of 1/(pi h^3) (Monaghan & Lattanzio 1985)."""
    return np.where(q < 1, 1 - 1.5 * q ** 2 + 0.75 * q ** 3,
                    np.where(q < 2, 0.25 * (2 - q) ** 3, 0.0))


def m4_dwdq(q):
    return np.where(q < 1, -3 * q + 2.25 * q ** 2, np.where(q < 2, -0.75 * (2 - q) ** 2, 0.0))


@pytest.fixture(scope="module")
def lat():
    ijk = np.stack(np.meshgrid(*[np.arange(G)] * 3, indexing="ij"), -1).reshape(-1, 3)
    r = ijk * SP
    n = len(r)
    fb2 = np.full(n, TWOH * TWOH)
    lists = O.ball_gather(r, fb2)
    m = np.full(n, SP ** 3)              # rho = 1 in the continuum

    def inner(nh):                       # at least nh smoothing lengths (h) from the edge
        k = int(np.ceil(nh * TWOH / 2 / SP))
        return ((ijk >= k) & (ijk <= G - 1 - k)).all(1)
    return dict(r=r - r.mean(0), n=n, fb2=fb2, lists=lists, m=m, inner=inner)


def hydro(lat, v, u, **kw):
    smf = O.default_smf(**kw)
    d = O.dendvdx(lat["m"], v, u, lat["fb2"], lat["lists"], smf)
    pr2, c, dt = O.gas_pressure(u, d["rho"], lat["fb2"], smf)
    h = O.sph_pressure_sym(lat["m"], v, u, d["rho"], pr2, c, d["balsara"], d["diff"], dt,
                           lat["fb2"], lat["lists"], smf)
    return d, pr2, c, h


def test_lattice_density_is_the_continuum_density(lat):
    d = O.dendvdx(lat["m"], np.zeros((lat["n"], 3)), np.ones(lat["n"]), lat["fb2"],
                  lat["lists"], O.default_smf())
    rho = d["rho"][lat["inner"](2)]
    assert np.ptp(rho) == 0.0 or np.ptp(rho) < 1e-14
    assert rho[0] == pytest.approx(1.0, rel=2e-3)


def test_pressure_force_is_minus_grad_p_over_rho(lat):
    # u linear, v = 0: P = (gamma-1) rho u with rho uniform inside, so the
    # continuum force per unit mass is -(gamma-1) grad u; the symmetric sum
    # takes half from each end's list (fNorm's 0.5, smoothfcn.c:2303-2306)
    gu = np.array([0.3, -0.2, 0.5])
    u = 10.0 + lat["r"] @ gu
    _, _, _, h = hydro(lat, np.zeros((lat["n"], 3)), u)
    a = h["acc"][lat["inner"](4)]
    want = -(GAMMA - 1) * gu
    assert np.abs(a / want - 1).max() < 5e-3
    # no viscosity without relative motion
    assert (h["uDotAV"] == 0).all()


def test_pdv_work_is_minus_p_over_rho_div_v(lat):
    # uniform compression v = -eps x: du/dt = -(P/rho) div v = 3 eps (gamma-1) u
    eps, u0 = 0.1, 2.0
    v = -eps * lat["r"]
    _, _, _, h = hydro(lat, v, np.full(lat["n"], u0))
    pdv = h["uDotPdV"][lat["inner"](4)]
    assert np.abs(pdv / (3 * eps * (GAMMA - 1) * u0) - 1).max() < 5e-3
    # the compression is isotropic: no net force inside
    assert np.abs(h["acc"][lat["inner"](4)]).max() < 1e-9 * 3 * eps * u0 / TWOH


@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.0, 2.0), (1.0, 2.0)])
def test_viscous_heating_is_the_continuum_monaghan_integral(lat, alpha, beta):
    # v = -eps x: every pair approaches with dv.dx = -eps r^2, mu = eps r;
    # the pair viscosity (alpha (c_p+c_q) + 1.5 beta mu) mu / (rho_p+rho_q)
    # heats each end by half of visc * (v_pq . grad W).  In the continuum:
    # du/dt = rho/2 * int visc(r) eps r |dW/dr| 4 pi r^2 dr
    eps, u0 = 0.3, 2.0
    v = -eps * lat["r"]
    _, _, c, h = hydro(lat, v, np.full(lat["n"], u0), alpha=alpha, beta=beta)
    ins = lat["inner"](4)
    c0 = np.sqrt(GAMMA * (GAMMA - 1) * u0)
    assert np.allclose(c[ins], c0, rtol=1e-15)
    hh = TWOH / 2
    rr = np.linspace(0, TWOH, 400001)
    dwdr = m4_dwdq(rr / hh) / (np.pi * hh ** 4)
    visc = (alpha * 2 * c0 + 1.5 * beta * eps * rr) * eps * rr / 2.0    # rho = 1
    want = 0.5 * np.trapezoid(visc * eps * rr * (-dwdr) * 4 * np.pi * rr ** 2, rr)
    got = h["uDotAV"][ins]
    assert np.abs(got / want - 1).max() < 5e-3
    # mu = eps r peaks at the farthest neighbour (sqrt(19) spacings: the
    # largest sum of three squares within (2h/spacing)^2 = 19.36)
    assert np.abs(h["mumax"][ins] / (eps * np.sqrt(19) * SP) - 1).max() < 1e-12


def velocity_gradient_case(lat, A, gu, H):
    u = 5.0 + lat["r"] @ gu
    v = lat["r"] @ A.T
    smf = O.default_smf(H=H, iViscosityLimiter=1)
    d = O.dendvdx(lat["m"], v, u, lat["fb2"], lat["lists"], smf)
    return {k: (d[k][lat["inner"](2)]) for k in ("divv", "fDivv_t", "curlv", "dvds", "diff",
                                                 "balsara")}


@pytest.mark.parametrize("case", ["expanding", "compressing"])
def test_dendvdx_outputs_of_a_linear_field(lat, case):
    # the estimated gradient tensor equals A: divv = tr A + 3H, fDivv_t = tr A,
    # curl v, the derivative along the u gradient g (dvds, with H on the
    # diagonal; 1.5 (dvdr - divv/3) under compression), the shear diffusion
    # coefficient h^2 |S| with S the traceless symmetric part of A
    rng = np.random.default_rng(5 if case == "expanding" else 6)
    A = rng.normal(0, 1, (3, 3))
    A += np.eye(3) * (2.0 if case == "expanding" else -2.0)
    gu = rng.normal(0, 1, 3)
    H = 0.07
    o = velocity_gradient_case(lat, A, gu, H)
    tr = np.trace(A)
    divv = tr + 3 * H
    assert np.allclose(o["divv"], divv, rtol=1e-11, atol=1e-12)
    assert np.allclose(o["fDivv_t"], tr, rtol=1e-11, atol=1e-12)
    curl = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    assert np.allclose(o["curlv"], curl, rtol=1e-11, atol=1e-12)
    gh = gu / np.linalg.norm(gu)
    dvdr = gh @ (A + H * np.eye(3)) @ gh
    dvds = 1.5 * (dvdr - divv / 3) if divv < 0 else dvdr
    assert np.allclose(o["dvds"], dvds, rtol=1e-10, atol=1e-12)
    S = 0.5 * (A + A.T) - tr / 3 * np.eye(3)
    diff = 0.25 * TWOH ** 2 * np.sqrt(2 * (S * S).sum())
    assert np.allclose(o["diff"], diff, rtol=1e-10)
    bal = abs(divv) / (abs(divv) + np.linalg.norm(curl))
    assert np.allclose(o["balsara"], bal, rtol=1e-10)


def test_balsara_switch_limits(lat):
    # pure rotation: no compression, the switch turns viscosity off; pure
    # compression: no curl, fully on (iViscosityLimiter 1, smoothfcn.c:3542-3569)
    W = np.array([[0, -0.4, 0.3], [0.4, 0, -0.2], [-0.3, 0.2, 0]])
    rot = velocity_gradient_case(lat, W, np.array([1.0, 0.0, 0.0]), 0.0)
    assert np.abs(rot["divv"]).max() < 1e-12
    assert rot["balsara"].max() < 1e-11
    comp = velocity_gradient_case(lat, -0.5 * np.eye(3), np.array([1.0, 0.0, 0.0]), 0.0)
    assert np.allclose(comp["balsara"], 1.0, rtol=0, atol=1e-12)
    # limiter 2 (dvds-based): planar compression along the u gradient is a
    # shock (switch 1), with rotation added |dvds| / (|dvds| + |curl|) where
    # dvds = 1.5 (dvdr - divv/3) = -0.5; no motion at all gives dvds = 0 and
    # the floor ALPHAMIN (pkd.h:41)
    r, u = lat["r"], 5.0 + lat["r"][:, 0]
    cw = np.linalg.norm([W[2, 1] - W[1, 2], W[0, 2] - W[2, 0], W[1, 0] - W[0, 1]])
    P = np.diag([-0.5, 0.0, 0.0])
    for A, want in ((P, 1.0), (P + W, 0.5 / (0.5 + cw)), (np.zeros((3, 3)), 0.01)):
        d = O.dendvdx(lat["m"], r @ A.T, u, lat["fb2"], lat["lists"],
                      O.default_smf(iViscosityLimiter=2))
        b = d["balsara"][lat["inner"](2)]
        assert np.allclose(b, want, rtol=0, atol=1e-9), (A, b[:3])


def two_gas_particles(vrel, h=(0.1, 0.1), u=(1.0, 1.0)):
    """Two gas particles 0.1 apart on the x axis, particle 1 moving at vrel
    along +x (vrel < 0: approaching).  Each list holds both particles (the
    owner itself included, as smSmooth's does)."""
    r = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]])
    v = np.array([[0.0, 0.0, 0.0], [vrel, 0.0, 0.0]])
    fb2 = (2 * np.asarray(h)) ** 2
    lists = O.ball_gather(r, fb2)
    assert np.array_equal(np.diff(lists[0]), [2, 2])
    m = np.array([1.0, 1.0])
    smf = O.default_smf()
    uu = np.asarray(u, float)
    d = O.dendvdx(m, v, uu, fb2, lists, smf)
    pr2, c, dt = O.gas_pressure(uu, d["rho"], fb2, smf)
    out = O.sph_pressure_sym(m, v, uu, d["rho"], pr2, c, d["balsara"], d["diff"], dt, fb2,
                             lists, smf)
    return out, c, dt, smf


def test_pair_mumax_is_the_closing_speed():
    out, _, _, _ = two_gas_particles(-0.7)
    assert np.allclose(out["mumax"], 0.7, rtol=1e-14)
    out, _, _, _ = two_gas_particles(+0.7)
    assert (out["mumax"] == 0).all()


def test_pair_dtnew_courant_forms():
    # receding: the signal speed is 2 max(c) (SphPressureTerms.h:122); each
    # list is evaluated with its owner's h, so dtNew takes the smaller h;
    # and never more than 4 dt of the other particle (SETDTNEW_PQ)
    out, c, dt, smf = two_gas_particles(+0.3, h=(0.1, 0.08), u=(1.0, 1.5))
    k = smf.dtFacCourant
    want = k * 0.08 / (2 * c.max())
    assert np.allclose(out["dtNew"], min(want, 4 * dt.min()), rtol=1e-14)
    # approaching at closing speed w: signal speed 1.25 cbar + 0.375 (alpha
    # (c_p+c_q) + 1.5 beta w) with mu = w, the smaller h again
    w = 0.5
    out, c, dt, smf = two_gas_particles(-w, h=(0.1, 0.08), u=(1.0, 1.5))
    sig = 0.625 * c.sum() + 0.375 * (smf.alpha * c.sum() + 1.5 * smf.beta * w)
    self_term = smf.dtFacCourant * np.array([0.1, 0.08]) / (2 * c)   # the pair of a particle with itself
    want = min(smf.dtFacCourant * 0.08 / sig, self_term.min(), 4 * dt.min())
    assert np.allclose(out["dtNew"], want, rtol=1e-14)


def test_dtgrav_of_an_isolated_pair():
    # beyond the softening the pair term is (m_i + m_j) / r^3 (grav.c:211-242)
    r = np.array([[0.0, 0.0, 0.0], [0.3, 0.4, 0.0]])
    m = np.array([2.0, 3.0])
    T = O.Tree(r, m, np.array([0.01, 0.02]), nBucket=16)
    _, _, dtg, _ = T.grav_all()
    assert np.allclose(dtg, 5.0 / 0.5 ** 3, rtol=1e-14)


def cluster_and_probe(soft_probe, D, soft_cl=0.002, seed=3):
    """A compact cluster of 200 particles (radius 1e-3) at the origin and one
    probe particle at distance D on the x axis."""
    rng = np.random.default_rng(seed)
    cl = rng.normal(0, 3e-4, (200, 3))
    r = np.vstack([cl, [[D, 0.0, 0.0]]])
    m = np.concatenate([np.full(200, 1.0 / 200), [1e-6]])
    soft = np.concatenate([np.full(200, soft_cl), [soft_probe]])
    T = O.Tree(r, m, soft, nBucket=16, dTheta=0.7, iOrder=2)
    return T, r, m, soft


def probe_lists(T):
    """The probe's bucket, its walk lists and the root's other child."""
    perm = T.perm()
    slot = int(np.flatnonzero(perm == 200)[0])
    ints, dbls = T.nodes()
    b = [i for i in T.buckets() if ints[i, 1] <= slot <= ints[i, 2]][0]
    return b, T.walk(b), ints, dbls


def test_dtgrav_of_a_distant_newtonian_cell():
    D = 1.0
    T, r, m, soft = cluster_and_probe(1e-3, D)
    b, (pp, cs, cn), ints, dbls = probe_lists(T)
    assert len(cn) == 1 and len(cs) == 0 and len(pp) == 0
    com = dbls[cn[0], 9:12]
    R = np.linalg.norm(r[200] - com)
    _, _, dtg, _ = T.grav_all()
    # 1/r of a listed cell comes from v_sqrt1 (accurate to ~1e-9, v_sqrt1.c)
    assert dtg[200] == pytest.approx((1e-6 + 1.0) / R ** 3, rel=3e-9)


def test_soft_or_newtonian_cell_on_each_side_of_its_boundaries():
    # an accepted cell is softened when the softening sphere twoh = h_cell +
    # fSoftMax reaches past its opening radius (twoh^2 >= fOpen2) AND the
    # bucket lies within twoh of the cell's centre of mass; otherwise
    # Newtonian.  The probe (a bucket of one particle) sits at distance D.
    D = 0.05
    T, r, m, soft = cluster_and_probe(1e-3, D)
    b, (pp, cs, cn), ints, dbls = probe_lists(T)
    cell = (list(cs) + list(cn))[0]
    hc, fo2 = dbls[cell, 8], dbls[cell, 12]
    R = np.linalg.norm(r[200] - dbls[cell, 9:12])
    assert fo2 < R * R                               # the cell is accepted
    cases = {
        "inside twoh, twoh^2 >= fOpen2": (R * (1 + 1e-9) - hc, "soft"),
        "just outside twoh": (R * (1 - 1e-9) - hc, "newt"),
        "twoh^2 just below fOpen2": (np.sqrt(fo2) * (1 - 1e-9) - hc, "newt"),
        "twoh^2 just above fOpen2 but outside": (np.sqrt(fo2) * (1 + 1e-9) - hc, "newt"),
        "deep inside twoh": (3 * R - hc, "soft"),
    }
    for name, (sp, kind) in cases.items():
        T2, r2, m2, s2 = cluster_and_probe(sp, D)
        b2, (pp2, cs2, cn2), ints2, dbls2 = probe_lists(T2)
        assert dbls2[cell, 8] == hc and dbls2[cell, 12] == fo2   # same cell, same radii
        got = "soft" if cell in cs2 else ("newt" if cell in cn2 else "none")
        assert got == kind, name
        # either way the force on the probe is the softened direct sum's to
        # the cluster's multipole error ((1e-3 / D)^3)
        a, _, _, _ = T2.grav_all()
        ad, _ = O.direct_sum(r2, m2, s2, idx=[200])
        assert np.linalg.norm(a[200] - ad[0]) / np.linalg.norm(ad[0]) < 1e-5, name
        if name == "deep inside twoh":
            # where the softening matters, a Newtonian cell would be far off
            assert np.linalg.norm(a[200]) < 0.5 * 1.0 / R ** 2
