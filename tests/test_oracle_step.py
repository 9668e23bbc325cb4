"""Pins of the step oracle (oracle/step_oracle.c) from what the mathematics
fixes: the leapfrog of a uniform acceleration is exact, a drift by a whole
period is the identity, each timestep criterion at values where it has a
closed form, and the rung of dt = dDelta/2^k."""
import numpy as np
import pytest

from oracle import oracle as O


def fields(n, seed, gas=0.5):
    rng = np.random.default_rng(seed)
    bits = np.where(rng.uniform(size=n) < 0.7, 1, 0) | np.where(rng.uniform(size=n) < gas, 256, 0)
    return dict(iActive=bits.astype(np.uint32), r=rng.uniform(-0.5, 0.5, (n, 3)),
                v=rng.normal(size=(n, 3)), vPred=rng.normal(size=(n, 3)),
                a=rng.normal(size=(n, 3)), u=rng.uniform(1, 2, n), uPred=rng.uniform(1, 2, n),
                uDotPdV=rng.normal(size=n), uDotAV=rng.uniform(0, 1, n),
                uDotDiff=rng.normal(0, 0.1, n), fPot=rng.normal(size=n),
                dtGrav=rng.uniform(0.5, 4, n), fSoft=rng.uniform(0.01, 0.02, n),
                fBall2=rng.uniform(0.01, 0.04, n), fDensity=rng.uniform(0.5, 2, n),
                c=rng.uniform(0.5, 1.5, n), diff=rng.uniform(0, 0.1, n),
                dt=np.full(n, 0.1), dtNew=rng.uniform(0.01, 0.2, n),
                iRung=np.zeros(n, np.int32), dtOld=np.zeros(n))


def test_leapfrog_of_uniform_acceleration_is_exact():
    # kick dt/2, drift dt, kick dt/2 with a constant a: x0 + v0 t + a t^2/2, v0 + a t;
    # values chosen dyadic so every step is exact in FP64
    n = 64
    rng = np.random.default_rng(1)
    f = fields(n, 1, gas=0)
    f["iActive"][:] = 1
    f["r"] = rng.integers(-64, 64, (n, 3)) / 64.0
    f["v"] = rng.integers(-64, 64, (n, 3)) / 64.0
    f["a"] = rng.integers(-64, 64, (n, 3)) / 64.0
    dt = 0.125
    k = O.kick(f, 1.0, 0.5 * dt, 1.0, 0.5 * dt, 0.5 * dt, 0.5 * dt)
    g = dict(f, v=k["v"])
    d = O.drift(g, dt)
    k2 = O.kick(dict(g, r=d["r"]), 1.0, 0.5 * dt, 1.0, 0.5 * dt, 0.5 * dt, 0.5 * dt)
    assert np.array_equal(d["r"], f["r"] + f["v"] * dt + 0.5 * f["a"] * dt * dt)
    assert np.array_equal(k2["v"], f["v"] + f["a"] * dt)


def test_kick_touches_only_active_and_energy_only_of_gas():
    f = fields(5000, 2)
    k = O.kick(f, 1.0, 0.01, 1.0, 0.02, 0.01, 0.02)
    act = (f["iActive"] & 1) != 0
    gas = act & ((f["iActive"] & 256) != 0)
    assert np.array_equal(k["v"][~act], f["v"][~act])
    assert np.array_equal(k["u"][~gas], f["u"][~gas])
    # vPred is predicted for active gas only: a GASOLINE build leaves dark
    # and star vPred alone (pkd.c:3953-3957)
    assert np.array_equal(k["vPred"][~gas], f["vPred"][~gas])
    assert np.allclose(k["vPred"][gas], f["v"][gas] + 0.02 * f["a"][gas], rtol=1e-14)
    assert not np.array_equal(k["v"][act & ~gas], f["v"][act & ~gas])
    ud = f["uDotPdV"] + f["uDotAV"] + f["uDotDiff"]
    assert np.allclose(k["u"][gas], f["u"][gas] + 0.01 * ud[gas], rtol=1e-15, atol=0)
    assert np.allclose(k["uPred"][gas], f["u"][gas] + 0.02 * ud[gas], rtol=1e-15, atol=0)
    iso = O.kick(f, 1.0, 0.01, 1.0, 0.02, 0.01, 0.02, bIsothermal=True)
    assert np.array_equal(iso["u"], f["u"])


def test_kick_vpred_of_gas_only():
    f = fields(5000, 3)
    k = O.kick_vpred(f, 1.0, 0.05, 0.05)
    gas = (f["iActive"] & 256) != 0
    assert np.array_equal(k["vPred"][~gas], f["vPred"][~gas])
    assert np.allclose(k["vPred"][gas], f["vPred"][gas] + 0.05 * f["a"][gas], rtol=1e-14)
    assert k["bad"] == 0


def test_periodic_drift_by_a_period_is_identity():
    f = fields(4000, 4)
    L = np.array([1.0, 2.0, 0.5])
    # dyadic coordinates: x + L and x + L - L are exact
    f["r"] = np.random.default_rng(4).integers(-512, 512, (4000, 3)) / 1024.0 * L
    f["v"] = np.tile(L, (len(f["r"]), 1)) * np.random.default_rng(5).choice([-1.0, 1.0], (4000, 3))
    d = O.drift(f, 1.0, fCenter=[0, 0, 0], fPeriod=L)
    assert d["bad"] == 0
    assert np.array_equal(d["r"], f["r"])


def test_timestep_criteria_closed_forms():
    n = 16
    f = fields(n, 6)
    f["iActive"][:] = 1 | 256
    f["dt"][:] = 1e30
    f["dtGrav"][:] = 4.0
    assert np.array_equal(O.grav_step(f, 0.2)["dt"], np.full(n, 0.1))   # dEta/sqrt(4)
    f["fDensity"][:] = 0.25
    assert np.array_equal(O.density_step(f, 0.3, 4.0)["dt"], np.full(n, 0.3))
    # |a| = 4, fBall2 = 4 h^2 with h = 1: gas dt = dEta sqrt(h/|a|) = dEta/2
    f["a"] = np.tile([0.0, 0.0, 4.0], (n, 1))
    f["fBall2"][:] = 4.0
    assert np.array_equal(O.accel_step(f, 0.5, 1.0, 1.0, True, False)["dt"], np.full(n, 0.25))
    g = dict(f, iActive=np.full(n, 1, np.uint32), fSoft=np.full(n, 0.25))
    assert np.array_equal(O.accel_step(g, 0.5, 1.0, 1.0, True, False)["dt"], np.full(n, 0.125))
    # sqrt-phi criterion: dEta 3.5 sqrt(|phi|)/|a| with phi = -16
    g["fPot"] = np.full(n, -16.0)
    assert np.array_equal(O.accel_step(g, 1.0, 1.0, 1.0, False, True)["dt"], np.full(n, 3.5))
    # inactive particles keep their dt
    g["iActive"][:] = 0
    assert np.array_equal(O.accel_step(g, 1.0, 1.0, 1.0, False, True)["dt"], g["dt"])


def test_sph_step_limits():
    n = 8
    f = fields(n, 7)
    f["iActive"][:] = 1 | 256
    f["dt"][:] = 10.0
    f["dtNew"][:] = 0.5
    f["uDotPdV"][:] = -1.0            # cooling, no heating: the Courant estimate stands
    f["uDotAV"][:] = 0.0
    f["uDotDiff"][:] = 0.0
    f["diff"][:] = 0.0
    s = O.sph_step(f, 1.0, 0.4, 0.0, 0.0, 0.0, 0.1, 0.0)
    assert np.array_equal(s["dt"], np.full(n, 0.5)) and s["dtMinGas"] == 0.5
    assert np.all(s["dtNew"] == np.float32(3.402823466e38))
    # heating gas: dtFacCourant sqrt(h^2/(4 (c^2 + gamma uDot dt))), here
    # h = 1, c = 1, 5/3 uDot dt = 3: 0.4*2/1.6 * sqrt(1/16) = 0.125
    f["fBall2"][:] = 4.0
    f["c"][:] = 1.0
    f["uDotPdV"][:] = 0.18
    f["dt"][:] = 10.0
    f["dtNew"][:] = 1.0
    s = O.sph_step(f, 1.0, 0.4, 0.0, 0.0, 0.0, 0.1, 0.0)
    assert np.allclose(s["dt"], 0.125, rtol=1e-15)
    # adiabatic cooling limit dEtauDot u / |uDotPdV| with u = 2, uDotPdV = -4
    f["uDotPdV"][:] = -4.0
    f["u"][:] = 2.0
    s = O.sph_step(f, 1.0, 0.4, 0.25, 0.0, 0.0, 0.1, 0.0)
    assert np.array_equal(s["dt"], np.full(n, 0.125))
    # diffusion limit dEtaDiffusion h^2/(dDiffCoeff diff), h = 1
    f["uDotPdV"][:] = 0.0
    f["diff"][:] = 2.0
    s = O.sph_step(f, 1.0, 0.4, 0.0, 0.0, 0.5, 0.1, 0.0)
    assert np.array_equal(s["dt"], np.full(n, 0.1))
    # dtMinGas covers inactive gas too, and nothing else
    f["iActive"] = np.array([256, 257, 1, 0, 256, 256, 257, 257], np.uint32)
    f["dt"] = np.array([0.3, 10, 0.01, 0.02, 0.4, 0.5, 10, 10])
    s = O.sph_step(f, 1.0, 0.4, 0.0, 0.0, 0.5, 0.1, 0.0)
    assert s["dtMinGas"] == 0.1


def test_one_particle_rung():
    dD = 0.1
    for k in range(0, 20):
        dt = dD / 2 ** k
        assert O.one_dt_to_rung(0, dD, dt) == k, k         # the boundary goes to the lower rung
        assert O.one_dt_to_rung(0, dD, np.nextafter(dt, 0)) == k + 1, k
        assert O.one_dt_to_rung(3, dD, dt) == 3 + k
    rng = np.random.default_rng(8)
    for dt in dD * 2.0 ** -rng.uniform(0, 25, 2000):
        s = dD / dt
        if np.fmod(dD, dt) == 0:
            continue
        want = 0 if s < 1 else int(np.floor(np.log2(np.floor(s)))) + 1
        assert O.one_dt_to_rung(0, dD, dt) == want
    assert O.one_dt_to_rung(0, dD, 2 * dD) == 0


def test_dt_to_rung_bookkeeping():
    n = 1000
    f = fields(n, 9)
    f["iActive"][:] = 1
    rng = np.random.default_rng(10)
    k = rng.integers(0, 12, n)
    f["dt"] = 1.0 / 2.0 ** k
    r = O.dt_to_rung(f, 0, 1.0, 8, True)
    want = np.minimum(k, 7)
    assert np.array_equal(r["iRung"], want)
    assert np.array_equal(r["dtOld"], f["dt"])
    assert r["iMaxRung"] == 7 and r["nMaxRung"] == int((want == 7).sum())
    assert r["iMaxRungIdeal"] == k.max() + 1
    # without bAll: iRung where dDelta <= dt, else iRung + 1; lower rungs untouched
    f["iRung"] = np.where(np.arange(n) % 2 == 0, 2, 0).astype(np.int32)
    r = O.dt_to_rung(f, 1, 0.25, 8, False)
    on = f["iRung"] >= 1
    assert np.array_equal(r["iRung"][~on], f["iRung"][~on])
    assert np.array_equal(r["iRung"][on], np.where(0.25 <= f["dt"][on], 1, 2))
    # an inactive particle on an active rung is the reference's assertion
    f["iActive"][0] = 0
    assert O.dt_to_rung(f, 1, 0.25, 8, False)["iMaxRung"] == -1
