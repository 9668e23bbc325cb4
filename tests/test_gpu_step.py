"""The updates on either side of the force path through the C ABI against
the oracle's restatement of pkd.c: pkdInitAccel, pkdKick, pkdKickVpred,
pkdDrift, pkdInitDt, pkdGravStep, pkdAccelStep, pkdDensityStep, pkdSphStep
and pkdDtToRung.

Bar: bit for bit (FP64 elementwise arithmetic without contraction, as the
reference's x86-64 build rounds it), the reductions (dtMinGas, largest rung,
its count, the ideal rung) equal; the reference's assertions come back as
errors."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

N = 100_003


def fields(seed):
    rng = np.random.default_rng(seed)
    n = N
    bits = (np.where(rng.uniform(size=n) < 0.7, 1, 0) | np.where(rng.uniform(size=n) < 0.5, 256, 0)
            | 6)
    return dict(iActive=bits.astype(np.uint32), r=rng.uniform(-0.5, 0.5, (n, 3)),
                v=rng.normal(0, 0.1, (n, 3)), vPred=rng.normal(size=(n, 3)),
                a=rng.normal(size=(n, 3)), u=rng.uniform(1, 2, n), uPred=rng.uniform(1, 2, n),
                uDotPdV=rng.normal(size=n), uDotAV=rng.uniform(0, 1, n),
                uDotDiff=rng.normal(0, 0.1, n), fPot=rng.normal(size=n),
                dtGrav=rng.uniform(0.5, 4, n), fSoft=rng.uniform(0.01, 0.02, n),
                fBall2=rng.uniform(0.01, 0.04, n), fDensity=rng.uniform(0.5, 2, n),
                c=rng.uniform(0.5, 1.5, n), diff=rng.uniform(0, 0.1, n),
                dt=rng.uniform(0.001, 0.2, n), dtNew=rng.uniform(0.001, 0.2, n),
                iRung=rng.integers(0, 3, n).astype(np.int32), dtOld=np.zeros(n))


def load(pk, f):
    for k, v in f.items():
        pk.put(k, v)


def same(pk, want):
    for k, v in want.items():
        if k in pk_fields():
            got = pk.get(k)
            assert np.array_equal(got.reshape(np.shape(v)), v), k


def pk_fields():
    import gasoline_amd as gg
    return gg.FIELDS


@pytest.fixture(scope="module")
def f0():
    return fields(1)


def test_kick_and_init_accel(gpu, f0):
    with gpu.Pkd(N) as pk:
        load(pk, f0)
        pk.kick(0.99, 0.013, 1.01, 0.027, 0.011, 0.023)
        same(pk, O.kick(f0, 0.99, 0.013, 1.01, 0.027, 0.011, 0.023))
        load(pk, f0)
        pk.kick(0.99, 0.013, 1.01, 0.027, 0.011, 0.023, bIsothermal=True)
        same(pk, O.kick(f0, 0.99, 0.013, 1.01, 0.027, 0.011, 0.023, bIsothermal=True))
        pk.init_accel()
        same(pk, O.init_accel(f0))


def test_kick_vpred(gpu, f0):
    with gpu.Pkd(N) as pk:
        load(pk, f0)
        pk.kick_vpred(1.02, 0.031, 0.017)
        want = O.kick_vpred(f0, 1.02, 0.031, 0.017)
        assert want["bad"] == 0
        same(pk, want)
        g = dict(f0, uDotPdV=np.full(N, -1e6))            # uPred < 0: the reference asserts
        load(pk, g)
        with pytest.raises(gpu.GGError, match="uPred < 0"):
            pk.kick_vpred(1.0, 0.0, 1.0)


@pytest.mark.parametrize("periodic", [False, True])
def test_drift(gpu, f0, periodic):
    with gpu.Pkd(N) as pk:
        load(pk, f0)
        if periodic:
            L, c = np.array([1.0, 1.5, 2.0]), np.array([0.0, 0.1, -0.2])
            pk.drift(0.37, period=L, center=c)
            want = O.drift(f0, 0.37, fCenter=c, fPeriod=L)
        else:
            pk.drift(0.37)
            want = O.drift(f0, 0.37)
        assert want["bad"] == 0
        assert np.array_equal(pk.get("r"), want["r"])


def test_timestep_criteria(gpu, f0):
    with gpu.Pkd(N) as pk:
        load(pk, f0)
        pk.grav_step(0.17)
        w = O.grav_step(f0, 0.17)
        assert np.array_equal(pk.get("dt"), w["dt"])
        for eps, phi in ((True, False), (False, True), (True, True)):
            load(pk, f0)
            pk.accel_step(0.23, 1.0, 1.3, True, eps, phi, 0.0)
            assert np.array_equal(pk.get("dt"), O.accel_step(f0, 0.23, 1.0, 1.3, eps, phi)["dt"])
        load(pk, f0)
        pk.density_step(0.31, 2.5)
        assert np.array_equal(pk.get("dt"), O.density_step(f0, 0.31, 2.5)["dt"])
        pk.init_dt(0.05)
        assert np.array_equal(pk.get("dt"), O.init_dt(dict(f0, dt=O.density_step(
            f0, 0.31, 2.5)["dt"]), 0.05)["dt"])
        g = dict(f0, dtGrav=np.where(np.arange(N) == 5, 0.0, f0["dtGrav"]),
                 iActive=np.full(N, 7, np.uint32))
        load(pk, g)
        with pytest.raises(gpu.GGError, match="dtGrav <= 0"):
            pk.grav_step(0.1)


@pytest.mark.parametrize("limits", ["courant", "all"])
def test_sph_step(gpu, f0, limits):
    args = (1.0, 0.4, 0.0, 0.0, 0.0, 0.1, 0.0) if limits == "courant" else \
        (0.8, 0.4, 0.25, 0.3, 0.5, 0.1, 3.0)
    with gpu.Pkd(N) as pk:
        load(pk, f0)
        m = pk.sph_step(*args)
        want = O.sph_step(f0, *args)
        assert np.array_equal(pk.get("dt"), want["dt"])
        assert np.array_equal(pk.get("dtNew"), want["dtNew"])
        assert m == want["dtMinGas"]


def test_dt_to_rung(gpu, f0):
    rng = np.random.default_rng(3)
    # dt over 14 octaves, some exactly on a rung boundary
    dt = 0.1 * 2.0 ** -rng.uniform(0, 14, N)
    dt[::7] = 0.1 * 2.0 ** -rng.integers(0, 14, len(dt[::7]))
    f = dict(f0, dt=dt, iActive=(f0["iActive"] | 1).astype(np.uint32))
    with gpu.Pkd(N) as pk:
        load(pk, f)
        top, nmax, ideal = pk.dt_to_rung(1, 0.1, 10, bAll=True)
        want = O.dt_to_rung(f, 1, 0.1, 10, True)
        assert np.array_equal(pk.get("iRung"), want["iRung"])
        assert np.array_equal(pk.get("dtOld"), want["dtOld"])
        assert (top, nmax, ideal) == (want["iMaxRung"], want["nMaxRung"], want["iMaxRungIdeal"])
        assert top == 9 and ideal > 10               # capped at iMaxRung - 1
        load(pk, f)
        top, nmax, ideal = pk.dt_to_rung(1, 0.01, 10, bAll=False)
        want = O.dt_to_rung(f, 1, 0.01, 10, False)
        assert np.array_equal(pk.get("iRung"), want["iRung"])
        assert (top, nmax, ideal) == (want["iMaxRung"], want["nMaxRung"], want["iMaxRungIdeal"])
        g = dict(f, iActive=np.where(np.arange(N) == 11, 0, 1).astype(np.uint32),
                 iRung=np.full(N, 2, np.int32))
        load(pk, g)
        with pytest.raises(gpu.GGError, match="inactive"):
            pk.dt_to_rung(1, 0.1, 10)
