"""Block time steps with gas through the C host (gasoline_amd/host/gg_msr.c:
ggMsrTopStepKDK, the msrTopStepKDK recursion of master.c:8242-8450 with its
SPH branches), driving the device entry points: type bits, emergency
adjust, timestep criteria, rungs, kicks, drifts with the gas's predicted
velocity and energy, gravity and msrSph on the kicked rungs.

Bars (no oracle restates the master's recursion; these are what the scheme
fixes): a collisionless system with rungs held fixed is time-reversible
(N top steps forward and back) to 1e-9; with gas and fixed rungs every
particle's force is evaluated 2^rung times per top step and only gas on the
kicked rungs gets SPH forces; one top step with a single rung equals the
Python single-rung KDK (evolve.KDK) to 1e-10; with rungs from the gravity
and Courant criteria the total energy (kinetic + potential + thermal) of an
Evrard collapse is conserved to 2e-3 over two top steps and its error falls
when the criteria's eta halves."""
import numpy as np
import pytest

from gasoline_amd import evolve, ics

pytestmark = pytest.mark.gpu

GAS = 1 | 2 | 4 | 256
DARK = 1 | 2 | 512


def evrard_pkd(gpu, n, seed, dark=0, rungs=None, soft=0.02):
    d = ics.evrard(n, seed=seed)
    pk = gpu.Pkd(n, prec=gpu.PREC_FP64)
    evolve.start(pk, d, soft, np.where(np.arange(n) < n - dark, GAS, DARK).astype(np.uint32))
    if rungs is not None:
        pk.put("iRung", rungs)
    return pk, d


def energy(pk, m, gas):
    v, pot, u = pk.get("v"), pk.get("fPot"), pk.get("u")
    return (0.5 * np.sum(m * np.sum(v * v, axis=1)) + 0.5 * np.sum(m * pot)
            + np.sum((m * u)[gas]))


def test_collisionless_fixed_rungs_reversible(gpu):
    n = 4000
    r0, v0, m = ics.plummer(n, seed=3)
    rad = np.linalg.norm(r0, axis=1)
    rungs = np.clip(np.floor(np.log2(2.0 / np.maximum(rad, 1e-3))), 0, 3).astype(np.int32)
    with gpu.Pkd(n, prec=gpu.PREC_FP64) as pk:
        pk.put("r", r0)
        pk.put("v", v0)
        pk.put("fMass", m)
        pk.put("fSoft", np.full(n, 0.02))
        pk.put("iActive", np.full(n, DARK, np.uint32))
        pk.put("iRung", rungs)
        drv = gpu.BlockSteps(pk, 0.02, iMaxRung=4, bDoGas=False, bAdjust=False)
        drv.forces(0)
        drv.s.iCurrMaxRung = 3
        n0 = drv.s.nForce
        drv.top_step()
        # one top step: 8 force evaluations at the bottom, each particle's 2^rung times
        log = drv.force_log()[n0:]
        assert len(log) == 8 and sum(c for _, c in log) == int((2 ** rungs).sum())
        for _ in range(3):
            drv.top_step()
        pk.put("v", -pk.get("v"))
        for _ in range(4):
            drv.top_step()
        back = pk.get("r")
        assert np.array_equal(pk.get("iRung"), rungs)
    assert np.abs(back - r0).max() < 1e-9


def test_gas_fixed_rungs_force_counts(gpu):
    # gas and dark matter on rungs 0..2: one top step evaluates each
    # particle's force 2^rung times (gravity), and SPH runs for the kicked
    # rungs' gas only
    n, nd = 6000, 1500
    rng = np.random.default_rng(2)
    rungs = rng.integers(0, 3, n).astype(np.int32)
    pk, d = evrard_pkd(gpu, n, 7, dark=nd, rungs=rungs)
    with pk:
        drv = gpu.BlockSteps(pk, 1e-3, iMaxRung=3, bAdjust=False)
        drv.forces(0)
        drv.s.iCurrMaxRung = 2
        n0 = drv.s.nForce
        drv.top_step()
        log = drv.force_log()[n0:]
        assert len(log) == 4
        assert sum(c for _, c in log) == int((2 ** rungs).sum())
        # the last evaluation (rungs >= 0) gave every gas particle its hydro
        # terms and no dark particle any
        bits = pk.get("iActive")
        assert np.array_equal(pk.get("iRung"), rungs)
        gas = (bits & 256) != 0
        assert (pk.get("uDotPdV")[~gas] == 0).all() and (pk.get("uDotPdV")[gas] != 0).all()


def test_single_rung_matches_the_python_driver(gpu):
    n = 8000
    out = []
    for drive in ("c", "py"):
        pk, d = evrard_pkd(gpu, n, 9)
        with pk:
            smf = gpu.Smf.default(iViscosityLimiter=1)
            if drive == "c":
                drv = gpu.BlockSteps(pk, 2e-3, iMaxRung=1, smf=smf, nSmooth=32, bAdjust=False)
                drv.forces(0)
                drv.top_step()
            else:
                run = evolve.KDK(pk, smf, nSmooth=32)
                run.forces(2e-3)
                run.step(2e-3)
            out.append({f: pk.get(f) for f in ("r", "v", "u", "a", "uDotPdV", "uDotAV")})
    for f in out[0]:
        a, b = out[0][f], out[1][f]
        scale = np.abs(b).max()
        assert np.abs(a - b).max() <= 1e-10 * scale, f


def _evrard_energy_error(gpu, eta):
    n = 6000
    pk, d = evrard_pkd(gpu, n, 11)
    m = d["m"]
    gas = np.ones(n, bool)
    with pk:
        drv = gpu.BlockSteps(pk, 0.1, iMaxRung=6, dEta=eta, dEtaCourant=eta * 2,
                             smf=gpu.Smf.default(iViscosityLimiter=1), dTheta=0.4, iOrder=4,
                             nSmooth=32)
        drv.forces(0)
        e0 = energy(pk, m, gas)
        for _ in range(2):
            drv.top_step()
        e1 = energy(pk, m, gas)
        top = drv.s.iCurrMaxRung
        rungs = pk.get("iRung")
    return abs(e1 - e0) / abs(e0), top, rungs, drv.s.nEmergency


def test_gas_adjusted_rungs_conserve_energy(gpu):
    e1, top1, r1, em1 = _evrard_energy_error(gpu, 0.2)
    e2, top2, r2, em2 = _evrard_energy_error(gpu, 0.1)
    assert top1 >= 2 and top2 >= top1, (top1, top2)
    assert e1 < 2e-3 and e2 < e1 / 2, (e1, e2)
